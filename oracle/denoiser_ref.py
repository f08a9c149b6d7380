"""CPU oracle of the histogram denoiser (include/pach_denoise.h) -- TEST
INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU leg use it
to check the HIP network; the product never calls it.

The network has no counterpart in the reference (Pachyderm converges by ray
count alone); this is the plain FP64 statement of the same network the HIP
kernels compute: forward pass, mean-squared loss, and the backward pass by the
chain rule.  `bf16=True` rounds every GEMM operand to BF16 (round to nearest
even) where the HIP path does, so the GPU can be compared with it at FP32
accumulation accuracy.

Pins (tests/test_denoiser_ref.py): the backward pass against central finite
differences in FP64; the im2col product against a direct kernel-3
convolution loop; GELU and its derivative against finite differences.
"""
from __future__ import annotations

import numpy as np

GELU_K = 0.7978845608028654  # sqrt(2 / pi)


def bf16_round(a):
    """Round float values to the nearest BF16 (ties to even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_K * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_K * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K * (1.0 + 3 * 0.044715 * x * x)


class Layout:
    """Parameter-vector layout of pach_dn (offsets in elements)."""

    def __init__(self, bins, patch, width, blocks, n_desc):
        self.bins, self.patch, self.D, self.NB, self.n_desc = bins, patch, width, blocks, n_desc
        self.P = bins // patch
        self.F = 8 * patch
        self.KE = (self.F + n_desc + 31) // 32 * 32
        D, P, KE, F = self.D, self.P, self.KE, self.F
        shapes = [("We", (KE, D)), ("be", (D,)), ("Pos", (P, D))]
        for i in range(blocks):
            shapes += [(f"Wc{i}", (3 * D, D)), (f"bc{i}", (D,)), (f"W1{i}", (D, 2 * D)), (f"b1{i}", (2 * D,)),
                       (f"W2{i}", (2 * D, D)), (f"b2{i}", (D,))]
        shapes += [("Wh", (D, F)), ("bh", (F,))]
        self.shapes = dict(shapes)
        self.offsets = {}
        o = 0
        for name, shp in shapes:
            self.offsets[name] = o
            n = int(np.prod(shp))
            o += (n + 7) // 8 * 8 if name == "Pos" else n
        self.n = o

    def unpack(self, vec):
        return {k: vec[o:o + int(np.prod(self.shapes[k]))].reshape(self.shapes[k]) for k, o in self.offsets.items()}

    def pack(self, tensors):
        v = np.zeros(self.n)
        for k, o in self.offsets.items():
            a = np.asarray(tensors[k], dtype=np.float64).ravel()
            v[o:o + a.size] = a
        return v


def patchify(L, x, desc):
    """[B, 8, bins], [B, n_desc] -> feature rows [B * P, KE]."""
    B = x.shape[0]
    xp = x.reshape(B, 8, L.P, L.patch).transpose(0, 2, 1, 3).reshape(B, L.P, L.F)
    X = np.zeros((B, L.P, L.KE))
    X[:, :, :L.F] = xp
    if L.n_desc:
        X[:, :, L.F:L.F + L.n_desc] = desc[:, None, :]
    return X.reshape(B * L.P, L.KE)


def unpatchify(L, Y, B):
    return Y.reshape(B, L.P, 8, L.patch).transpose(0, 2, 1, 3).reshape(B, 8, L.bins)


def im2col(L, H, B):
    """C[m, tap * D + c] = H[m + tap - 1, c] within a sample, else 0."""
    Hs = H.reshape(B, L.P, L.D)
    z = np.zeros((B, 1, L.D))
    Hp = np.concatenate([z, Hs, z], axis=1)
    C = np.concatenate([Hp[:, 0:L.P], Hp[:, 1:L.P + 1], Hp[:, 2:L.P + 2]], axis=2)
    return C.reshape(B * L.P, 3 * L.D)


def col2im(L, dC, B):
    d = dC.reshape(B, L.P, 3, L.D)
    out = d[:, :, 1].copy()
    out[:, :-1] += d[:, 1:, 0]
    out[:, 1:] += d[:, :-1, 2]
    return out.reshape(B * L.P, L.D)


def forward(L, params, x, desc, bf16=False):
    """Returns (out [B, 8, bins], cache)."""
    r = bf16_round if bf16 else (lambda a: a)
    W = L.unpack(params)
    B = x.shape[0]
    X0 = r(patchify(L, x, desc))
    H = X0 @ r(W["We"]) + W["be"] + np.tile(W["Pos"], (B, 1))
    Hb = [r(H)]
    cache = {"X0": X0, "Hb": Hb, "Upre": [], "Ub": [], "Vpre": [], "Vb": []}
    for i in range(L.NB):
        C1 = im2col(L, Hb[i], B)
        Upre = C1 @ r(W[f"Wc{i}"]) + W[f"bc{i}"]
        Ub = r(gelu(Upre))
        Vpre = Ub @ r(W[f"W1{i}"]) + W[f"b1{i}"]
        Vb = r(gelu(Vpre))
        H = H + Vb @ r(W[f"W2{i}"]) + W[f"b2{i}"]
        Hb.append(r(H))
        for k, v in (("Upre", Upre), ("Ub", Ub), ("Vpre", Vpre), ("Vb", Vb)):
            cache[k].append(v)
    Y = Hb[L.NB] @ r(W["Wh"]) + W["bh"]
    out = x + unpatchify(L, Y, B)
    cache["B"] = B
    return out, cache


def loss_grad(L, params, x, desc, target, bf16=False, w=None):
    """Mean of w[bin] (out - target)^2 and its gradient (parameter-vector layout)."""
    r = bf16_round if bf16 else (lambda a: a)
    W = L.unpack(params)
    out, c = forward(L, params, x, desc, bf16)
    B = c["B"]
    diff = out - target
    w = np.ones(L.bins) if w is None else np.asarray(w, dtype=np.float64)
    loss = float(np.mean(w * diff ** 2))
    dout = 2 * w * diff / diff.size
    dY = dout.reshape(B, 8, L.P, L.patch).transpose(0, 2, 1, 3).reshape(B * L.P, L.F)
    dYb = r(dY)
    G = {}
    G["bh"] = dY.sum(0)
    G["Wh"] = c["Hb"][L.NB].T @ dYb
    dH = dYb @ r(W["Wh"]).T
    dHb = r(dH)
    for i in reversed(range(L.NB)):
        G[f"b2{i}"] = dH.sum(0)
        G[f"W2{i}"] = c["Vb"][i].T @ dHb
        dV = (dHb @ r(W[f"W2{i}"]).T) * gelu_grad(c["Vpre"][i])
        dVb = r(dV)
        G[f"b1{i}"] = dV.sum(0)
        G[f"W1{i}"] = c["Ub"][i].T @ dVb
        dU = (dVb @ r(W[f"W1{i}"]).T) * gelu_grad(c["Upre"][i])
        dUb = r(dU)
        G[f"bc{i}"] = dU.sum(0)
        C1 = im2col(L, c["Hb"][i], B)
        G[f"Wc{i}"] = C1.T @ dUb
        dC1 = dUb @ r(W[f"Wc{i}"]).T
        dH = dH + col2im(L, dC1, B)
        dHb = r(dH)
    G["be"] = dH.sum(0)
    G["Pos"] = dH.reshape(B, L.P, L.D).sum(0)
    G["We"] = c["X0"].T @ dHb
    return loss, L.pack(G)


def conv3_direct(L, H, Wc, B):
    """Kernel-3 convolution over patches by explicit loops (pins im2col)."""
    Hs = H.reshape(B, L.P, L.D)
    K = Wc.reshape(3, L.D, -1)
    out = np.zeros((B, L.P, K.shape[2]))
    for b in range(B):
        for p in range(L.P):
            for tap in range(3):
                q = p + tap - 1
                if 0 <= q < L.P:
                    out[b, p] += Hs[b, q] @ K[tap]
    return out.reshape(B * L.P, -1)
