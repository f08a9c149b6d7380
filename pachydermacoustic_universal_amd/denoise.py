"""Learned denoiser of receiver energy histograms (include/pach_denoise.h).

No counterpart in the reference: Pachyderm converges a trace by ray count
alone.  This module holds the host side of the network libpach.so trains and
runs on the GPU (BF16 MFMA GEMMs, FP32 accumulation, AdamW on FP32 master
weights): parameter initialisation, the histogram normalisation, room
descriptors, low/high ray-count training pairs traced by the tracer itself,
the training loop, and inference that returns denoised energy histograms.

Representation: each band of a histogram E[8, bins] (energy per bin, already
divided by the ray count as Receiver_Bank.Scale does) enters as its
normalised log backward-integrated (Schroeder) curve,
x(t) = (log10(max(Sch(t) / S, 1e-10)) + 5) / 2.5 with Sch(t) = sum over
tau >= t of E(tau) and S = Sch(0) of the low-ray input: T30 and EDT are line
fits to exactly this curve (in dB) and C80 is a ratio of two of its values,
so a mean-squared error on it weighs the bins by what the room-acoustic
parameters see.  The network's output y maps back to a curve
S * 10^(2.5 y - 5) (zero within 0.05 of the floor), made non-increasing, and
its backward differences are the denoised histogram.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import _abi as A
from . import rooms as RM

FLOOR = 1e-10
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "denoiser.npz")
DEFAULTS = dict(bins=3000, patch=24, width=256, blocks=4, n_desc=12, max_batch=256)


def layout(bins, patch, width, blocks, n_desc, **_):
    """Offsets and shapes of the parameter vector (order of include/pach_denoise.h)."""
    P, F = bins // patch, 8 * patch
    KE = (F + n_desc + 31) // 32 * 32
    D = width
    shapes = [("We", (KE, D)), ("be", (D,)), ("Pos", (P, D))]
    for i in range(blocks):
        shapes += [(f"Wc{i}", (3 * D, D)), (f"bc{i}", (D,)), (f"W1{i}", (D, 2 * D)), (f"b1{i}", (2 * D,)),
                   (f"W2{i}", (2 * D, D)), (f"b2{i}", (D,))]
    shapes += [("Wh", (D, F)), ("bh", (F,))]
    out, o = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (o, shp)
        o += (n + 7) // 8 * 8 if name == "Pos" else n
    return out, o


def init_params(cfg, seed=0):
    """He-normal convolution and MLP weights, a small last layer per residual
    branch and a zero head, so the untrained network returns its input."""
    rng = np.random.default_rng(seed)
    lay, n = layout(**cfg)
    v = np.zeros(n, dtype=np.float32)
    D = cfg["width"]
    for name, (o, shp) in lay.items():
        k = int(np.prod(shp))
        if name == "We":
            w = rng.normal(0, 1 / np.sqrt(shp[0]), shp)
        elif name == "Pos":
            w = rng.normal(0, 0.02, shp)
        elif name.startswith("Wc"):
            w = rng.normal(0, np.sqrt(2 / (3 * D)), shp)
        elif name.startswith("W1"):
            w = rng.normal(0, np.sqrt(2 / D), shp)
        elif name.startswith("W2"):
            w = rng.normal(0, 0.1 / np.sqrt(2 * D), shp)
        else:
            continue  # biases and the head start at zero
        v[o:o + k] = w.ravel()
    return v


def normalize(E):
    """E [..., 8, bins] -> (x float32 [..., 8, bins], S [..., 8, 1])."""
    E = np.asarray(E, dtype=np.float64)
    sch = np.cumsum(E[..., ::-1], axis=-1)[..., ::-1]
    S = np.maximum(sch[..., :1], 1e-300)
    x = (np.log10(np.maximum(sch / S, FLOOR)) + 5.0) / 2.5
    return x.astype(np.float32), S


def denormalize(y, S):
    """Network output (normalised log Schroeder curve) -> energy histogram."""
    y = np.asarray(y, dtype=np.float64)
    sch = S * 10.0 ** (2.5 * y - 5.0)
    sch[y <= (np.log10(FLOOR) + 5.0) / 2.5 + 0.05] = 0.0
    sch = np.minimum.accumulate(sch, axis=-1)
    E = sch.copy()
    E[..., :-1] -= sch[..., 1:]
    return E


def ray_descriptor(rays):
    """The trace's source-ray count as a network input: log10(rays) - 4.5."""
    return np.float32(np.log10(float(rays)) - 4.5)


def descriptors(room: RM.Room, receivers=None, rays=None):
    """Per-receiver descriptors [R, 12]: log10 volume / 4, log10 surface area
    / 4, area-weighted mean absorption per band (8), log10 source-receiver
    distance / 2, area-weighted mean scattering; with `rays`, a 13th column
    holds the trace's ray count (ray_descriptor)."""
    v = room.verts
    area_t = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    area_s = np.bincount(room.surf, weights=area_t, minlength=room.n_surfaces)
    total = area_s.sum()
    vol = room.volume if room.volume > 0 else RM._tri_volume(room.verts)
    abs_mean = (area_s[:, None] * room.alpha).sum(0) / total
    scat_mean = float((area_s[:, None] * room.scat).sum() / (8 * total))
    recs = np.asarray(room.receivers if receivers is None else receivers, dtype=np.float64).reshape(-1, 3)
    dist = np.maximum(np.linalg.norm(recs - room.source, axis=1), 1e-3)
    head = np.concatenate([[np.log10(vol) / 4, np.log10(total) / 4], abs_mean])
    out = np.empty((recs.shape[0], 12))
    out[:, :10] = head
    out[:, 10] = np.log10(dist) / 2
    out[:, 11] = scat_mean
    if rays is not None:
        out = np.concatenate([out, np.full((out.shape[0], 1), ray_descriptor(rays))], axis=1)
    return out.astype(np.float32)


class Denoiser:
    """One pach_dn network on one GPU."""

    def __init__(self, device=0, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, **cfg):
        self.cfg = dict(DEFAULTS)
        self.cfg.update(cfg)
        c = A.DnConfig(self.cfg["bins"], self.cfg["patch"], self.cfg["width"], self.cfg["blocks"],
                       self.cfg["n_desc"], self.cfg["max_batch"], beta1, beta2, eps, weight_decay)
        h = C.c_void_p()
        rc = A.lib().pach_dn_create(C.byref(h), device, C.byref(c))
        if rc != 0:
            raise A.PachError(rc, "pach_dn_create")
        self.h = h
        self.n_params = int(A.lib().pach_dn_param_count(h))
        lay, n = layout(**self.cfg)
        assert n == self.n_params, (n, self.n_params)

    def close(self):
        if self.h:
            A.lib().pach_dn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise A.PachError(rc, A.lib().pach_dn_last_error(self.h).decode())

    @staticmethod
    def _f(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def set_params(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        assert v.size == self.n_params
        self._chk(A.lib().pach_dn_set_params(self.h, self._f(v)))

    def get_params(self):
        v = np.zeros(self.n_params, dtype=np.float32)
        self._chk(A.lib().pach_dn_get_params(self.h, self._f(v)))
        return v

    def _desc(self, desc, B):
        if self.cfg["n_desc"] == 0:
            return None
        return np.ascontiguousarray(np.asarray(desc, dtype=np.float32).reshape(B, self.cfg["n_desc"]))

    def forward(self, x, desc):
        """x [B, 8, bins] normalised; returns the network output [B, 8, bins]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B = x.shape[0]
        d = self._desc(desc, B)
        out = np.zeros_like(x)
        mb = self.cfg["max_batch"]
        for s in range(0, B, mb):
            e = min(B, s + mb)
            xs = np.ascontiguousarray(x[s:e])
            ds = None if d is None else np.ascontiguousarray(d[s:e])
            os_ = np.zeros_like(xs)
            self._chk(A.lib().pach_dn_forward(self.h, self._f(xs), None if ds is None else self._f(ds), e - s,
                                              self._f(os_)))
            out[s:e] = os_
        return out

    def loss_grad(self, x, desc, target, grad=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        d = self._desc(desc, x.shape[0])
        loss = C.c_double()
        g = np.zeros(self.n_params, dtype=np.float32) if grad else None
        self._chk(A.lib().pach_dn_loss_grad(self.h, self._f(x), None if d is None else self._f(d), self._f(t),
                                            x.shape[0], C.byref(loss), None if g is None else self._f(g)))
        return loss.value, g

    def set_loss_weights(self, w):
        """Per-bin loss weights [bins] (None: uniform)."""
        if w is None:
            self._chk(A.lib().pach_dn_set_loss_weights(self.h, None))
            return
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.cfg["bins"]
        self._chk(A.lib().pach_dn_set_loss_weights(self.h, self._f(w)))

    def set_dataset(self, x, desc, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        d = self._desc(desc, x.shape[0])
        self._chk(A.lib().pach_dn_dataset(self.h, self._f(x), None if d is None else self._f(d), self._f(t),
                                          x.shape[0]))
        self.ds_n = x.shape[0]

    def train_step(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        loss = C.c_double()
        self._chk(A.lib().pach_dn_train_step(self.h, idx.ctypes.data_as(C.POINTER(C.c_int64)), idx.size,
                                             float(lr), C.byref(loss)))
        return loss.value

    def gemm(self, a_bits, b_bits, M, N, K, ta=False, tb=False):
        a = np.ascontiguousarray(a_bits, dtype=np.uint16)
        b = np.ascontiguousarray(b_bits, dtype=np.uint16)
        c = np.zeros((M, N), dtype=np.float32)
        self._chk(A.lib().pach_dn_gemm(self.h, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                       b.ctypes.data_as(C.POINTER(C.c_uint16)), M, N, K, int(ta), int(tb),
                                       self._f(c)))
        return c

    def save(self, path, meta=None):
        keys = sorted(self.cfg)
        np.savez_compressed(path, params=self.get_params(), cfg_keys=np.array(keys),
                            cfg=np.array([self.cfg[k] for k in keys]), meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device=0, max_batch=None):
        z = np.load(path, allow_pickle=False)
        cfg = {str(k): int(v) for k, v in zip(z["cfg_keys"], z["cfg"])}
        if max_batch:
            cfg["max_batch"] = max_batch
        dn = cls(device=device, **cfg)
        dn.set_params(z["params"])
        return dn

    def denoise(self, E, desc, rays=None):
        """Energy histograms [B, 8, bins] (scaled) -> denoised energy histograms.
        desc: [B or 1, n_desc] descriptors; a network that also takes the ray
        count (n_desc 13) accepts the 12 room descriptors plus `rays`."""
        x, S = normalize(E)
        d = np.asarray(desc, dtype=np.float32)
        d = d.reshape(-1, d.shape[-1])
        if d.shape[1] == self.cfg["n_desc"] - 1:
            if rays is None:
                raise ValueError("this network also takes the trace's source-ray count (rays=)")
            d = np.concatenate([d, np.full((d.shape[0], 1), ray_descriptor(rays), dtype=np.float32)], axis=1)
        if d.shape[1] != self.cfg["n_desc"]:
            raise ValueError(f"descriptors have {d.shape[1]} columns, the network takes {self.cfg['n_desc']}")
        y = self.forward(x, np.broadcast_to(d, (x.shape[0], self.cfg["n_desc"])))
        return denormalize(y, S)


# ---------------------------------------------------------------- training pairs

def trace_energy(tr, packed, rays, cutoff_s, seed, precision=A.PREC_FP32, is_order=-1):
    """Scaled energy histograms of every receiver of a packed batch, [R, 8, bins]."""
    A.trace_batch(tr, packed, rays, cutoff_s=cutoff_s, seed=seed, precision=precision, is_order=is_order)
    return tr.histogram(0, scaled=True)


def direct_part(room, fs=1000):
    """Deterministic direct sound of a room's receivers: (energy [R, 8] with
    zeros where the path is blocked, arrival bin [R] = ceil(d / c * fs)), what
    IR_Construction.ETCurve adds to a trace made with IS_Order = 0."""
    E, t, aud = A.direct_sound(room)
    return E * aud[:, None], np.ceil(t * fs).astype(np.int64)


def with_direct(E, dE, dbin):
    """Energy histograms [N, 8, bins] plus their direct sound (ETCurve)."""
    out = np.array(E, dtype=np.float64, copy=True)
    for i in range(out.shape[0]):
        if dbin[i] < out.shape[-1]:
            out[i, :, dbin[i]] += dE[i]
    return out


def make_pairs(tr, room_list, lo_rays, hi_rays, cutoff_s=3.0, lo_repeats=1, seed=0, pack=16, log=None, is_order=-1,
               ray_desc=False, room_ids=None):
    """Traces every room at lo_rays (lo_repeats independent seeds each; lo_rays
    may also be a list of ray counts, one input per entry) and hi_rays.
    Returns dict(x [N, 8, bins] normalised low-ray input, t [N, 8, bins]
    normalised high-ray target (same S), S [N, 8, 1] the input's band
    energies, desc [N, n_desc] (with ray_desc, the input's ray count as a
    13th column), lo [N] the input's ray count, room [N], rec [N], the direct
    sound dE [N, 8] and its bin dbin [N] for the ETCurve)."""
    los = [int(r) for r in np.atleast_1d(lo_rays) for _ in range(lo_repeats)]
    ids = list(range(len(room_list))) if room_ids is None else list(room_ids)
    xs, ts, ss, ds, ri, rr, des, dbs, lr = [], [], [], [], [], [], [], [], []
    for g0 in range(0, len(room_list), pack):
        grp = room_list[g0:g0 + pack]
        pk = RM.pack_rooms(grp)
        n = len(grp)
        t0 = time.perf_counter()
        Ehi = trace_energy(tr, pk, np.full(n, hi_rays), cutoff_s, seed + 7919 * (g0 + 1), is_order=is_order)
        Elo = [trace_energy(tr, pk, np.full(n, lo), cutoff_s, seed + 7919 * (g0 + 1) + 104729 * (k + 1),
                            is_order=is_order)
               for k, lo in enumerate(los)]
        r0 = 0
        for i, room in enumerate(grp):
            k = int(pk.rec_count[i])
            dE, db = direct_part(room)
            for lo, E in zip(los, Elo):
                d = descriptors(room, rays=lo if ray_desc else None)
                x, S = normalize(E[r0:r0 + k])
                sch_hi = np.cumsum(Ehi[r0:r0 + k][..., ::-1], axis=-1)[..., ::-1]
                t = (np.log10(np.maximum(sch_hi / S, FLOOR)) + 5.0) / 2.5
                xs.append(x)
                ss.append(S)
                ts.append(t.astype(np.float32))
                ds.append(d)
                des.append(dE)
                dbs.append(db)
                lr.append(np.full(k, lo))
                ri.append(np.full(k, ids[g0 + i]))
                rr.append(np.arange(k))
            r0 += k
        if log:
            log(f"pairs: rooms {g0}-{g0 + n - 1} traced in {time.perf_counter() - t0:.1f} s")
    return dict(x=np.concatenate(xs), t=np.concatenate(ts), S=np.concatenate(ss), desc=np.concatenate(ds),
                lo=np.concatenate(lr), room=np.concatenate(ri), rec=np.concatenate(rr), dE=np.concatenate(des),
                dbin=np.concatenate(dbs))


def merge_pairs(*sets):
    """Concatenates make_pairs results."""
    return {k: np.concatenate([s[k] for s in sets]) for k in sets[0]}


def early_weights(bins, fs=1000, gain=20.0, tau_s=0.15):
    """Loss weights 1 + gain exp(-t / tau): the first few hundred ms (EDT,
    C80) weigh as much as the rest of the decay."""
    t = np.arange(bins) / fs
    return (1.0 + gain * np.exp(-t / tau_s)).astype(np.float32)


def train(dn, data, steps, batch=256, lr=1e-3, warmup=100, seed=0, log=None, log_every=100, val=None,
          eval_every=250, weights=None):
    """AdamW with linear warm-up and cosine decay.  With validation pairs the
    parameters with the lowest validation loss (checked every eval_every
    steps) are kept.  Returns (loss history, validation history)."""
    dn.set_dataset(data["x"], data["desc"], data["t"])
    vw = 1.0 if weights is None else weights
    dn.set_loss_weights(weights)
    rng = np.random.default_rng(seed)
    n = data["x"].shape[0]
    hist, vhist = [], []
    best = (np.inf, None, -1)
    t0 = time.perf_counter()
    for s in range(steps):
        a = min(1.0, (s + 1) / warmup) * 0.5 * (1 + np.cos(np.pi * s / steps))
        idx = rng.choice(n, size=min(batch, n), replace=False)
        loss = dn.train_step(idx, lr * a)
        hist.append(loss)
        if val is not None and ((s + 1) % eval_every == 0 or s == steps - 1):
            vl = float(np.mean(vw * (dn.forward(val["x"], val["desc"]) - val["t"]) ** 2))
            vhist.append((s + 1, vl))
            if vl < best[0]:
                best = (vl, dn.get_params(), s + 1)
            if log:
                log(f"step {s + 1}: validation loss {vl:.3e} (best {best[0]:.3e} at {best[2]})")
        if log and (s % log_every == 0 or s == steps - 1):
            log(f"train step {s}: loss {loss:.5f} ({time.perf_counter() - t0:.1f} s)")
    if best[1] is not None:
        dn.set_params(best[1])
    return np.array(hist), vhist


def decay_params(E, fs=1000, start_bin=None):
    """T30, EDT, C80 per (receiver, band) of energy histograms [R, 8, bins].
    C80 is AcousticalMath.Clarity (Classes_Math.cs:559-585) with its
    startTime at bin start_bin[r] (the direct sound's arrival, ETCurve's
    ceil(d / c fs)), so the 80 ms early window opens with the direct sound as
    ISO 3382-1 defines it; with startTime 0 on these absolute-time
    histograms the window would end before the direct sound of every
    receiver more than 27 m from the source."""
    out = np.full(E.shape[:2] + (3,), np.nan)
    for r in range(E.shape[0]):
        t0 = 0.0 if start_bin is None else (float(start_bin[r]) + 1e-6) / fs
        for b in range(E.shape[1]):
            if E[r, b].sum() <= 0:
                continue
            sch = A.schroeder_integral(E[r, b])
            out[r, b] = (A.t_x(sch, 30, fs), A.early_decay_time(sch, fs), A.clarity(E[r, b], fs, start_time=t0))
    return out


def errors(P, Pref):
    """Relative T30 / EDT errors and absolute C80 error (dB) against a reference."""
    with np.errstate(invalid="ignore"):
        return dict(t30=np.abs(P[..., 0] - Pref[..., 0]) / Pref[..., 0],
                    edt=np.abs(P[..., 1] - Pref[..., 1]) / Pref[..., 1],
                    c80=np.abs(P[..., 2] - Pref[..., 2]))


def summarize(err, t30=0.05, edt=0.05, c80=1.0):
    """Mean / median / 90th percentile of each error and the share of
    (receiver, band) pairs inside the just-noticeable differences."""
    lim = dict(t30=t30, edt=edt, c80=c80)
    out = {}
    for k, e in err.items():
        e = e[np.isfinite(e)]
        out[k] = dict(mean=float(e.mean()), median=float(np.median(e)), p90=float(np.percentile(e, 90)),
                      within=float((e <= lim[k]).mean()))
    return out


def trace_with_progress(tr, log, every_s=20.0, **kw):
    """pach_trace on the context's worker thread, logging progress; returns wall seconds."""
    t0 = time.perf_counter()
    tr.trace(async_=True, **kw)
    last = t0
    while tr.poll()["running"]:
        time.sleep(0.05)
        if time.perf_counter() - last > every_s:
            p = tr.poll()
            log(f"trace: {p['rays_done']} of {p['rays_total']} source rays "
                f"({time.perf_counter() - t0:.0f} s)")
            last = time.perf_counter()
    rc = tr.wait()
    if rc != 0:
        raise A.PachError(rc, A.lib().pach_last_error(tr.h).decode())
    return time.perf_counter() - t0
