"""GPU checks of the denoiser kernels against the FP64 CPU oracle.

  * GEMM: every operand orientation, edge tiles and split K, against an FP64
    product of the same BF16 operands (FP32 accumulation: 1e-5 of the scale);
    tile counts of 30 and 36 (not multiples of the 8 XCDs the tile order
    deals over) cover the XCD-aware tile order.
  * Forward pass: the oracle with BF16 rounding at the same points; 2e-3
    absolute on outputs of order one.
  * Loss (with per-bin weights) and gradient: relative 2-norm error per
    parameter tensor below 2e-2 (BF16 operands; the rounding of an
    intermediate can flip one BF16 step).
  * AdamW training on a fixed batch lowers the loss.
"""
import numpy as np
import pytest

import denoiser_ref as R
from pachydermacoustic_universal_amd import denoise as DN

pytestmark = pytest.mark.gpu

SMALL = dict(bins=240, patch=24, width=64, blocks=2, n_desc=10)


def bf16_bits(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bits_to_f64(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def net():
    dn = DN.Denoiser(max_batch=8, **SMALL)
    yield dn
    dn.close()


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(136, 200, 96), (64, 128, 1088), (256, 72, 40), (1160, 264, 200), (264, 264, 2056)])
def test_gemm_matches_fp64(net, ta, tb, M, N, K):
    print(f"[case] gemm M={M} N={N} K={K} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(M + N + K)
    a = bf16_bits(rng.normal(size=(K, M) if ta else (M, K)))
    b = bf16_bits(rng.normal(size=(N, K) if tb else (K, N)))
    c = net.gemm(a, b, M, N, K, ta=ta, tb=tb)
    A_ = bits_to_f64(a)
    B_ = bits_to_f64(b)
    ref = (A_.T if ta else A_) @ (B_.T if tb else B_)
    np.testing.assert_allclose(c, ref, rtol=0, atol=1e-5 * np.abs(ref).max() + 1e-6)


def random_problem(seed, B=3):
    L = R.Layout(**SMALL)
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.1, L.n).astype(np.float32).astype(np.float64)
    x = rng.normal(0, 1, (B, 8, L.bins)).astype(np.float32).astype(np.float64)
    desc = rng.normal(0, 1, (B, L.n_desc)).astype(np.float32).astype(np.float64)
    tgt = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32).astype(np.float64)
    return L, params, x, desc, tgt


def test_forward_matches_oracle(net):
    L, params, x, desc, tgt = random_problem(1)
    print("[case] denoiser forward B=3", flush=True)
    net.set_params(params)
    out = net.forward(x, desc)
    ref, _ = R.forward(L, params, x, desc, bf16=True)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-3)


def test_gradient_matches_oracle(net):
    L, params, x, desc, tgt = random_problem(2)
    print("[case] denoiser loss/grad B=3", flush=True)
    w = np.random.default_rng(5).uniform(0.5, 3.0, L.bins).astype(np.float32)
    net.set_params(params)
    net.set_loss_weights(w)
    try:
        loss, g = net.loss_grad(x, desc, tgt)
    finally:
        net.set_loss_weights(None)
    rloss, rg = R.loss_grad(L, params, x, desc, tgt, bf16=True, w=w.astype(np.float64))
    assert abs(loss - rloss) <= 1e-3 * rloss
    for name, o in L.offsets.items():
        n = int(np.prod(L.shapes[name]))
        a, b = g[o:o + n].astype(np.float64), rg[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err < 2e-2, (name, err)


def test_training_lowers_the_loss(net):
    L, params, x, desc, tgt = random_problem(3, B=8)
    tgt = x + 0.5 * np.tanh(x)  # a learnable map of the input
    print("[case] denoiser adamw 40 steps", flush=True)
    net.set_params(DN.init_params(SMALL, seed=1))
    net.set_dataset(x, desc, tgt)
    idx = np.arange(8)
    first = net.train_step(idx, 1e-2)
    for _ in range(39):
        last = net.train_step(idx, 1e-2)
    assert last < 0.5 * first, (first, last)
