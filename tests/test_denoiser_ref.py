"""CPU pins of the denoiser oracle (oracle/denoiser_ref.py) and of the host
side of the denoiser module: the oracle's backward pass against central finite
differences, its im2col product against a direct convolution loop, GELU's
derivative, the parameter layout shared with the library, and the histogram
normalisation round trip."""
import numpy as np
import pytest

import denoiser_ref as R
from pachydermacoustic_universal_amd import denoise as DN


def tiny(seed=0, bins=48, patch=8, width=64, blocks=2, n_desc=3, B=2):
    L = R.Layout(bins, patch, width, blocks, n_desc)
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.2, L.n)
    x = rng.normal(0, 1, (B, 8, bins))
    desc = rng.normal(0, 1, (B, n_desc))
    tgt = rng.normal(0, 1, (B, 8, bins))
    return L, params, x, desc, tgt


def test_backward_matches_finite_differences():
    L, params, x, desc, tgt = tiny()
    w = np.random.default_rng(7).uniform(0.5, 2.0, L.bins)
    loss, g = R.loss_grad(L, params, x, desc, tgt, w=w)
    rng = np.random.default_rng(1)
    # every tensor gets a few probes
    probes = []
    for name, o in L.offsets.items():
        n = int(np.prod(L.shapes[name]))
        probes += list(o + rng.choice(n, size=min(n, 4), replace=False))
    h = 1e-6
    for i in probes:
        p1 = params.copy(); p1[i] += h
        p2 = params.copy(); p2[i] -= h
        fd = (R.loss_grad(L, p1, x, desc, tgt, w=w)[0] - R.loss_grad(L, p2, x, desc, tgt, w=w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_im2col_product_is_a_kernel3_convolution():
    L, params, x, desc, tgt = tiny(B=3)
    rng = np.random.default_rng(2)
    H = rng.normal(size=(3 * L.P, L.D))
    Wc = rng.normal(size=(3 * L.D, 16))
    np.testing.assert_allclose(R.im2col(L, H, 3) @ Wc, R.conv3_direct(L, H, Wc, 3), rtol=1e-12, atol=1e-12)
    # col2im is the adjoint of im2col: <im2col(H), G> == <H, col2im(G)>
    G = rng.normal(size=(3 * L.P, 3 * L.D))
    assert np.isclose(np.sum(R.im2col(L, H, 3) * G), np.sum(H * R.col2im(L, G, 3)), rtol=1e-12)


def test_gelu_derivative():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (R.gelu(x + h) - R.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(R.gelu_grad(x), fd, atol=1e-8)
    # the tanh form stays within 1e-3 of x * Phi(x)
    from scipy.special import erf
    np.testing.assert_allclose(R.gelu(x), 0.5 * x * (1 + erf(x / np.sqrt(2))), atol=1e-3)


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], dtype=np.float32)
    r = R.bf16_round(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5


def test_untrained_network_returns_its_input():
    L, params, x, desc, tgt = tiny()
    cfg = dict(bins=L.bins, patch=L.patch, width=L.D, blocks=L.NB, n_desc=L.n_desc)
    p = DN.init_params(cfg, seed=3).astype(np.float64)
    out, _ = R.forward(L, p, x, desc)
    np.testing.assert_allclose(out, x, atol=1e-12)


@pytest.mark.parametrize("cfg", [dict(bins=48, patch=8, width=64, blocks=2, n_desc=3),
                                 dict(bins=3000, patch=24, width=256, blocks=4, n_desc=10)])
def test_layout_matches_oracle(cfg):
    lay, n = DN.layout(**cfg)
    L = R.Layout(**cfg)
    assert n == L.n
    for k, (o, shp) in lay.items():
        assert L.offsets[k] == o and tuple(L.shapes[k]) == tuple(shp)


def test_default_network_size():
    _, n = DN.layout(**DN.DEFAULTS)
    assert 1.8e6 < n < 2.2e6


def test_normalisation_round_trip():
    rng = np.random.default_rng(4)
    E = rng.exponential(1.0, (3, 8, 300)) * np.exp(-np.arange(300) / 40.0)
    E[:, :, :5] = 0
    x, S = DN.normalize(E)
    # the input is the normalised log Schroeder curve: non-increasing, 2 at t = 0
    assert np.all(np.diff(x, axis=-1) <= 1e-6) and np.allclose(x[..., 0], 2.0)
    back = DN.denormalize(x, S)
    sch = np.cumsum(E[..., ::-1], axis=-1)[..., ::-1]
    sch_back = np.cumsum(back[..., ::-1], axis=-1)[..., ::-1]
    ok = sch / S > 1e-8
    np.testing.assert_allclose(sch_back[ok], sch[ok], rtol=1e-5)
    # bins are differences of the FP32-coded curve: exact to ~1e-6 of the curve
    assert np.all(np.abs(back - E)[ok] <= 1e-5 * sch[ok])
    assert np.all(back[:, :, :5] == 0) and np.all(back >= 0)
    np.testing.assert_allclose(back.sum(-1), E.sum(-1), rtol=1e-6)
