"""Pins of the oracle's Lambert split scheme and transmission
(Lambert_Scattering, Classes_Materials.cs:4463-4702; OctaveRay / BroadRay
SplitRay, Classes_Rays.cs:123-129, :345-351).

The single-event hooks run the oracle's own scatter functions (the ones
or_trace calls) on supplied uniforms, so every event's outcome is predicted
exactly from the reference's rules:
  * Lambert bounds: s < 0.25 -> (lo, hi) = (0.1, 0.4), else s -/+ |1 - s| / 5
    (:4473-4482), as literal known answers worked out by hand;
  * Scatter_Late (:4629-4702): selector u > hi -> specular, no split;
    lo < u <= hi -> a specular child with I (1 - s), the parent keeps I s and
    goes diffuse; u <= lo -> the parent goes diffuse with all of I; a
    transmitted event reverses the outgoing directions;
  * Scatter_VeryLate (:4572-4627): diffuse when u < s (not lo / hi), no split;
  * Scatter_Early (:4498-4570): per band a diffuse child E s, the ray keeps
    E (1 - s); with transmission tau the ray splits off tr = E (1 - s) tau
    along the incoming direction and the child splits off td = E s tau in the
    reversed child direction.
A swapped child / parent energy, a dropped clamp, a wrong outcome boundary or
band index fails one of them.  The end-to-end test traces a transmissive panel
with or_trace and checks the energy the panel passes straight through against
the same trace with no panel, to 1e-9.
"""
import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import rooms

TWO32 = 4294967296.0

# (s, lo, hi) worked out by hand from Classes_Materials.cs:4473-4482
KAT = [(0.0, 0.1, 0.4), (0.2, 0.1, 0.4), (0.2499, 0.1, 0.4), (0.25, 0.1, 0.4), (0.3, 0.16, 0.44),
       (0.5, 0.4, 0.6), (0.6, 0.52, 0.68), (0.9, 0.88, 0.92), (1.0, 1.0, 1.0)]


def scene_with(scat, trans=0.0, alpha=0.1):
    """Shoebox whose polygon 0 carries scattering scat[b] (and trans[b]) per band."""
    box = rooms.shoebox()
    S = box.alpha.shape[0]
    sc = np.full((S, 8), 0.5)
    sc[0] = scat
    tr = np.zeros((S, 8))
    tr[0] = trans
    box.scat, box.trans, box.alpha = sc, tr, np.full((S, 8), alpha)
    return O.Scene(box, grid=(2, 2, 2))


def test_lambert_bounds_known_answers():
    s_vals = np.array([k[0] for k in KAT])
    for start in range(0, len(KAT), 8):
        part = KAT[start:start + 8]
        scat = np.zeros(8)
        scat[:len(part)] = [k[0] for k in part]
        sc = scene_with(scat)
        for b, (s, lo, hi) in enumerate(part):
            got = sc.lambert_bounds(0, b)
            assert got == pytest.approx((lo, hi), abs=1e-15), (s, got)
    assert len(s_vals) == len(KAT)


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def is_mirror(d_in, d_out, N):
    """d_out is the mirror image of d_in in the plane of normal N."""
    return (abs(np.linalg.norm(d_out) - 1) < 1e-12 and abs(d_out @ N + d_in @ N) < 1e-12
            and np.linalg.norm(np.cross(d_out - d_in, N)) < 1e-12)


@pytest.mark.parametrize("s,lo,hi", [(0.2, 0.1, 0.4), (0.5, 0.4, 0.6), (0.9, 0.88, 0.92)])
def test_scatter_late_outcomes_per_event(s, lo, hi):
    band = 3
    scat = np.full(8, 0.7)
    scat[band] = s
    sc = scene_with(scat)
    rng = np.random.default_rng(17)
    N = unit([0.2, -0.3, 0.93])
    n_ev = 4000
    # selectors spread over [0, 1) plus the exact boundaries (u = lo, u = hi
    # belong to the lower branch: the reference tests u > hi, then u > lo)
    sel = rng.integers(0, 2**32, n_ev, dtype=np.uint64)
    sel[:2] = [int(round(lo * TWO32)), int(round(hi * TWO32))]
    counts = {"spec": 0, "split": 0, "diff": 0}
    e_diffuse = 0.0
    for i in range(n_ev):
        d_in = unit(rng.normal(size=3))
        if d_in @ N > 0:
            d_in = -d_in
        I = float(rng.uniform(0.1, 2.0))
        w = [rng.integers(0, 2**32), sel[i], rng.integers(0, 2**32), rng.integers(0, 2**32)]
        u = float(sel[i]) / TWO32
        (pd, pI, pmask, pband), child = sc.scatter_late(0, band, d_in, N, I, w, mask=0)
        assert pband == band
        if u > hi:
            counts["spec"] += 1
            assert child is None
            assert pI == I and pmask == 0
            assert is_mirror(d_in, pd, N)
        elif u > lo:
            counts["split"] += 1
            assert child is not None
            cd, cI, cmask, _ = child
            assert cI == pytest.approx(I * (1 - s), rel=1e-15)
            assert pI == pytest.approx(I * s, rel=1e-15)
            assert cmask == 0 and pmask == 3
            assert is_mirror(d_in, cd, N)
            assert pd @ N > 0 and abs(np.linalg.norm(pd) - 1) < 1e-12  # diffuse, back on the incoming side
            e_diffuse += pI / I
        else:
            counts["diff"] += 1
            assert child is None
            assert pI == I and pmask == 3
            assert pd @ N > 0
            e_diffuse += 1.0
    # the share of energy sent diffuse is lo + (hi - lo) s, not s
    assert e_diffuse / n_ev == pytest.approx(lo + (hi - lo) * s, abs=5 * 0.5 / np.sqrt(n_ev))
    assert counts["split"] / n_ev == pytest.approx(hi - lo, abs=5 * np.sqrt((hi - lo) / n_ev) + 1e-3)


def test_scatter_late_transmission_reverses_both_rays():
    band = 1
    scat = np.full(8, 0.5)
    sc = scene_with(scat)
    N = unit([0.0, 0.0, 1.0])
    d_in = unit([0.3, 0.4, -0.8])
    lo, hi = 0.4, 0.6
    # split branch (lo < u <= hi), transmitted
    w = [0, int(0.5 * TWO32), 123456789, 987654321]
    (pd, pI, pmask, _), child = sc.scatter_late(0, band, d_in, N, 1.0, w, transmit=True)
    (pd0, _, _, _), child0 = sc.scatter_late(0, band, d_in, N, 1.0, w, transmit=False)
    assert child is not None and child0 is not None
    np.testing.assert_array_equal(pd, -pd0)
    np.testing.assert_array_equal(child[0], -child0[0])
    assert pd @ N < 0 and child[0] @ N < 0  # both continue through the surface
    # specular branch
    w = [0, int(0.9 * TWO32), 1, 2]
    (pd, _, _, _), ch = sc.scatter_late(0, band, d_in, N, 1.0, w, transmit=True)
    assert ch is None
    np.testing.assert_allclose(pd, -(d_in - 2 * (d_in @ N) * N), atol=1e-15)


def test_scatter_verylate_selects_on_s_without_split():
    band = 5
    s = 0.3  # Lambert bounds (0.16, 0.44): the very-late rule must ignore them
    scat = np.full(8, 0.9)
    scat[band] = s
    sc = scene_with(scat)
    rng = np.random.default_rng(5)
    N = unit([0.6, 0.0, -0.8])
    for u in (0.0, 0.15, 0.299, 0.3, 0.301, 0.42, 0.5, 0.99):
        d_in = unit([0.1, 0.5, 0.4])
        w = [0, int(u * TWO32), rng.integers(0, 2**32), rng.integers(0, 2**32)]
        pd, pI, pmask, _ = sc.scatter_verylate(0, band, d_in, N, 0.77, w)
        assert pI == 0.77
        if float(w[1]) / TWO32 < s:
            assert pmask == 3 and (pd @ N) * (d_in @ N) < 0
        else:
            assert pmask == 0 and is_mirror(d_in, pd, N)
        pdt, _, _, _ = sc.scatter_verylate(0, band, d_in, N, 0.77, w, transmit=True)
        np.testing.assert_array_equal(pdt, -pd)


def test_scatter_early_splits_per_band():
    scat = np.array([0.0, 0.05, 0.2, 0.35, 0.5, 0.65, 0.8, 1.0])
    tau = np.array([0.0, 0.3, 0.0, 0.1, 0.5, 0.0, 0.25, 0.9])
    sc = scene_with(scat, tau)
    rng = np.random.default_rng(9)
    N = unit([0.1, 0.2, 0.97])
    d_in = unit([0.5, -0.2, 0.6])  # leaves through the normal's side: the normal is flipped
    E = rng.uniform(0.5, 2.0, 8)
    E[2] = 0.0  # a band with no energy spawns nothing
    w = rng.integers(0, 2**32, 16, dtype=np.uint64)
    d_out, E_out, ch = sc.scatter_early(0, d_in, N, E, w)
    assert is_mirror(d_in, d_out, N)
    k = 0
    total = E_out.copy()
    for b in range(8):
        s, t = scat[b], tau[b]
        if E[b] == 0:
            assert E_out[b] == 0
            continue
        if t > 0:
            (trd, trI, trm, trb), (tdd, tdI, tdm, tdb) = ch[k][:4], ch[k + 1][:4]
            R = ch[k + 2]
            k += 3
            assert trb == tdb == R[3] == b
            assert trI == pytest.approx(E[b] * (1 - s) * t, rel=1e-14, abs=1e-300)
            assert tdI == pytest.approx(E[b] * s * t, rel=1e-14, abs=1e-300)
            assert R[1] == pytest.approx(E[b] * s * (1 - t), rel=1e-14, abs=1e-300)
            assert E_out[b] == pytest.approx(E[b] * (1 - s) * (1 - t), rel=1e-14, abs=1e-300)
            np.testing.assert_array_equal(trd, d_in)        # straight through
            np.testing.assert_array_equal(tdd, -R[0])        # the reversed child direction
            assert trm == 0 and tdm == 3
            total[b] += trI + tdI + R[1]
        else:
            R = ch[k]
            k += 1
            assert R[3] == b
            assert R[1] == pytest.approx(E[b] * s, rel=1e-15, abs=1e-300)
            assert E_out[b] == pytest.approx(E[b] * (1 - s), rel=1e-15, abs=1e-300)
            total[b] += R[1]
        assert R[2] == 3
        assert (R[0] @ N) * (d_in @ N) < 0 or s == 0  # diffuse child on the incoming side
    assert k == len(ch)
    np.testing.assert_allclose(total, E, rtol=1e-14)


def panel_room(with_panel=True):
    """Square panel x = 5, |y|, |z| <= 5 (two triangles), source at the origin;
    receiver 0 behind the panel on the axis, receiver 1 in front."""
    if with_panel:
        quad = np.array([[5.0, -5, -5], [5.0, 5, -5], [5.0, 5, 5], [5.0, -5, 5]])
    else:  # a tiny polygon far off every path to the receivers
        quad = np.array([[0.0, 900, 900], [1e-4, 900, 900], [1e-4, 900 + 1e-4, 900], [0.0, 900 + 1e-4, 900]])
    tris = np.stack([quad[[0, 1, 2]], quad[[0, 2, 3]]])
    b = np.arange(8)
    return rooms.Room(name="panel", verts=tris, surf=np.zeros(2, np.int32),
                      alpha=(0.05 + 0.05 * b)[None, :], scat=(0.1 + 0.08 * b)[None, :],
                      trans=(0.1 + 0.05 * b)[None, :], source=np.zeros(3),
                      receivers=np.array([[10.0, 0.3, -0.2], [2.0, 1.0, 0.5]]), radius=0.5)


def test_transmissive_panel_passes_e_1ma_1ms_tau():
    """Energy the panel passes straight through (the specular-plane deposits
    behind it) is the free-field energy x (1 - alpha)(1 - s) tau per band."""
    kw = dict(rays=20000, cutoff_s=0.2, seed=4, mode=O.MODE_STOCHASTIC, grid=(4, 4, 4))
    with_p, cnt = O.trace(panel_room(True), **kw)
    free, _ = O.trace(panel_room(False), **kw)
    b = np.arange(8)
    factor = (1 - (0.05 + 0.05 * b)) * (1 - (0.1 + 0.08 * b)) * (0.1 + 0.05 * b)
    spec_behind = with_p[8, 0].sum(axis=1)
    free_behind = free[8, 0].sum(axis=1)
    assert free_behind.min() > 0
    np.testing.assert_allclose(spec_behind / free_behind, factor, rtol=1e-9)
    # the diffuse plane behind the panel holds only the diffusely transmitted
    # children; nothing reaches it in front except the reflected diffuse children
    assert with_p[9, 0].sum() > 0
    assert cnt["spawned"] > 0
