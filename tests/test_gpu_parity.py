"""GPU parity of the HIP tracer against the FP64 oracle (through the C ABI).

Tolerances:
  * FP64 geometry, deterministic specular mode: every histogram bin within
    1e-5 relative of the oracle (the bar for the deterministic mode), and
    bounce / deposit counters equal.
  * FP64 geometry, stochastic mode: the same per-ray Philox streams drive both,
    so bins again agree within 1e-5 relative.
  * FP32 geometry, stochastic mode: T30 and EDT within 2 % and C80 within
    0.2 dB per band of the FP64 result on the same seeds.
"""
import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tracer():
    t = A.Tracer(0)
    yield t
    t.close()


@pytest.fixture(scope="module")
def medium():
    at, c, rho = A.uniform_medium()
    return at, c, rho


def seeded_dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def run_both(tracer, medium, room, *, rays, cut, seed, mode, precision, dir_mode=A.DIR_GEODESIC,
             dirs=None, grid=(0, 0, 0), **kw):
    at, c, rho = medium
    print(f"[case] {room.name} rays={rays} cut={cut} seed={seed} mode={mode} prec={precision}",
          flush=True)
    st = A.trace_room(tracer, room, rays=rays, cutoff_s=cut, seed=seed, mode=mode,
                      precision=precision, dir_mode=dir_mode, dirs=dirs, grid=grid,
                      medium=medium, **kw)
    g = tracer.planes(scaled=False)
    ora, cnt = O.trace(room, rays=rays, cutoff_s=cut, seed=seed, mode=mode, dir_mode=dir_mode,
                       dirs=dirs, medium=O.Medium(c=c, atten=at, rho=rho), threads=8, **kw)
    return st, g, cnt, ora


def assert_bins(g, o, rtol, what):
    scale = max(np.abs(o).max(), 1e-300)
    diff = np.abs(g - o)
    bad = diff > rtol * np.abs(o) + 1e-12 * scale
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} bins off; worst rel "
                           f"{(diff / np.maximum(np.abs(o), 1e-300))[bad].max():.3e}")


def bounces(cnt):
    return cnt["bounces_broad"] + cnt["bounces_late"] + cnt["bounces_verylate"]


@pytest.mark.parametrize("room_fn,rays,cut", [(rooms.shoebox, 64, 2.0), (rooms.fan_hall, 48, 1.5)])
def test_specular_fp64_per_bin(tracer, medium, room_fn, rays, cut):
    room = room_fn()
    dirs = seeded_dirs(rays, 11)
    st, g, cnt, ora = run_both(tracer, medium, room, rays=rays, cut=cut, seed=3,
                               mode=A.MODE_SPECULAR, precision=A.PREC_FP64,
                               dir_mode=A.DIR_SUPPLIED, dirs=dirs)
    assert st["bounces"] == bounces(cnt)
    assert st["deposits"] == cnt["deposits"]
    assert st["deposits"] > 0
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")


# Directions with zero components: the walk's axis exits are infinite on those
# axes (selects, not branches), and its moves and empty-block jumps must not
# step along them.  In the box, no path of these directions meets an edge or
# corner within the cut-off (a path that did would slip through the crack in
# FP64 and not in FP32, whose edges are widened).
AXIS_DIRS = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                      (1, 1, 0), (1, 0, 1), (0, -1, 1), (-1, 1, 0)], dtype=np.float64)
AXIS_DIRS /= np.linalg.norm(AXIS_DIRS, axis=1, keepdims=True)


def axis_box():
    # source off the faces' diagonal planes (a ray along an axis would hit a
    # triangle edge exactly); receivers on the +x, +y and (1, 1, 0) lines
    src = (2.3, 3.1, 1.7)
    return rooms.shoebox(source=src, receivers=[(7.0, 3.1, 1.7), (2.3, 6.5, 1.7), (5.0, 5.8, 1.7)])


@pytest.mark.parametrize("room_fn", [axis_box, rooms.fan_hall])
def test_axis_aligned_rays(tracer, medium, room_fn):
    room = room_fn()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=len(AXIS_DIRS), cut=1.0, seed=3,
                               mode=A.MODE_SPECULAR, precision=A.PREC_FP64,
                               dir_mode=A.DIR_SUPPLIED, dirs=AXIS_DIRS)
    assert st["bounces"] == bounces(cnt)
    assert st["deposits"] == cnt["deposits"]
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")
    if room_fn is not axis_box:
        return
    assert st["deposits"] > 0
    # FP32 geometry in the box: reflections off axis-aligned walls keep these
    # directions exact and both triangles of a face share its normal, so the
    # paths, bounces and deposits are those of FP64
    st32 = A.trace_room(tracer, room, rays=len(AXIS_DIRS), cutoff_s=1.0, seed=3, mode=A.MODE_SPECULAR,
                        precision=A.PREC_FP32, dir_mode=A.DIR_SUPPLIED, dirs=AXIS_DIRS, medium=medium)
    g32 = tracer.planes(scaled=False)
    assert st32["bounces"] == st["bounces"]
    assert st32["deposits"] == st["deposits"]
    np.testing.assert_allclose(g32[0].sum(), g[0].sum(), rtol=1e-4)


@pytest.mark.parametrize("room_fn,rays,cut", [(rooms.shoebox, 6, 0.6), (rooms.fan_hall, 12, 1.0)])
def test_stochastic_fp64_per_bin(tracer, medium, room_fn, rays, cut):
    room = room_fn()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=rays, cut=cut, seed=5,
                               mode=A.MODE_STOCHASTIC, precision=A.PREC_FP64)
    assert st["bounces"] == bounces(cnt)
    assert st["octave_rays"] == cnt["octave_rays"]
    assert st["deposits"] == cnt["deposits"]
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")
    assert abs(st["lost"] - cnt["lost"]) <= 1e-9 * max(cnt["e_init"], 1)


def decay_params(E, fs=1000):
    """T30, EDT, C80 per (receiver, band) from an energy histogram [R, 8, n]."""
    out = np.zeros(E.shape[:2] + (3,))
    for r in range(E.shape[0]):
        for b in range(E.shape[1]):
            sch = A.schroeder_integral(E[r, b])
            out[r, b] = (A.t_x(sch, 30, fs), A.early_decay_time(sch, fs), A.clarity(E[r, b], fs))
    return out


def test_fp32_matches_fp64_statistics(tracer, medium):
    room = rooms.shoebox()
    kw = dict(rays=1000, cutoff_s=1.0, seed=9, medium=medium)
    A.trace_room(tracer, room, precision=A.PREC_FP64, **kw)
    E64 = tracer.histogram(0)
    s32 = A.trace_room(tracer, room, precision=A.PREC_FP32, **kw)
    E32 = tracer.histogram(0)
    p64, p32 = decay_params(E64), decay_params(E32)
    rel = np.abs(p32[..., :2] - p64[..., :2]) / p64[..., :2]
    assert rel.max() < 0.02, rel.max()
    assert np.abs(p32[..., 2] - p64[..., 2]).max() < 0.2
    assert s32["lost"] / s32["e_init"] < 1e-5


DEFAULT_KNOBS = {"walk_flat": 2, "walk_steps": 12, "receiver_masks": 1, "dist_lds": 1, "shade_min": 16,
                 "refill_min": 1, "dump_min": 256, "block": 256, "occupancy": -1, "shade_queue": 0,
                 "exact_queue": 1 << 25}


@pytest.mark.parametrize("knobs", [{"walk_flat": 0, "walk_steps": 4}, {"receiver_masks": 0, "dist_lds": 0},
                                   {"walk_steps": 1, "shade_min": 64},
                                   {"refill_min": 24, "block": 512, "dump_min": 16},
                                   {"occupancy": 4, "refill_min": 8, "shade_min": 8},
                                   {"shade_queue": 1}, {"occupancy": 5}, {"occupancy": 0},
                                   {"walk_flat": 1}, {"walk_flat": 2, "walk_steps": 5, "shade_queue": 1},
                                   {"shade_queue": 1, "block": 512, "refill_min": 16, "dump_min": 8,
                                    "walk_steps": 3}])
def test_launch_tuning_does_not_change_results(tracer, medium, knobs):
    room = rooms.fan_hall()
    try:
        for k, v in knobs.items():
            tracer.tune(k, v)
        st, g, cnt, ora = run_both(tracer, medium, room, rays=10, cut=0.8, seed=13,
                                   mode=A.MODE_STOCHASTIC, precision=A.PREC_FP64)
    finally:
        for k, v in DEFAULT_KNOBS.items():
            tracer.tune(k, v)
    assert st["bounces"] == bounces(cnt)
    assert st["deposits"] == cnt["deposits"]
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")


@pytest.mark.parametrize("is_order", [0, 1])
def test_is_order_gates_the_broadband_receiver_checks(tracer, medium, is_order):
    """IS_Order >= 0 leaves the first orders to deterministic methods: the
    broadband phase records only bounces of higher order (TaskAll:368, :378)."""
    room = rooms.shoebox()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=8, cut=0.6, seed=31, mode=A.MODE_STOCHASTIC,
                               precision=A.PREC_FP64, is_order=is_order)
    assert st["deposits"] == cnt["deposits"]
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")
    # the direct sound (order 0) is gone: nothing arrives before the first reflection
    d = np.linalg.norm(room.receivers - room.source, axis=1).min()
    assert g[0][:, :, : int(d / medium[1] * 1000)].sum() == 0
