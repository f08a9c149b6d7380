"""The mixed-precision tracer (PACH_PREC_MIXED: FP64 rays, hits and shading,
an FP32 screen for the voxel walk) against the FP64 oracle and the FP64 GPU
path, on the 2,032-triangle hall the ray-bounce rate is measured on.

Tolerances (the bar for the deterministic mode): every bin of all 11 planes
within 1e-5 relative of the oracle, bounce and deposit counts equal.  Against
the FP64 GPU path the decisions are the same, so the counters are equal and
the bins agree up to the order of the FP64 atomic additions (1e-12).
"""
import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

from test_gpu_parity import AXIS_DIRS, assert_bins, axis_box, bounces, run_both, seeded_dirs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tracer():
    t = A.Tracer(0)
    yield t
    t.close()


@pytest.fixture(scope="module")
def medium():
    return A.uniform_medium()


def check_planes(st, g, cnt, ora):
    assert st["bounces"] == bounces(cnt)
    assert st["deposits"] == cnt["deposits"]
    assert st["octave_rays"] == cnt["octave_rays"]
    for k in range(A.NPLANES):
        assert_bins(g[k], ora[k], 1e-5, f"plane {k}")


def test_specular_hall_1000_rays_per_bin(tracer, medium):
    """Deterministic mode on the hall: 1,000 seeded supplied directions, 1.5 s."""
    room = rooms.fan_hall()
    dirs = seeded_dirs(1000, 21)
    st, g, cnt, ora = run_both(tracer, medium, room, rays=1000, cut=1.5, seed=3, mode=A.MODE_SPECULAR,
                               precision=A.PREC_MIXED, dir_mode=A.DIR_SUPPLIED, dirs=dirs, grid=(32, 32, 32))
    check_planes(st, g, cnt, ora)
    assert st["deposits"] > 1000
    print(f"  FP64 fallbacks {st['exact_rewalks']}, grazing departures {st['exact_walks']} "
          f"of {st['bounces']} bounces", flush=True)


@pytest.mark.parametrize("room_fn,rays,cut", [(rooms.shoebox, 8, 0.8), (rooms.fan_hall, 24, 1.0)])
def test_stochastic_per_bin(tracer, medium, room_fn, rays, cut):
    room = room_fn()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=rays, cut=cut, seed=5, mode=A.MODE_STOCHASTIC,
                               precision=A.PREC_MIXED)
    check_planes(st, g, cnt, ora)
    assert abs(st["lost"] - cnt["lost"]) <= 1e-9 * max(cnt["e_init"], 1)


def test_axis_aligned_rays(tracer, medium):
    # rays parallel to the walls: the FP32 screen cannot place those hits, the
    # walks fall back to FP64 decisions
    room = axis_box()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=len(AXIS_DIRS), cut=1.0, seed=3,
                               mode=A.MODE_SPECULAR, precision=A.PREC_MIXED, dir_mode=A.DIR_SUPPLIED,
                               dirs=AXIS_DIRS)
    check_planes(st, g, cnt, ora)


def transmissive_room():
    """The shoebox with a transmissive partition across it and one
    transmissive wall: rays split at both (Scatter_Early and Scatter_Late)."""
    box = rooms.shoebox()
    quad = np.array([[5.0, 0, 0], [5.0, 8, 0], [5.0, 8, 4], [5.0, 0, 4]])
    part = np.stack([quad[[0, 1, 2]], quad[[0, 2, 3]]])
    S = box.alpha.shape[0]
    b = np.arange(8)
    trans = np.zeros((S + 1, 8))
    trans[S] = 0.2 + 0.05 * b
    trans[2] = 0.1
    return rooms.Room(name="partition", verts=np.concatenate([box.verts, part]),
                      surf=np.concatenate([box.surf, [S, S]]).astype(np.int32),
                      alpha=np.concatenate([box.alpha, np.full((1, 8), 0.05)]),
                      scat=np.concatenate([box.scat, np.full((1, 8), 0.3)]), trans=trans,
                      source=box.source, receivers=np.array([[7.5, 4.0, 2.0], [2.5, 5.5, 1.5], [8.0, 1.5, 3.0]]),
                      radius=0.5)


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_MIXED])
def test_transmission_per_bin(tracer, medium, precision):
    room = transmissive_room()
    st, g, cnt, ora = run_both(tracer, medium, room, rays=12, cut=0.6, seed=8, mode=A.MODE_STOCHASTIC,
                               precision=precision)
    check_planes(st, g, cnt, ora)
    assert cnt["spawned"] > 0
    # energy reaches the far side of the partition only through it
    assert g[0][0].sum() > 0


def test_full_exact_queue_drains_without_losing_rays(tracer, medium):
    """An exact queue of 100,000 rays beyond its per-wave reserve: waves drain
    soon after they hand bounces to the FP64 walk, with input left unclaimed;
    that input must reach the next launch, and the trace must equal FP64's."""
    room = rooms.fan_hall()
    kw = dict(rays=30000, cutoff_s=1.0, seed=23, grid=(32, 32, 32), medium=medium)
    s64 = A.trace_room(tracer, room, precision=A.PREC_FP64, **kw)
    g64 = tracer.planes(scaled=False)
    try:
        tracer.tune("exact_queue", 100_000)
        smx = A.trace_room(tracer, room, precision=A.PREC_MIXED, **kw)
        gmx = tracer.planes(scaled=False)
    finally:
        tracer.tune("exact_queue", 1 << 25)
    for k in ("bounces", "deposits", "octave_rays", "spawned", "misses"):
        assert s64[k] == smx[k], k
    assert smx["launches_octave"] > s64["launches_octave"]
    scale = np.abs(g64).max()
    assert np.all(np.abs(gmx - g64) <= 1e-12 * np.abs(g64) + 1e-15 * scale)


def test_mixed_equals_fp64_on_the_hall(tracer, medium):
    room = rooms.fan_hall()
    kw = dict(rays=20000, cutoff_s=2.0, seed=17, grid=(32, 32, 32), medium=medium)
    print("[case] hall 2e4 rays FP64", flush=True)
    s64 = A.trace_room(tracer, room, precision=A.PREC_FP64, **kw)
    g64 = tracer.planes(scaled=False)
    print("[case] hall 2e4 rays mixed", flush=True)
    smx = A.trace_room(tracer, room, precision=A.PREC_MIXED, **kw)
    gmx = tracer.planes(scaled=False)
    for k in ("bounces", "deposits", "octave_rays", "spawned", "misses"):
        assert s64[k] == smx[k], k
    scale = np.abs(g64).max()
    assert np.all(np.abs(gmx - g64) <= 1e-12 * np.abs(g64) + 1e-15 * scale)
    assert smx["lost"] == pytest.approx(s64["lost"], rel=1e-12, abs=1e-300)
    rate = smx["exact_rewalks"] / smx["bounces"]
    print(f"  FP64 fallbacks per bounce {rate:.2e}, grazing departures {smx['exact_walks'] / smx['bounces']:.2e}",
          flush=True)
    assert rate < 0.01


@pytest.mark.parametrize("knobs", [{"walk_flat": 1, "walk_steps": 5}, {"block": 512, "refill_min": 16},
                                   {"occupancy": 0}, {"occupancy": 3}, {"receiver_masks": 0, "dist_lds": 0},
                                   {"exact_queue": 100_000}])
def test_mixed_launch_tuning_keeps_results(tracer, medium, knobs):
    from test_gpu_parity import DEFAULT_KNOBS
    room = rooms.fan_hall()
    try:
        for k, v in knobs.items():
            tracer.tune(k, v)
        st, g, cnt, ora = run_both(tracer, medium, room, rays=10, cut=0.8, seed=13, mode=A.MODE_STOCHASTIC,
                                   precision=A.PREC_MIXED)
    finally:
        for k, v in DEFAULT_KNOBS.items():
            tracer.tune(k, v)
    check_planes(st, g, cnt, ora)
