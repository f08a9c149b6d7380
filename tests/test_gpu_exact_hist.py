"""Exact (order-independent) histogram accumulation, pach_exact_histograms.

The reference adds every deposit into shared FP64 histograms from many tasks
(Directional_Histogram.Add, Classes_Receivers.cs:1159-1166, called from the
task loop of TaskAll:217-282), so its sums depend on the schedule; the FP64
atomics of the default GPU path do too, in the last bits.  With exact
accumulation every deposit becomes an integer (scaled by a power of two) added
in 32-bit limbs, so:

  * FP64 geometry: every bin of all 11 planes equals the oracle's within
    1e-12 relative (the oracle sums in ray order; the integer sum is exact and
    rounded once), bounce and deposit counts equal;
  * any launch shape, workgroup size, batch size or register cap gives
    bit-for-bit the same histograms (FP32 hall, FP64 hall);
  * batched rooms: emptying one room leaves the bins of the rooms before it
    bit-for-bit unchanged;
  * a convergence-mode trace equals the plain trace of its ray count bit for
    bit (the fixed-point unit depends on the scene, not the ray count);
  * results agree with the FP64-atomic path within 1e-12 relative.
"""
import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu

DEFAULT_KNOBS = {"walk_flat": 2, "walk_steps": 12, "shade_min": 16, "refill_min": 1, "dump_min": 256,
                 "block": 256, "occupancy": -1, "shade_queue": 0}


@pytest.fixture(scope="module")
def tracer():
    t = A.Tracer(0)
    yield t
    t.exact_histograms(False)
    t.close()


def close(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-300)
    bad = np.abs(a - b) > rtol * np.abs(b) + 1e-15 * scale
    return int(bad.sum())


@pytest.mark.parametrize("room_fn,rays,cut", [(rooms.shoebox, 64, 1.0), (rooms.fan_hall, 24, 0.8)])
def test_exact_fp64_equals_the_oracle(tracer, room_fn, rays, cut):
    room = room_fn()
    at, c, rho = A.uniform_medium()
    print(f"[case] exact FP64 {room.name} rays={rays}", flush=True)
    tracer.exact_histograms(True)
    try:
        st = A.trace_room(tracer, room, rays=rays, cutoff_s=cut, seed=5, precision=A.PREC_FP64)
        g = tracer.planes(scaled=False)
    finally:
        tracer.exact_histograms(False)
    ora, cnt = O.trace(room, rays=rays, cutoff_s=cut, seed=5, medium=O.Medium(c=c, atten=at, rho=rho), threads=8)
    assert st["bounces"] == cnt["bounces_broad"] + cnt["bounces_late"] + cnt["bounces_verylate"]
    assert st["deposits"] == cnt["deposits"]
    for k in range(A.NPLANES):
        assert close(g[k], ora[k], 1e-12) == 0, f"plane {k}"
    assert np.abs(g[0]).sum() > 0


@pytest.mark.parametrize("precision,rays", [(A.PREC_FP32, 4000), (A.PREC_FP64, 1500)])
def test_exact_histograms_do_not_depend_on_the_launch(tracer, precision, rays):
    room = rooms.fan_hall()
    shapes = [{}, {"block": 512, "refill_min": 16, "dump_min": 8, "walk_steps": 3},
              {"occupancy": 0, "shade_queue": 1, "shade_min": 8}]
    out = []
    tracer.exact_histograms(True)
    try:
        for i, knobs in enumerate(shapes):
            print(f"[case] exact launch independence prec={precision} shape={knobs}", flush=True)
            for k, v in knobs.items():
                tracer.tune(k, v)
            A.trace_room(tracer, room, rays=rays, cutoff_s=1.0, seed=21, precision=precision,
                         batch_rays=[0, 700, 1 << 20][i])
            out.append(tracer.planes(scaled=False))
            for k, v in DEFAULT_KNOBS.items():
                tracer.tune(k, v)
    finally:
        tracer.exact_histograms(False)
        for k, v in DEFAULT_KNOBS.items():
            tracer.tune(k, v)
    for o in out[1:]:
        assert np.array_equal(o, out[0])
    # the FP64-atomic path agrees to rounding
    A.trace_room(tracer, room, rays=rays, cutoff_s=1.0, seed=21, precision=precision)
    atom = tracer.planes(scaled=False)
    for k in range(A.NPLANES):
        assert close(atom[k], out[0][k], 1e-12) == 0, f"plane {k}"


def test_exact_batch_isolation_is_bitwise(tracer):
    rng = np.random.default_rng(9)
    rl = [rooms.shoebox(*rng.uniform(4, 12, 3), alpha=rng.uniform(0.05, 0.3, (6, 8)),
                        scat=rng.uniform(0.1, 0.6, (6, 8)), source=(1.0, 1.0, 1.0),
                        receivers=[(2.0, 2.5, 1.2), (3.0, 1.5, 1.6)], name=f"iso{i}") for i in range(5)]
    pk = rooms.pack_rooms(rl)
    rays = np.array([40, 25, 30, 35, 20])
    print(f"[case] exact batch isolation rays={list(rays)}", flush=True)
    tracer.exact_histograms(True)
    try:
        A.trace_batch(tracer, pk, rays, cutoff_s=0.4, seed=2, precision=A.PREC_FP32)
        full = tracer.planes(scaled=False)
        r2 = rays.copy()
        r2[2] = 0
        A.trace_batch(tracer, pk, r2, cutoff_s=0.4, seed=2, precision=A.PREC_FP32)
        part = tracer.planes(scaled=False)
    finally:
        tracer.exact_histograms(False)
    k0 = int(pk.rec_count[:2].sum())
    assert np.array_equal(part[:, :k0], full[:, :k0])
    assert np.all(part[:, k0:k0 + int(pk.rec_count[2])] == 0)
    assert full[0, :k0].sum() > 0


def test_exact_convergence_trace_equals_plain_trace(tracer):
    """Shoebox 20 x 12 x 6 m with a block hiding receiver 1 from the source:
    Minimum_Convergence_Check on the farthest visible and occluded receivers."""
    box = rooms.shoebox(20, 12, 6, alpha=0.15, scat=0.3, source=(3, 6, 1.5),
                        receivers=[(16, 10.5, 1.5), (14, 6, 1.5), (9, 2, 1.5)], radius=0.5)
    block = rooms._box_tris((10.0, 4.0, 0.0), (12.0, 8.0, 5.0), bottom=False)
    S = 6 + block.shape[0] // 2
    room = rooms.Room(name="box_with_block", verts=np.concatenate([box.verts, block]),
                      surf=np.concatenate([box.surf, 6 + np.arange(block.shape[0], dtype=np.int32) // 2]),
                      alpha=np.full((S, 8), 0.15), scat=np.full((S, 8), 0.3), trans=np.zeros((S, 8)),
                      source=box.source, receivers=box.receivers, radius=0.5)
    tracer.exact_histograms(True)
    try:
        print("[case] exact convergence mode", flush=True)
        st = A.trace_room(tracer, room, rays=-1, cutoff_s=0.6, seed=3, precision=A.PREC_FP32)
        conv = tracer.planes(scaled=False)
        n = st["rays_done"]
        assert st["converged"] == 1 and n == 80 * st["revolutions"] and n >= 160
        A.trace_room(tracer, room, rays=n, cutoff_s=0.6, seed=3, precision=A.PREC_FP32)
        plain = tracer.planes(scaled=False)
    finally:
        tracer.exact_histograms(False)
    assert np.array_equal(conv, plain)


def test_exact_mode_rejects_a_map(tracer):
    room = rooms.shoebox()
    at, c, rho = A.uniform_medium()
    tracer.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
    tracer.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tracer.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 300)
    tracer.set_map(np.array([[3.0, 3.0, 1.2], [5.0, 4.0, 1.2]]), 0.5, True, 300)
    tracer.exact_histograms(True)
    try:
        with pytest.raises(A.PachError):
            tracer.trace(rays=8, cutoff_m=0.3 * c, seed=1)
    finally:
        tracer.exact_histograms(False)
        tracer.clear_map()
