"""Mapping receivers on the GPU (pach_map_receivers_set, PachMapReceiver)
against the oracle's restatement (oracle.c or_map_*), same scene, seeds and
receivers:

* FP64 geometry, stochastic and specular modes: every bin of the 8 map
  planes (energy, pressure, directions) within 1e-9 relative of the oracle's
  (largest bin), and the same deposit count, for a plane map of 48 receivers
  over a shoebox floor and for a map crossing a blocked room;
* FP32 geometry on the same seed: after the first 30 ms (octave rays) the
  map's energy within 1 % of FP64 and each receiver's within 5 %.  The early
  part is left out: the broadband walk steers by the signs of the segment
  origin's coordinates, so a ray leaving the floor (z = 0 up to rounding)
  walks differently for z = +1e-16 and z = -1e-16, and FP32 and FP64 hit
  points round differently (FP64 reproduces the oracle exactly, above);
* clearing the map gives the spherical receivers back unchanged.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu

SRC = (2.0, 2.0, 1.5)


def plane_centers(nx=8, ny=6, z=1.2, x0=3.0, y0=2.5, pitch=1.0):
    xs, ys = np.meshgrid(x0 + pitch * np.arange(nx), y0 + pitch * np.arange(ny), indexing="ij")
    return np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, z)], 1)


def box():
    return rooms.shoebox(14, 10, 4, alpha=0.15, scat=0.3, source=SRC, receivers=[(7.3, 5.2, 1.4)], radius=0.5)


def run_pair(tr, room, centers, inc, *, rays, mode, seed=5, cutoff=0.6, directional=True, time1pt=False):
    at, c, rho = A.uniform_medium()
    med = O.Medium(c=c, atten=at, rho=rho)
    sct = int(cutoff * 1000)
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (8, 8, 8))
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, sct)
    tr.set_map(centers, inc, directional, sct, time1pt=time1pt)
    tr.trace(rays=rays, cutoff_m=cutoff * c, seed=seed, mode=mode, precision=A.PREC_FP64)
    got = tr.map_planes(scaled=False)
    st = tr.stats()
    m = O.Map(centers, inc, room.source, med, directional=directional, sample_ct=sct, time1pt=time1pt)
    want, cnt = O.trace(room, rays=rays, cutoff_s=cutoff, seed=seed, mode=mode, medium=med, map_rx=m,
                        grid=(8, 8, 8), threads=8)
    return got, want, st, cnt


@pytest.fixture(scope="module")
def tracer():
    tr = A.Tracer(0)
    yield tr
    tr.close()


@pytest.mark.parametrize("mode", [A.MODE_STOCHASTIC, A.MODE_SPECULAR])
def test_plane_map_matches_oracle(tracer, mode):
    print(f"[case] map receivers plane map mode={mode}", flush=True)
    got, want, st, cnt = run_pair(tracer, box(), plane_centers(), 1.0, rays=2000, mode=mode)
    assert want[0].sum() > 0
    assert st["deposits"] == cnt["deposits"]
    for k in range(8):
        np.testing.assert_allclose(got[k], want[k], rtol=0, atol=1e-9 * np.abs(want[k]).max() + 1e-300)
    tracer.clear_map()


def test_map_in_blocked_room_matches_oracle(tracer):
    print("[case] map receivers blocked room", flush=True)
    b = box()
    blk = rooms._box_tris((6.0, 3.0, 0.0), (8.0, 7.0, 3.0), bottom=False)
    S = 6 + blk.shape[0] // 2
    room = rooms.Room(name="blocked", verts=np.concatenate([b.verts, blk]),
                      surf=np.concatenate([b.surf, 6 + np.arange(blk.shape[0], dtype=np.int32) // 2]),
                      alpha=np.full((S, 8), 0.15), scat=np.full((S, 8), 0.3), trans=np.zeros((S, 8)),
                      source=b.source, receivers=b.receivers, radius=0.5)
    centers = plane_centers(nx=11, ny=8, x0=1.5, y0=1.2, pitch=1.1, z=1.3)
    got, want, st, cnt = run_pair(tracer, room, centers, 0.8, rays=1500, mode=A.MODE_STOCHASTIC,
                                  directional=False)
    assert st["deposits"] == cnt["deposits"]
    for k in (0, 1):
        np.testing.assert_allclose(got[k], want[k], rtol=0, atol=1e-9 * np.abs(want[k]).max())
    tracer.clear_map()


def test_fp32_map_close_to_fp64(tracer):
    print("[case] map receivers fp32", flush=True)
    room = box()
    at, c, rho = A.uniform_medium()
    centers = plane_centers()
    tracer.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (8, 8, 8))
    tracer.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tracer.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 600)
    tracer.set_map(centers, 1.0, True, 600)
    E = {}
    for prec in (A.PREC_FP64, A.PREC_FP32):
        tracer.trace(rays=100000, cutoff_m=0.6 * c, seed=9, precision=prec)
        E[prec] = tracer.map_histogram(0, scaled=True)[:, :, 30:].sum((1, 2))
    assert E[A.PREC_FP32].sum() == pytest.approx(E[A.PREC_FP64].sum(), rel=0.01)
    np.testing.assert_allclose(E[A.PREC_FP32], E[A.PREC_FP64], rtol=0.05)
    tracer.clear_map()
    # spherical receivers are back
    tracer.trace(rays=2000, cutoff_m=0.6 * c, seed=9, precision=A.PREC_FP64)
    assert tracer.histogram(0).sum() > 0


def test_split_ray_tracer_with_a_map_mesh():
    """PachMapReceiver over a 6 x 4 quad mesh at 1.2 m as RecMain of a
    SplitRayTracer: the map histograms equal a direct pach_map_receivers_set
    trace with the same receivers and seed, and the SPL map is the energy sum
    of each receiver's window in dB."""
    print("[case] map receivers through SplitRayTracer", flush=True)
    from pachydermacoustic_universal_amd import pachyderm as P
    room = box()
    sc = P.Polygon_Scene.from_room(room)
    src = P.GeodesicSource([120] * 8, room.source)
    nx, ny = 6, 4
    V = np.array([(3 + 1.5 * i, 2 + 1.5 * j, 0.2) for i in range(nx + 1) for j in range(ny + 1)])
    quads = [[i * (ny + 1) + j, (i + 1) * (ny + 1) + j, (i + 1) * (ny + 1) + j + 1, i * (ny + 1) + j + 1]
             for i in range(nx) for j in range(ny)]
    mr = P.PachMapReceiver((V, quads), src, 1000, 1.0, sc, 500.0)
    np.testing.assert_allclose(mr.Origins[:, 2], 1.2)  # polygon centres + unit normal (+z)
    rt = P.SplitRayTracer(src, mr, sc, 0.5 * sc.Sound_speed(0), [0, 7], -1, 5000, precision="fp64", seed=4)
    rt.Begin()
    rt.Combine_ThreadLocal_Results()
    got = mr.planes[0]
    with A.Tracer(0) as tr:
        at, c, rho = A.uniform_medium()
        tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
        tr.set_source(room.source, src.SoundPower, A.DIR_GEODESIC)
        tr.set_receivers(mr.Origins[:1], 0.5, at, c, c * rho, 1000, mr.SampleCT)
        tr.set_map(mr.Origins, 1.0, True, mr.SampleCT)
        tr.trace(rays=5000, cutoff_m=0.5 * c, seed=4, precision=A.PREC_FP64)
        want = tr.map_histogram(0, scaled=True)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    spl = P.PachMapReceiver.Get_SPL_Map([mr], [500, 0], 8, [0])
    E = want.sum(1)  # all bands
    for i in range(mr.Count):
        arrival = int(mr.Direct_Time[i] * 1000)
        e = E[i][: max(0, 500 - arrival)].sum() * np.sqrt(405.0)
        assert spl[i] == pytest.approx(10 * np.log10(e / 1e-12), rel=1e-12)
    rt.close()


def test_time1pt_map_matches_oracle_and_the_windowed_sums(tracer):
    """Time1Pt (Histogram_1Pt): one sample per receiver and band holding every
    deposit's energy; equal to the oracle within 1e-9 of the largest value,
    the same deposit count, no pressure or direction planes, and at least the
    time sum of the windowed map of the same trace (which drops deposits
    outside [0, SampleCT))."""
    print("[case] map receivers Time1Pt", flush=True)
    got, want, st, cnt = run_pair(tracer, box(), plane_centers(), 1.0, rays=2000, mode=A.MODE_STOCHASTIC,
                                  time1pt=True)
    assert got.shape == (8, 48, 8, 1)
    assert st["deposits"] == cnt["deposits"] and want[0].sum() > 0
    np.testing.assert_allclose(got[0], want[0], rtol=0, atol=1e-9 * np.abs(want[0]).max())
    assert not got[1:].any()
    win, _, st_w, _ = run_pair(tracer, box(), plane_centers(), 1.0, rays=2000, mode=A.MODE_STOCHASTIC)
    tot = win[0].sum(-1)
    assert np.all(got[0][..., 0] >= tot * (1 - 1e-12))
    assert st["deposits"] >= st_w["deposits"]
    tracer.clear_map()
