"""The reference's SplitRayTracer call sequence, traced on the GPU and checked
against the FP64 oracle (FP64 geometry: bins within 1e-5 relative)."""
import time

import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import (AcousticalMath, Basic_Material, GeodesicSource,
                                             Lambert_Scattering, Polygon_Scene, Receiver_Bank,
                                             SplitRayTracer, rooms)

pytestmark = pytest.mark.gpu


def shoebox_model():
    """The 10 x 8 x 4 m box as one object of six quad polygons."""
    x1, y1, z1 = 10.0, 8.0, 4.0
    P = np.array([[0, 0, 0], [x1, 0, 0], [x1, y1, 0], [0, y1, 0],
                  [0, 0, z1], [x1, 0, z1], [x1, y1, z1], [0, y1, z1]])
    faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [3, 2, 6, 7], [0, 3, 7, 4], [1, 5, 6, 2]]
    return [[P[f] for f in faces]]


def run(rays, seed, cancel_after=None, batch=0):
    scene = Polygon_Scene(20, 50, 1013.25, 0, False, True)
    scene.Construct(shoebox_model(), [Basic_Material([0.1] * 8)], [Lambert_Scattering([0.2] * 8)],
                    [np.zeros(8)])
    scene.partition()
    src = GeodesicSource([120] * 8, (2, 3, 1.5), 0, False)
    recs = rooms.shoebox().receivers
    rb = Receiver_Bank(recs, src, scene, 1000, 600.0, radius=0.5)
    rt = SplitRayTracer(src, rb, scene, 0.6 * scene.Sound_speed(0), [0, 7], -1, rays, None,
                        precision="fp64", seed=seed, batch_rays=batch)
    rt.Begin()
    if cancel_after is not None:
        time.sleep(cancel_after)
        rt.Conclude_Simulation()
    while rt.ThreadState() == "Running":
        time.sleep(0.01)
    rt.Combine_ThreadLocal_Results()
    return rt, scene


def test_split_ray_tracer_call_sequence_matches_oracle():
    rt, scene = run(8, seed=21)
    rb = rt.GetReceiver
    assert rt.Sim_Type() == "Stochastic Ray Tracing"
    ora, cnt = O.trace(rooms.shoebox(), rays=8, cutoff_s=0.6, seed=21, threads=8,
                       medium=O.Medium(c=scene.c, atten=scene.atten, rho=scene.rho))
    ref = O.scale(ora, 8)
    for r in range(3):
        for o in range(8):
            g = rb.GetEnergyHistogram(o, 0, r)
            np.testing.assert_allclose(g, ref[0, r, o], rtol=1e-5, atol=1e-12 * ref[0].max())
    np.testing.assert_allclose(rb.GetEnergyHistogram(8, 5, 1)[5:], ref[0, 1].sum(0), rtol=1e-5,
                               atol=1e-12 * ref[0].max())
    assert abs(rt.PercentLost - 100 * cnt["lost"] / cnt["e_init"]) < 1e-9
    etc = rb.GetEnergyHistogram(3, 0, 0)
    sch = AcousticalMath.Schroeder_Integral(etc)
    assert 0.2 < AcousticalMath.T_X(sch, 30, 1000) < 3.0
    rt.close()


def test_cancel_keeps_completed_rays_only():
    rt, scene = run(40000, seed=5, cancel_after=0.05, batch=16)
    st = rt.Stats
    k = st["rays_done"]
    assert 0 <= k < 40000
    if k == 0:
        pytest.skip("cancelled before the first batch")
    ora, _ = O.trace(rooms.shoebox(), rays=k, cutoff_s=0.6, seed=5, threads=8,
                     medium=O.Medium(c=scene.c, atten=scene.atten, rho=scene.rho))
    ref = O.scale(ora, k)[0]
    np.testing.assert_allclose(rt.GetReceiver.planes[0], ref, rtol=1e-5, atol=1e-12 * ref.max())
    rt.close()


def test_progress_counts_completed_source_rays_while_tracing():
    """pach_poll's rays_done advances while the trace runs, as the reference
    advances its completed-ray count per finished source ray (TaskAll:274-312):
    it never falls, stays below the ray count until the trace ends, takes many
    values on the way, and ends at the ray count."""
    from pachydermacoustic_universal_amd import _abi as A

    room = rooms.fan_hall()
    at, c, rho = A.uniform_medium()
    tr = A.Tracer(0)
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (32, 32, 32))
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 1500)
    n = 300_000
    print(f"[case] async hall trace n={n} batch 4096", flush=True)
    tr.trace(rays=n, cutoff_m=1.5 * c, seed=3, precision=A.PREC_FP32, batch_rays=4096, async_=True)
    seen = []
    while True:
        p = tr.poll()
        seen.append((p["rays_done"], p["running"]))
        if not p["running"]:
            break
        time.sleep(0.002)
    assert tr.wait() == 0
    done = [d for d, _ in seen]
    assert all(b >= a for a, b in zip(done, done[1:])), done
    mid = sorted({d for d, running in seen if running and 0 < d < n})
    print(f"  {len(seen)} polls, {len(mid)} distinct intermediate counts", flush=True)
    assert len(mid) >= 5, seen[:20]
    assert tr.poll()["rays_done"] == n and tr.stats()["rays_done"] == n
    tr.close()


def test_receiver_bank_write_back_scales_every_plane():
    """Combine_ThreadLocal_Results writes all 11 planes back into the bank and
    scales them as Receiver_Bank.Scale does: Energy, Dir_Rec_Pos / Neg and
    Reflection_Analysis_Data / N, Pressure / sqrt(N) (Classes_Receivers.cs:
    1774-1786, :1088-1102), N the completed source rays; the accessors read
    those planes, and every plane equals the oracle's."""
    rt, scene = run(8, seed=21)
    rb = rt.GetReceiver
    n = rt.Stats["rays_done"]
    assert n == 8
    raw = rt._tr.planes(scaled=False)
    for k in range(11):
        want = raw[k] / (np.sqrt(n) if k == 1 else n)
        np.testing.assert_array_equal(rb.planes[k], want)
    ora, _ = O.trace(rooms.shoebox(), rays=8, cutoff_s=0.6, seed=21, threads=8,
                     medium=O.Medium(c=scene.c, atten=scene.atten, rho=scene.rho))
    ref = O.scale(ora, 8)
    for k in range(11):
        np.testing.assert_allclose(rb.planes[k], ref[k], rtol=1e-5, atol=1e-12 * np.abs(ref[k]).max())
        if k in (0, 1, 2, 5, 8, 9):
            assert np.abs(ref[k]).max() > 0, k
    for r in range(rb.Count):
        for o in range(8):
            np.testing.assert_array_equal(rb.GetPressureHistogram(o, r), rb.planes[1, r, o])
            np.testing.assert_array_equal(rb.Directions_Pos(o, r), rb.planes[2:5, r, o])
            np.testing.assert_array_equal(rb.Directions_Neg(o, r), rb.planes[5:8, r, o])
        np.testing.assert_array_equal(rb.Reflection_Analysis_Data(r), rb.planes[8:11, r])
    rt.close()


def test_pac1_file_of_a_gpu_trace(tmp_path):
    """Write_Pac1 (Classes_IO.cs:57-261) of a GPU trace: the direct sound and
    the ray-traced receiver bank go to a .pac1 file whose ray-traced section
    reads back as the bank's scaled planes, and whose energy plane equals the
    oracle's trace of the same rays (bins within 1e-5 relative)."""
    from pachydermacoustic_universal_amd import Direct_Sound, pac1

    rt, scene = run(8, seed=21)
    rb = rt.GetReceiver
    ds = Direct_Sound(rt.Source, rb, scene, [0, 7])
    ds.Begin()
    f = tmp_path / "gpu_trace.pac1"
    pac1.write_pac1(f, ds, rb, date="12/04/2026 10:00:00 AM")
    d = pac1.read_pac1(f)
    sec = d["ray_traced"][0]
    assert sec["sample_ct"] == rb.SampleCT
    np.testing.assert_array_equal(sec["planes"], rb.planes[:8])
    ora, _ = O.trace(rooms.shoebox(), rays=8, cutoff_s=0.6, seed=21, threads=8,
                     medium=O.Medium(c=scene.c, atten=scene.atten, rho=scene.rho))
    ref = O.scale(ora, 8)
    np.testing.assert_allclose(sec["planes"][0], ref[0], rtol=1e-5, atol=1e-12 * ref[0].max())
    assert len(d["direct"]) == 1 and d["direct"][0]["receivers"][0]["time"] == ds.Time(0)
    rt.close()
