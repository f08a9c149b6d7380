"""Batched rooms: many rooms traced in one launch sequence (pach_rooms_set).

Each room's receivers see only that room's rays, and every source ray keeps
its global index as its Philox stream, so in FP64 each room's slice of the
batch must equal the oracle tracing that room alone from the same first ray
index (1e-5 relative per bin, the deterministic bar).  The FP32 batch is held
to the statistical bar (T30 / EDT within 2 %, C80 within 0.2 dB per band) of
the FP64 batch on the same seeds.
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


def assert_bins(g, o, rtol, what):
    scale = max(np.abs(o).max(), 1e-300)
    diff = np.abs(g - o)
    bad = diff > rtol * np.abs(o) + 1e-12 * scale
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} bins off; worst rel "
                           f"{(diff / np.maximum(np.abs(o), 1e-300))[bad].max():.3e}")


def small_batch(seed, n, rays_lo, rays_hi, max_side=None):
    rl, _ = rooms.batch_rooms(seed=seed, n=8 * n if max_side else n)
    if max_side:
        rl = [r for r in rl if (r.bbox()[1] - r.bbox()[0]).max() <= max_side][:n]
    rng = np.random.default_rng(seed + 100)
    rays = rng.integers(rays_lo, rays_hi + 1, n)
    return rooms.pack_rooms(rl), rays


def test_batch_fp64_matches_oracle_per_room(tracer):
    at, c, rho = A.uniform_medium()
    pk, rays = small_batch(seed=7, n=8, rays_lo=4, rays_hi=12)
    cut = 0.6
    print(f"[case] batch fp64 rooms={len(rays)} rays={list(rays)} cut={cut}", flush=True)
    st = A.trace_batch(tracer, pk, rays, cutoff_s=cut, seed=5, mode=A.MODE_STOCHASTIC,
                       precision=A.PREC_FP64, medium=(at, c, rho))
    G = tracer.planes(scaled=False)
    base = np.concatenate([[0], np.cumsum(rays)])
    r0 = 0
    total = 0
    for i in range(len(rays)):
        m = pk.member(i)
        ora, cnt = O.trace(m, rays=int(rays[i]), cutoff_s=cut, seed=5, mode=A.MODE_STOCHASTIC,
                           medium=O.Medium(c=c, atten=at, rho=rho), threads=8, ray_begin=int(base[i]))
        total += cnt["bounces_broad"] + cnt["bounces_late"] + cnt["bounces_verylate"]
        k = int(pk.rec_count[i])
        for p in range(A.NPLANES):
            assert_bins(G[p, r0:r0 + k], ora[p], 1e-5, f"room {i} plane {p}")
        r0 += k
    assert st["bounces"] == total
    # Receiver_Bank.Scale per room: energy / own ray count, pressure / sqrt
    E = tracer.histogram(0, scaled=True)
    P = tracer.histogram(A.PLANES["pressure"], scaled=True)
    r0 = 0
    for i in range(len(rays)):
        k = int(pk.rec_count[i])
        np.testing.assert_allclose(E[r0:r0 + k], G[0, r0:r0 + k] / rays[i], rtol=1e-14)
        np.testing.assert_allclose(P[r0:r0 + k], G[A.PLANES["pressure"], r0:r0 + k] / np.sqrt(rays[i]), rtol=1e-14)
        r0 += k


def test_batch_rooms_are_isolated(tracer):
    """Dropping every ray of one room leaves the other rooms' bins unchanged: the
    same deposits, up to the order in which FP64 atomics add two deposits into
    one bin (a last-bit difference, so 1e-14 relative)."""
    pk, rays = small_batch(seed=9, n=5, rays_lo=20, rays_hi=40)
    print(f"[case] batch isolation rooms={len(rays)} rays={list(rays)}", flush=True)
    A.trace_batch(tracer, pk, rays, cutoff_s=0.3, seed=2, precision=A.PREC_FP32)
    full = tracer.planes(scaled=False)
    r2 = rays.copy()
    r2[2] = 0
    # emptying room 2 moves the first global ray index (the Philox stream) of
    # rooms 3-4, so only rooms 0-1 keep their streams and must be unchanged
    A.trace_batch(tracer, pk, r2, cutoff_s=0.3, seed=2, precision=A.PREC_FP32)
    part = tracer.planes(scaled=False)
    k0 = int(pk.rec_count[:2].sum())
    np.testing.assert_allclose(part[:, :k0], full[:, :k0], rtol=1e-14, atol=0)
    assert np.array_equal(part[:, :k0] != 0, full[:, :k0] != 0)
    k1 = k0 + int(pk.rec_count[2])
    assert np.all(part[:, k0:k1] == 0)
    assert full[0, k0:k1].sum() > 0


def decay_params(E, fs=1000):
    """T30, EDT, C80 per (receiver, band) from an energy histogram [R, 8, n]."""
    out = np.zeros(E.shape[:2] + (3,))
    for r in range(E.shape[0]):
        for b in range(E.shape[1]):
            sch = A.schroeder_integral(E[r, b])
            out[r, b] = (A.t_x(sch, 30, fs), A.early_decay_time(sch, fs), A.clarity(E[r, b], fs))
    return out


def test_batch_fp32_statistics_match_fp64(tracer):
    pk, rays = small_batch(seed=11, n=4, rays_lo=2000, rays_hi=3000, max_side=20.0)
    cut = 1.0
    print(f"[case] batch fp32 vs fp64 rooms={len(rays)} rays={list(rays)} cut={cut}", flush=True)
    A.trace_batch(tracer, pk, rays, cutoff_s=cut, seed=4, precision=A.PREC_FP64)
    e64 = A.batch_histograms(tracer, pk, 0)
    A.trace_batch(tracer, pk, rays, cutoff_s=cut, seed=4, precision=A.PREC_FP32)
    e32 = A.batch_histograms(tracer, pk, 0)
    for i in range(len(rays)):
        p32, p64 = decay_params(e32[i]), decay_params(e64[i])
        rel = np.abs(p32[..., :2] - p64[..., :2]) / p64[..., :2]
        assert rel.max() < 0.02, (i, rel.max())
        assert np.abs(p32[..., 2] - p64[..., 2]).max() < 0.2, i


def test_batch_256_rooms_at_shape(tracer):
    """The batched workload's shape: all 256 rooms of batch_rooms(seed=3) in one
    launch sequence, each at 1 % of its log-uniform [1e4, 1e6] ray count.
    Every room's receivers record energy, the bank scales each room by its own
    rays, almost no energy leaves the closed rooms, and emptying the last
    room changes no other room's bins."""
    members, base_rays = rooms.batch_rooms(seed=3, n=256)
    pk = rooms.pack_rooms(members)
    rays = np.maximum(1, np.round(np.array(base_rays) * 0.01)).astype(np.int64)
    cut = 1.0
    print(f"[case] batch 256 rooms, {rays.sum()} rays, cut={cut}", flush=True)
    st = A.trace_batch(tracer, pk, rays, cutoff_s=cut, seed=8, precision=A.PREC_FP32)
    assert st["rays_done"] == rays.sum()
    raw = tracer.planes(scaled=False)
    E = A.batch_histograms(tracer, pk, 0)
    r0 = 0
    for i in range(len(rays)):
        k = int(pk.rec_count[i])
        assert raw[0, r0:r0 + k].sum(axis=(1, 2)).min() > 0, f"room {i}: a receiver got nothing"
        np.testing.assert_allclose(E[i], raw[0, r0:r0 + k] / rays[i], rtol=1e-14)
        r0 += k
    assert st["lost"] / st["e_init"] < 1e-4, st["lost"] / st["e_init"]
    r2 = rays.copy()
    r2[-1] = 0
    A.trace_batch(tracer, pk, r2, cutoff_s=cut, seed=8, precision=A.PREC_FP32)
    part = tracer.planes(scaled=False)
    k_last = int(pk.rec_count[-1])
    np.testing.assert_allclose(part[:, :-k_last], raw[:, :-k_last], rtol=1e-12, atol=0)
    assert np.all(part[:, -k_last:] == 0)


def test_batch_cancel_scales_each_room_by_its_completed_rays(tracer):
    """A batch cancelled part-way: the rooms whose rays all completed are
    scaled by their own ray counts, the room in progress by its completed
    rays, and the rooms not reached hold nothing (Receiver_Bank.Scale per
    room with the completed count, TaskAll:768-774)."""
    import time

    members, base_rays = rooms.batch_rooms(seed=3, n=64)
    pk = rooms.pack_rooms(members)
    rays = np.maximum(1, np.round(np.array(base_rays) * 0.01)).astype(np.int64)
    print(f"[case] batch 64 rooms cancelled, {rays.sum()} rays", flush=True)
    at, c, rho = A.uniform_medium()
    room = pk.room
    tracer.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
    tracer.set_source(pk.sources[0], np.ones(8), A.DIR_GEODESIC)
    tracer.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 1000)
    tracer.set_rooms(pk.sources, np.ones((len(rays), 8)), rays, pk.surf_count, pk.rec_count)
    try:
        tracer.trace(rays=0, cutoff_m=1.0 * c, seed=3, precision=A.PREC_FP32, batch_rays=512, async_=True)
        time.sleep(0.3)
        tracer.cancel()
        assert tracer.wait() in (0, -5)
    finally:
        tracer.clear_rooms()
    done = tracer.stats()["rays_done"]
    if done in (0, rays.sum()):
        pytest.skip(f"cancel landed at {done} rays")
    base = np.concatenate([[0], np.cumsum(rays)])
    completed = np.clip(done - base[:-1], 0, rays)
    raw = tracer.planes(scaled=False)
    E = A.batch_histograms(tracer, pk, 0)
    r0 = 0
    for i in range(len(rays)):
        k = int(pk.rec_count[i])
        if completed[i] == 0:
            assert np.all(raw[:, r0:r0 + k] == 0), i
        else:
            np.testing.assert_allclose(E[i], raw[0, r0:r0 + k] / completed[i], rtol=1e-14)
        r0 += k
