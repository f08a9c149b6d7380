"""pach_autotune: the tuner times sample traces over voxel resolution, launch
shape and loop thresholds, applies the fastest settings, and the tracer's
results under those settings still match the oracle bin for bin (FP64)."""
import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu


def test_autotune_hall_then_parity():
    at, c, rho = A.uniform_medium()
    room = rooms.fan_hall()
    tr = A.Tracer(0)
    try:
        A.trace_room(tr, room, rays=16, cutoff_s=0.8, seed=1)  # upload scene / receivers (800 bins)
        print("[case] autotune hall sample=1024 cut=1.0", flush=True)
        t = tr.autotune(rays=100000, cutoff_m=1.0 * c, sample_rays=1024, seed=3)
        print("[tuning]", t, flush=True)
        assert t["trials"] >= 10
        assert t["rate_best"] >= t["rate_start"] > 0
        assert t["block"] in (256, 512) and t["occupancy"] in (-1, 0, 4, 5) and t["shade_queue"] in (0, 1)
        assert 1 <= t["walk_steps"] <= 256 and 1 <= t["shade_min"] <= 64 and 1 <= t["refill_min"] <= 64
        assert all(1 <= g <= 1024 for g in t["grid"])
        # trace under the tuned settings (same context, no re-upload) against the oracle
        tr.trace(rays=10, cutoff_m=0.8 * c, seed=13, precision=A.PREC_FP64)
        g = tr.planes(scaled=False)
        ora, cnt = O.trace(room, rays=10, cutoff_s=0.8, seed=13, medium=O.Medium(c=c, atten=at, rho=rho),
                           threads=8)
        for k in range(A.NPLANES):
            o = ora[k]
            bad = np.abs(g[k] - o) > 1e-5 * np.abs(o) + 1e-12 * max(np.abs(o).max(), 1e-300)
            assert not bad.any(), f"plane {k}: {bad.sum()} bins off"
    finally:
        tr.close()


def test_autotune_batch_keeps_room_rays():
    rl, _ = rooms.batch_rooms(seed=13, n=12)
    pk = rooms.pack_rooms(rl)
    rays = np.random.default_rng(1).integers(200, 2000, len(rl))
    tr = A.Tracer(0)
    try:
        at, c, rho = A.uniform_medium()
        A.trace_batch(tr, pk, rays, cutoff_s=0.3, seed=2)
        ref = tr.planes(scaled=False)
        room = pk.room
        tr.set_rooms(pk.sources, np.ones((len(rl), 8)), rays, pk.surf_count, pk.rec_count)
        print("[case] autotune batch rooms=12 sample=2000", flush=True)
        t = tr.autotune(rays=0, cutoff_m=0.3 * c, sample_rays=2000, seed=2)
        print("[tuning]", t, flush=True)
        assert t["rate_best"] >= t["rate_start"] > 0
        # the rooms' own ray counts are back in force after tuning
        tr.trace(rays=0, cutoff_m=0.3 * c, seed=2)
        st = tr.stats()
        assert st["rays_done"] == rays.sum()
        got = tr.planes(scaled=False)
        # FP32 results depend on the grid only through rounding at voxel
        # faces; the energy per room agrees to well within 1 %
        r0 = 0
        for k in pk.rec_count:
            a, b = got[0, r0:r0 + k].sum(), ref[0, r0:r0 + k].sum()
            assert abs(a - b) <= 0.01 * b, (a, b)
            r0 += int(k)
    finally:
        tr.close()


def test_autotune_default_sample_is_sized_from_bounces():
    """With no sample size given, the tuner sizes its trials from the bounces
    per source ray of a first trial (here the whole, small ray count) and
    leaves the rooms' own ray counts in force afterwards."""
    rl, _ = rooms.batch_rooms(seed=17, n=6)
    pk = rooms.pack_rooms(rl)
    rays = np.random.default_rng(2).integers(100, 400, len(rl))
    tr = A.Tracer(0)
    try:
        at, c, rho = A.uniform_medium()
        A.trace_batch(tr, pk, rays, cutoff_s=0.2, seed=2)
        tr.set_rooms(pk.sources, np.ones((len(rl), 8)), rays, pk.surf_count, pk.rec_count)
        print("[case] autotune batch rooms=6 default sample", flush=True)
        t = tr.autotune(rays=0, cutoff_m=0.2 * c, sample_rays=0, seed=2)
        print("[tuning]", t, flush=True)
        assert t["trials"] >= 10 and t["rate_best"] >= t["rate_start"] > 0
        tr.trace(rays=0, cutoff_m=0.2 * c, seed=2)
        assert tr.stats()["rays_done"] == rays.sum()
    finally:
        tr.close()
