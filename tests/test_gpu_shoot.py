"""Closest-hit queries (Polygon_Scene.shoot) on the GPU against the oracle's
brute force over every triangle (Moller-Trumbore in FP64, t > 1e-10, lower
triangle on equal t), through pach_shoot.

  * FP64 and mixed walks: the same triangle and the same distance, bit for bit,
    for every ray -- from points inside the rooms and from points on their
    surfaces leaving in random directions (the departures of a trace), on the
    2,032-triangle hall, the 50k-triangle concert hall and packed batch rooms
    hundreds of metres from the origin (large FP32 rounding).
  * FP32 walk (the FP32 tracer's): the same triangle except where the ray
    passes within 1e-4 (barycentric) of an edge of the FP64 hit, where the two
    neighbours are equally valid to FP32; distances within 1e-5 relative + 2e-4 m.
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


def random_rays(room, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = room.bbox()
    o = rng.uniform(lo + 0.05 * (hi - lo), hi - 0.05 * (hi - lo), (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def departures(room, sc, o, d, tri, t, seed):
    """Rays leaving the hit points in random directions on the incoming side."""
    rng = np.random.default_rng(seed)
    hit = tri >= 0
    p = o[hit] + d[hit] * t[hit, None]
    v = room.verts[tri[hit]]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    side = -np.sign(np.einsum("ij,ij->i", d[hit], n))[:, None]
    q = rng.normal(size=p.shape)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q = np.where(np.einsum("ij,ij->i", q, n * side)[:, None] < 0, -q, q)
    return p, q


def barycentric_edge_distance(room, tri, o, d, t):
    v = room.verts[tri]
    p = o + d * t[:, None]
    e1, e2, w = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], p - v[:, 0]
    d11, d12, d22 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    w1, w2 = (w * e1).sum(1), (w * e2).sum(1)
    den = d11 * d22 - d12 * d12
    b1 = (d22 * w1 - d12 * w2) / den
    b2 = (d11 * w2 - d12 * w1) / den
    return np.minimum(np.minimum(b1, b2), 1 - b1 - b2)


def packed_batch():
    rs, _ = rooms.batch_rooms(seed=3, n=12)
    return rooms.pack_rooms(rs)


CASES = [("hall", rooms.fan_hall, 100_000), ("concert_hall", rooms.concert_hall, 20_000),
         ("batch12", packed_batch, 30_000)]


@pytest.mark.parametrize("name,make,n", CASES)
def test_fp64_and_mixed_equal_brute_force(tracer, name, make, n):
    room = make()
    room = getattr(room, "room", room)  # packed batch rooms
    tracer.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
    sc = O.Scene(room, grid=(1, 1, 1))
    o, d = random_rays(room, n, 1)
    bt, bd = sc.shoot_brute_many(o, d)
    p, q = departures(room, sc, o, d, bt, bd, 2)
    bt2, bd2 = sc.shoot_brute_many(p, q)
    for prec in (A.PREC_FP64, A.PREC_MIXED):
        print(f"[case] shoot {name} n={n} precision={prec}", flush=True)
        gt, gd, n_ex = tracer.shoot(o, d, prec)
        assert np.array_equal(gt, bt), (prec, np.flatnonzero(gt != bt)[:5])
        assert np.array_equal(gd, bd), (prec, np.abs(gd - bd).max())
        gt2, gd2, n_ex2 = tracer.shoot(p, q, prec)
        assert np.array_equal(gt2, bt2), (prec, np.flatnonzero(gt2 != bt2)[:5])
        assert np.array_equal(gd2, bd2), (prec, np.abs(gd2 - bd2).max())
        if prec == A.PREC_MIXED:
            # the FP32 screen's margins grow with the distance from the origin
            # (the packed rooms reach hundreds of metres), the FP64 fallbacks with them
            print(f"  FP64 fallbacks: {n_ex} of {n}, departures {n_ex2} of {len(p)}", flush=True)
            if name == "hall":
                assert n_ex < 0.01 * n


def test_fp32_walk_against_brute_force(tracer):
    room = rooms.fan_hall()
    tracer.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (32, 32, 32))
    sc = O.Scene(room, grid=(1, 1, 1))
    o, d = random_rays(room, 100_000, 3)
    bt, bd = sc.shoot_brute_many(o, d)
    print("[case] shoot hall FP32", flush=True)
    gt, gd, _ = tracer.shoot(o, d, A.PREC_FP32)
    same = gt == bt
    hit = bt >= 0
    assert np.array_equal(gt < 0, bt < 0)
    diff = hit & ~same
    if diff.any():
        near = barycentric_edge_distance(room, bt[diff], o[diff], d[diff], bd[diff])
        assert near.max() < 1e-4, near.max()
    both = hit & same
    # FP32 distances: a few ulp of coordinates up to ~40 m, more at grazing incidence
    np.testing.assert_allclose(gd[both], bd[both], rtol=1e-5, atol=2e-4)
    print(f"  FP32 triangle differs at an edge for {diff.sum()} of {hit.sum()} hits", flush=True)
