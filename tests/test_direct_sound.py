"""Deterministic direct sound (Classes_Direct.cs:654-840) and the ETCurve
assembly (Classes_Math.cs:1417-1474), host code checked on the CPU:

* unobstructed receivers get P 10^(-0.1 a d) / (4 pi d^2) per band at d / c;
* a receiver behind a block is not audible (IsOccluded False) while its
  energy is still computed, as the reference does;
* through a transmissive block the coefficients of every transmissive face
  crossed multiply the energy, and the search goes on from each hit point
  still comparing with the whole source-receiver distance (the reference's
  rule): audible only if the next opaque surface lies that far beyond;
* ETCurve adds the direct energy at ceil(d / c * fs), all bands summed for
  Octave 8, onto the ray tracer's scaled histogram.
"""
import numpy as np

from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import pachyderm as P
from pachydermacoustic_universal_amd import rooms


def blocked_room(L=20.0, tau=None):
    box = rooms.shoebox(L, 12, 6, alpha=0.15, scat=0.3, source=(3, 6, 1.5),
                        receivers=[(16, 10.5, 1.5), (14, 6, 1.5), (9, 2, 1.5)], radius=0.5)
    block = rooms._box_tris((10.0, 4.0, 0.0), (12.0, 8.0, 5.0), bottom=False)
    nb = block.shape[0] // 2
    S = 6 + nb
    trans = np.zeros((S, 8))
    if tau is not None:
        trans[6:] = tau
    return rooms.Room(name="blocked", verts=np.concatenate([box.verts, block]),
                      surf=np.concatenate([box.surf, 6 + np.arange(block.shape[0], dtype=np.int32) // 2]),
                      alpha=np.full((S, 8), 0.15), scat=np.full((S, 8), 0.3), trans=trans,
                      source=box.source, receivers=box.receivers, radius=0.5)


def closed_form(room, k, pw, at, c):
    d = np.linalg.norm(room.receivers[k] - room.source)
    return pw * 10 ** (-0.1 * at * d) / (4 * np.pi * d * d), d / c


def test_open_paths_closed_form():
    at, c, _ = A.uniform_medium()
    room = rooms.shoebox()
    pw = np.linspace(0.5, 2.0, 8)
    E, t, aud = A.direct_sound(room, pw, at, c)
    assert aud.all()
    for k in range(3):
        e, tt = closed_form(room, k, pw, at, c)
        np.testing.assert_allclose(E[k], e, rtol=1e-13)
        assert abs(t[k] - tt) < 1e-15


def test_blocked_receiver_not_audible_energy_kept():
    at, c, _ = A.uniform_medium()
    room = blocked_room()
    E, t, aud = A.direct_sound(room, np.ones(8), at, c)
    assert list(aud) == [True, False, True]
    np.testing.assert_allclose(E[1], closed_form(room, 1, np.ones(8), at, c)[0], rtol=1e-13)


def test_transmission_through_the_block():
    at, c, _ = A.uniform_medium()
    tau = np.zeros(8)
    tau[3] = 0.5
    # 20 m room: from the block's far face (x = 12) the wall at x = 20 is 8 m
    # away, less than the 11 m source-receiver distance -> still blocked
    room = blocked_room(20.0, tau)
    E, _, aud = A.direct_sound(room, np.ones(8), at, c)
    e, _ = closed_form(room, 1, np.ones(8), at, c)
    assert not aud[1]
    np.testing.assert_allclose(E[1], e * np.where(np.arange(8) == 3, 0.25, 0.0), rtol=1e-13)
    # 40 m room: the wall is 28 m beyond the far face -> audible, through two faces
    room = blocked_room(40.0, tau)
    E, _, aud = A.direct_sound(room, np.ones(8), at, c)
    assert aud[1]
    np.testing.assert_allclose(E[1], e * np.where(np.arange(8) == 3, 0.25, 0.0), rtol=1e-13)


def test_etcurve_places_the_direct_sound():
    room = rooms.shoebox()
    sc = P.Polygon_Scene.from_room(room)
    src = P.GeodesicSource([120] * 8, room.source)
    rb = P.Receiver_Bank(room.receivers, src, sc, 1000, 500.0, radius=room.radius)
    rng = np.random.default_rng(0)
    rb.planes[0] = rng.uniform(0, 1e-6, rb.planes[0].shape)
    ds = P.Direct_Sound(src, rb, sc, [0, 7])
    ds.Begin()
    at, c, _ = A.uniform_medium()
    for rec in range(3):
        e, tt = closed_form(room, rec, np.asarray(src.SoundPower), at, c)
        k = int(np.ceil(tt * 1000))
        for band in (0, 5):
            H = P.IR_Construction.ETCurve([ds], [None], [rb], 500.0, 1000, band, rec)
            expect = rb.planes[0, rec, band].copy()
            expect[k] += e[band]
            np.testing.assert_allclose(H, expect, rtol=1e-15)
        H8 = P.IR_Construction.ETCurve([ds], [None], [rb], 500.0, 1000, 8, rec)
        expect = rb.planes[0, rec].sum(0)
        expect[k] += e.sum()
        np.testing.assert_allclose(H8, expect, rtol=1e-13)
