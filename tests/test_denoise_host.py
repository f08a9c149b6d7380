"""Host-side arithmetic of the denoiser pipeline (no GPU): descriptors with
the ray count, C80 measured from the direct sound's arrival (Clarity's
startTime, Classes_Math.cs:559-585) against its closed form, and the
training-pair assembly over several input ray counts."""
import numpy as np

from pachydermacoustic_universal_amd import denoise as DN
from pachydermacoustic_universal_amd import rooms


def test_descriptors_take_the_ray_count():
    room = rooms.concert_variant(3)
    d12 = DN.descriptors(room)
    d13 = DN.descriptors(room, rays=50_000)
    assert d12.shape == (64, 12) and d13.shape == (64, 13)
    np.testing.assert_array_equal(d13[:, :12], d12)
    np.testing.assert_allclose(d13[:, 12], np.log10(5e4) - 4.5, rtol=1e-6)


def test_c80_is_measured_from_the_direct_sound():
    # direct sound D at bin 120 (a receiver 41 m away), then an exponential
    # decay e^(-k (n - 120)) from the next bin on
    n, d0, D, k = 3000, 120, 5.0, 0.01
    E = np.zeros((1, 8, n))
    t = np.arange(n)
    E[0, :, d0] = D
    E[0, :, d0 + 1:] = np.exp(-k * (t[d0 + 1:] - d0))
    P = DN.decay_params(E, start_bin=[d0])
    # the start sits 1e-6 of a bin past the direct sound's bin, so Clarity's
    # split bin (80 after the start) gives that fraction to the early sum
    f = 1e-6
    early = D + np.exp(-k * np.arange(1, 80)).sum() + f * np.exp(-80 * k)
    late = np.exp(-k * np.arange(81, n - d0)).sum() + (1 - f) * np.exp(-80 * k)
    np.testing.assert_allclose(P[0, :, 2], 10 * np.log10(early / late), rtol=1e-9)
    # from time 0 the 80 ms window ends before the direct sound arrives
    assert np.all(np.isneginf(DN.decay_params(E)[0, :, 2]))


def test_pairs_over_several_ray_counts():
    def fake(tr, pk, rays, cutoff_s, seed, precision=None, is_order=-1):
        rng = np.random.default_rng(seed)
        R = pk.room.receivers.shape[0]
        return rng.random((R, 8, 3000)) * np.exp(-np.arange(3000) / 300.0)

    real = DN.trace_energy
    DN.trace_energy = fake
    try:
        rl = [rooms.concert_variant(s) for s in range(2)] + [rooms.hall_variant(0)]
        a = DN.make_pairs(None, rl[:2], [20_000, 20_000, 50_000], 200_000, pack=1, ray_desc=True, is_order=0)
        b = DN.make_pairs(None, rl[2:], [20_000], 200_000, pack=1, ray_desc=True, is_order=0, room_ids=[7])
    finally:
        DN.trace_energy = real
    d = DN.merge_pairs(a, b)
    assert d["x"].shape == (2 * 64 * 3 + 16, 8, 3000) and d["desc"].shape[1] == 13
    np.testing.assert_array_equal(np.unique(d["lo"]), [20_000, 50_000])
    np.testing.assert_allclose(d["desc"][:, 12], np.log10(d["lo"]) - 4.5, rtol=1e-6)
    assert set(np.unique(d["room"])) == {0, 1, 7}
    # the target shares the input's normalisation: x(0) = t(0) only when the totals agree
    assert np.all(np.abs(d["x"][:, :, 0] - 2.0) < 1e-6)
