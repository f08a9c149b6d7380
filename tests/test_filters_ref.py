"""The auralisation-filter oracle (oracle/filters_ref.py: Pach_SP.ETCToFilter
and Directional_Histogram.Create_Filter) pinned on the CPU:

* FFT_Convolution_double equals a direct convolution (numpy.convolve);
* each carrier's L1 norm is 2 / length (GenerateBandLimitedSincPulse's
  normalisation), and its spectrum peaks inside its band;
* an impulse envelope in band b at sample t0 gives back the carrier of band
  b, scaled by 10^((120 - SWL) / 20) and delayed by t0 (band 7 adds the
  16 kHz carrier scaled by the extra air attenuation at 343 t0 / fs metres);
* the filter is linear in the envelopes' magnitudes;
* the ISO 9613-1 formula, evaluated with the pressure in kPa, reproduces the
  standard's table at 20 C / 50 % (4.66, 9.86, 29.7, 105 dB/km at 1, 2, 4,
  8 kHz); the reference passes pascals, which the 16 kHz band keeps;
* Create_Filter's channels: pressure, then per axis P neg / E and P pos / E,
  zero where E = 0.
"""
import numpy as np
import pytest

from oracle import filters_ref as F


def test_fft_convolution_is_a_convolution():
    rng = np.random.default_rng(1)
    for n, m in ((1, 5), (37, 64), (300, 1000), (1000, 16384)):
        a, b = rng.normal(size=n), rng.normal(size=m)
        got = F.fft_convolution(a, b)
        want = np.convolve(a, b)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_carrier_normalisation_and_band():
    fs = 44100
    car = F.carriers(fs)
    assert car.shape == (9, F.CARRIER_LEN)
    for b in range(9):
        assert np.abs(car[b]).sum() == pytest.approx(2 / F.CARRIER_LEN, rel=1e-12)
        spec = np.abs(np.fft.rfft(car[b]))
        f = np.fft.rfftfreq(F.CARRIER_LEN, 1 / fs)[np.argmax(spec)]
        fc = 16000.0 if b == 8 else 62.5 * 2 ** b
        assert fc / np.sqrt(2) <= f <= fc * np.sqrt(2)


@pytest.mark.parametrize("band", [0, 3, 7])
def test_impulse_envelope_returns_the_carrier(band):
    fs, T, t0 = 1000, 200, 37
    swl = np.array([120, 110, 100, 95, 90, 85, 80, 75.0])
    prms = np.zeros((8, T))
    prms[band, t0] = -2.5  # magnitudes are taken
    ir = F.etc_to_filter(prms, swl, fs)
    car = F.carriers(fs)
    want = np.zeros(T + F.CARRIER_LEN - 1)
    want[t0:t0 + F.CARRIER_LEN] += 2.5 * 10 ** ((120 - swl[band]) / 20) * car[band]
    if band == 7:
        extra = F.iso9613_scalar(16000.0, 293.15, 101325.0, 50) - F.iso9613_scalar(8000.0, 293.15, 101325.0, 50)
        want[t0:t0 + F.CARRIER_LEN] += 2.5 * 10 ** ((120 - swl[7]) / 20) * 10 ** (-extra * t0 / fs * 343 / 20) * car[8]
    np.testing.assert_allclose(ir, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_linear_in_envelope_magnitudes():
    rng = np.random.default_rng(2)
    a, b = rng.random((8, 150)), rng.random((8, 150))
    swl = np.full(8, 120.0)
    got = F.etc_to_filter(2 * a + 3 * b, swl, 1000)
    want = 2 * F.etc_to_filter(a, swl, 1000) + 3 * F.etc_to_filter(b, swl, 1000)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_iso9613_table_values_at_20c_50pct():
    table = {1000: 4.66, 2000: 9.86, 4000: 29.7, 8000: 105.0}  # dB/km
    for f, v in table.items():
        assert F.iso9613_scalar(f, 293.15, 101.325, 50) * 1000 == pytest.approx(v, rel=0.01)


def test_create_filter_channels():
    E = np.array([[0.0, 2.0, 4.0]] * 8)
    P = np.array([[1.0, 3.0, 8.0]] * 8)
    pos = np.stack([E * 0.25, E * 0.5, E * 0.0])
    neg = np.stack([E * 0.75, E * 0.0, E * 1.0])
    ch = F.direction_channels(E, P, pos, neg)
    assert ch.shape == (7, 8, 3)
    np.testing.assert_array_equal(ch[0], P)
    for d, (fn, fp) in enumerate(((0.75, 0.25), (0.0, 0.5), (1.0, 0.0))):
        np.testing.assert_allclose(ch[1 + 2 * d][:, 1:], P[:, 1:] * fn)
        np.testing.assert_allclose(ch[2 + 2 * d][:, 1:], P[:, 1:] * fp)
        assert (ch[1 + 2 * d][:, 0] == 0).all() and (ch[2 + 2 * d][:, 0] == 0).all()
    f = F.receiver_filters(E, P, pos, neg, np.full(8, 120.0), 1000)
    assert f.shape == (7, 3 + F.CARRIER_LEN - 1 - 4096)
