"""CPU FP64 restatement of the auralisation-filter path (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
filters (pach_etc_to_filter, pach_receiver_filters).  The product never
imports it.

Restated (paths relative to Pachyderm_Acoustic_Universal/):
  Pach_SP.ETCToFilter                 Pach_SP.cs:1584-1676
  GenerateBandLimitedSincPulse        Pach_SP.cs:1678-1700
  FFT_Convolution_double              Pach_SP.cs:1948-1975
  Medium_Properties.ISO9613_1_attencoef (scalar form)
                                      Classes_Medium_Props.cs:193-202
  Directional_Histogram.Create_Filter(SWL, Rho_C)
                                      Classes_Receivers.cs:1911-1936
  Directional_Histogram.Create_Filter(Rho_C)
                                      Classes_Receivers.cs:1969-1995

Pinned by tests/test_filters_ref.py: the FFT convolution against a direct
convolution, the carrier's L1 norm, an impulse envelope giving back the
scaled carrier, linearity, the ISO 9613-1 attenuation formula against the
standard's table at 20 C / 50 % (with kPa), and the channel rules of
Create_Filter on hand-built histograms.
"""
from __future__ import annotations

import math

import numpy as np

CARRIER_LEN = 16384
RT2 = math.sqrt(2)


def iso9613_scalar(freq, tk, pa, hr):
    """ISO9613_1_attencoef (scalar overload), C# integer exponents kept:
    -1/2 = 0, -1/3 = 0, -5/2 = -2, 1/2 = 0."""
    h = hr * 10 ** (-6.8346 * (273.16 / tk) ** 1.261 + 4.6151) / (pa / 101.325)
    fro = (pa / 101.325) * (24 + 4.04 * 10 ** 4 * h * ((0.02 + h) / (0.391 + h)))
    frn = (pa / 101.325) * (tk / 293.15) ** 0 * (9 + 280 * h * math.exp(-4.170 * ((tk / 293.15) ** 0 - 1)))
    z = 0.1068 * math.exp(-3352 / tk) * (frn + freq * freq / frn) ** -1
    y = (tk / 293.15) ** -2 * (0.01275 * math.exp(-2239.1 / tk) * (fro + freq * freq / fro) ** -1 + z)
    return 8.686 * freq * freq * ((1.84 * 10 ** -11 * (pa / 101.325) ** -1 * (tk / 293.15) ** 0) + y)


def carrier(length, fs, fl, fu):
    """GenerateBandLimitedSincPulse: sinc(bw t) cos(2 pi fc t) about the
    centre sample, scaled by 2 / (L1 norm x length)."""
    fc = (fl + fu) / 2.0
    bw = fu - fl
    t0 = (length - 1) / 2.0
    t = (np.arange(length) - t0) / fs
    x = np.pi * bw * t
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x)) if bw != 0 else np.ones(length)
    c = sinc * np.cos(2 * np.pi * fc * t)
    norm = 2 / (np.abs(c).sum() * length)
    return c * norm if norm > 0 else c


def carriers(fs):
    """The 8 octave-band carriers (62.5 Hz .. 8 kHz) and the 16 kHz one."""
    out = []
    for b in range(8):
        fc = 62.5 * 2 ** b
        out.append(carrier(CARRIER_LEN, fs, fc / RT2, fc * RT2))
    out.append(carrier(CARRIER_LEN, fs, 16000.0 / RT2, 16000.0 * RT2))
    return np.array(out)


def fft_convolution(sig, filt):
    """FFT_Convolution_double: zero-pad to the next power of two, multiply
    spectra, inverse, keep len(sig) + len(filt) - 1 samples."""
    n = len(sig) + len(filt) - 1
    W = 1 << int(math.ceil(math.log2(n)))
    return np.fft.irfft(np.fft.rfft(sig, W) * np.fft.rfft(filt, W), W)[:n]


def band16k_factor(T, fs):
    """Per-sample pressure factor of the 16 kHz band made from the 8 kHz one:
    the extra air attenuation over the distance 343 m/s x t (ISO 9613-1 at
    20 C, 50 %, with the pressure passed in Pa as the reference does)."""
    tk, pa, hr = 20 + 273.15, 101.325 * 1000, 50.0
    extra = iso9613_scalar(16000.0, tk, pa, hr) - iso9613_scalar(8000.0, tk, pa, hr)
    d = np.arange(T) * (1.0 / fs) * 343.0
    return 10 ** (-extra * d / 20.0)


def etc_to_filter(prms, swl, fs=44100):
    """Pach_SP.ETCToFilter of one set of 8 band envelopes prms [8][T]."""
    prms = np.asarray(prms, dtype=np.float64)
    T = prms.shape[1]
    p_mod = 10 ** ((120 - np.asarray(swl, dtype=np.float64)) / 20.0)
    car = carriers(fs)
    out_len = T + CARRIER_LEN - 1
    ir = np.zeros(out_len)
    for b in range(8):
        ir += fft_convolution(np.abs(prms[b]) * p_mod[b], car[b])[:out_len]
    env16 = np.abs(prms[7]) * p_mod[7] * band16k_factor(T, fs)
    ir += fft_convolution(env16, car[8])[:out_len]
    return ir


def direction_channels(E, P, pos, neg):
    """Directional_Histogram.Create_Filter(SWL, Rho_C): channel 0 the
    pressure envelopes, then per axis d the envelopes P neg_d / E and
    P pos_d / E (0 where E = 0).  E, P [8][T]; pos, neg [3][8][T]."""
    with np.errstate(invalid="ignore", divide="ignore"):
        chans = [P]
        for d in range(3):
            chans.append(np.where(E == 0, 0.0, P * neg[d] / np.where(E == 0, 1.0, E)))
            chans.append(np.where(E == 0, 0.0, P * pos[d] / np.where(E == 0, 1.0, E)))
    return np.array(chans)


def receiver_filters(E, P, pos, neg, swl, fs):
    """Create_Filter(SWL, Rho_C) of one receiver: [7][T + 16383 - 4096]."""
    return np.array([etc_to_filter(ch, swl, fs)[:-4096] for ch in direction_channels(E, P, pos, neg)])
