// filters.h -- launcher of the auralisation-filter convolution (filters.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace pach {

constexpr int kCarrierLen = 16384;  // GenerateBandLimitedSincPulse length (Pach_SP.cs:1586)
constexpr int kFiltBands = 9;       // 8 octave bands + the 16 kHz band
constexpr int kFiltCh = 8;          // channels per channel block

// out[ch][i] = sum over bands b and samples k of env[b][ch][k] car[b][i - k],
// i in [0, T + kCarrierLen - 1).  env is laid out [b][ch / 8][k][ch % 8]
// (n_chb blocks of 8 channels, zero beyond n_ch); car is [b][kCarrierLen];
// part holds kFiltBands x n_ch x (T + kCarrierLen - 1) band sums.
hipError_t launch_etc_conv(hipStream_t st, const double *env, const double *car, int T, int n_chb, int n_ch,
                           double *part, double *out);

// The envelopes of Directional_Histogram.Create_Filter for R receivers (7
// channels each) from the trace histograms hist (planes plane_stride apart,
// [R][8][T] each): scale[r] = 1 / (source rays of receiver r's room),
// p_mod [8], f16 [T] the 16 kHz band's factors; env as launch_etc_conv reads it.
hipError_t launch_receiver_env(hipStream_t st, const double *hist, size_t plane_stride, int R, int T,
                               const double *scale, const double *p_mod, const double *f16, int n_chb, double *env);

// The envelopes of pach_etc_to_filter from a device array prms [n_ch][8][T]
// (every element of env written, padding channels zero).
hipError_t launch_env_from_prms(hipStream_t st, const double *prms, int n_ch, int T, const double *p_mod,
                                const double *f16, int n_chb, double *env);

// Allocation-free filters: a fixed device array of kNoAllocScratch bytes per
// GPU, defined in filters.hip; noalloc_scratch returns its device address on
// `device` (looked up once per device with that device current).
constexpr size_t kNoAllocScratch = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch(int device, unsigned char **ptr);

}  // namespace pach
