// filters.hip -- auralisation filters: the band-carrier convolution of
// Pach_SP.ETCToFilter (Pach_SP.cs:1584-1676) for many channels at once.
//
// The reference convolves each octave band's pressure envelope with a
// 16,384-tap band-limited carrier through a zero-padded FFT
// (FFT_Convolution_double, Pach_SP.cs:1948-1975) and sums the bands.  Here the
// nine band convolutions of eight channels run in one FP64 direct-convolution
// pass per output tile: every carrier tap meets every envelope sample once
// (2 T L flops per band and channel, at the FP64 rate; no transform round-off).
//
// Work layout: a workgroup owns kTile consecutive output samples of one block
// of 8 channels in one band (the nine bands' partial sums go to separate
// planes and a second pass adds them in band order, so the grid is nine
// times wider and the sum deterministic); each thread keeps kOut consecutive
// outputs x 8 channels in registers.  For a chunk of kKC envelope samples the carrier window the tile
// needs is staged in LDS; a thread walks the chunk with a 4-value sliding
// window of carrier taps (one LDS read per step), and the 8 channels'
// envelope values of the step, the same for every lane, come through scalar
// loads.
#include "filters.h"

#include <algorithm>
#include <mutex>

namespace pach {
namespace {

constexpr int kThreads = 128;
constexpr int kOut = 4;                 // consecutive outputs per thread
constexpr int kTile = kThreads * kOut;  // outputs per workgroup
constexpr int kKC = 256;                // envelope samples per staged chunk
constexpr int kPad = 4;                 // LDS slack below the window (read, never used)

__global__ void __launch_bounds__(kThreads) k_etc_conv(const double *__restrict__ env,
                                                       const double *__restrict__ car, int T, int n_chb, int n_ch,
                                                       double *__restrict__ part)
{
    __shared__ double cs[kPad + kTile + kKC];
    const int out_len = T + kCarrierLen - 1;
    const int i0 = blockIdx.x * kTile;
    const int chb = blockIdx.y;
    const int base = kOut * threadIdx.x;  // local index of this thread's first output
    double acc[kOut][kFiltCh];
#pragma unroll
    for (int r = 0; r < kOut; ++r)
#pragma unroll
        for (int c = 0; c < kFiltCh; ++c) acc[r][c] = 0.0;
    // envelope samples that reach outputs [i0, i0 + kTile)
    const int k_lo = max(0, i0 - kCarrierLen + 1);
    const int k_hi = min(T, i0 + kTile);
    {
        const int b = blockIdx.z;
        const double *cb = car + (size_t)b * kCarrierLen;
        const double *eb = env + ((size_t)b * n_chb + chb) * (size_t)T * kFiltCh;
        for (int k0 = k_lo; k0 < k_hi; k0 += kKC) {
            const int kc = min(kKC, k_hi - k0);
            // cs[kPad + m] = carrier tap i0 - k0 - (kKC - 1) + m (zero outside the carrier)
            __syncthreads();
            for (int m = threadIdx.x; m < kTile + kKC; m += kThreads) {
                const int j = i0 - k0 - (kKC - 1) + m;
                cs[kPad + m] = (j >= 0 && j < kCarrierLen) ? cb[j] : 0.0;
            }
            __syncthreads();
            // output base + r at envelope sample k0 + kk uses tap m = base + r + kKC - 1 - kk
            const double *wp = cs + kPad + base + kKC - 1;
            double w[kOut];
#pragma unroll
            for (int r = 0; r < kOut; ++r) w[r] = wp[r];
            const double *e = eb + (size_t)k0 * kFiltCh;
            auto step = [&](int kk) {
                double ev[kFiltCh];
#pragma unroll
                for (int c = 0; c < kFiltCh; ++c) ev[c] = e[(size_t)kk * kFiltCh + c];
#pragma unroll
                for (int r = 0; r < kOut; ++r)
#pragma unroll
                    for (int c = 0; c < kFiltCh; ++c) acc[r][c] = fma(ev[c], w[r], acc[r][c]);
                // slide the window one tap down (after the chunk's last step
                // this reads at most one slot below the window: the slack)
#pragma unroll
                for (int r = kOut - 1; r > 0; --r) w[r] = w[r - 1];
                w[0] = wp[-1 - kk];
            };
            int kk = 0;
            for (; kk + 4 <= kc; kk += 4) {
                step(kk);
                step(kk + 1);
                step(kk + 2);
                step(kk + 3);
            }
            for (; kk < kc; ++kk) step(kk);
        }
    }
#pragma unroll
    for (int c = 0; c < kFiltCh; ++c) {
        const int ch = chb * kFiltCh + c;
        if (ch >= n_ch) break;
#pragma unroll
        for (int r = 0; r < kOut; ++r) {
            const int i = i0 + base + r;
            if (i < out_len) part[((size_t)blockIdx.z * n_ch + ch) * out_len + i] = acc[r][c];
        }
    }
}

// out = the nine band planes added in band order (bands 0-7, then 16 kHz).
__global__ void __launch_bounds__(256) k_band_sum(const double *__restrict__ part, size_t n, double *__restrict__ out)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double v = part[i];
#pragma unroll
        for (int b = 1; b < kFiltBands; ++b) v += part[(size_t)b * n + i];
        out[i] = v;
    }
}

// Directional_Histogram.Create_Filter's envelopes straight from the trace's
// histograms: channel 7 r + k of receiver r is its pressure (k = 0) or, for
// axis d, P neg_d / E (k = 2d + 1) and P pos_d / E (k = 2d + 2), 0 where E = 0,
// with Receiver_Bank.Scale's factors (1 / n; 1 / sqrt(n) for pressure);
// |value| x p_mod[band], band 8 from band 7 with f16[t].
__global__ void __launch_bounds__(256) k_receiver_env(const double *__restrict__ hist, size_t plane_stride, int R,
                                                      int T, const double *__restrict__ scale,
                                                      const double *__restrict__ p_mod,
                                                      const double *__restrict__ f16, int n_chb, double *__restrict__ env)
{
    const size_t n = (size_t)R * 7 * 8 * T;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int t = (int)(i % T);
        const int b = (int)((i / T) % 8);
        const int ch = (int)(i / ((size_t)8 * T));
        const int r = ch / 7, k = ch % 7;
        const size_t at = ((size_t)r * 8 + b) * T + t;
        const double sE = scale[r];
        const double P = hist[plane_stride + at] * sqrt(sE);
        double v = P;
        if (k > 0) {
            const int d = (k - 1) / 2;
            const double E = hist[at] * sE;
            const double D = hist[(size_t)(k % 2 == 1 ? 5 + d : 2 + d) * plane_stride + at] * sE;
            v = E == 0 ? 0.0 : P * D / E;
        }
        const double a = fabs(v) * p_mod[b];
        const int cb = ch / kFiltCh, cl = ch % kFiltCh;
        env[(((size_t)b * n_chb + cb) * T + t) * kFiltCh + cl] = a;
        if (b == 7) env[(((size_t)8 * n_chb + cb) * T + t) * kFiltCh + cl] = a * f16[t];
    }
}

// Envelopes of pach_etc_to_filter in launch_etc_conv's layout from the
// caller's device array prms [n_ch][8][T]: |prms| x p_mod[band], the 16 kHz
// band from band 7 with f16[t], zero for the padding channels of the last
// block of 8 (every element is written, so nothing needs clearing first).
__global__ void __launch_bounds__(256) k_env_from_prms(const double *__restrict__ prms, int n_ch, int T,
                                                       const double *__restrict__ p_mod,
                                                       const double *__restrict__ f16, int n_chb,
                                                       double *__restrict__ env)
{
    const size_t n = (size_t)kFiltBands * n_chb * kFiltCh * T;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int cl = (int)(i % kFiltCh);
        const int t = (int)((i / kFiltCh) % T);
        const size_t rest = i / ((size_t)kFiltCh * T);
        const int cb = (int)(rest % n_chb);
        const int b = (int)(rest / n_chb);
        const int ch = cb * kFiltCh + cl;
        double v = 0.0;
        if (ch < n_ch) {
            const int bs = b < 8 ? b : 7;
            v = fabs(prms[((size_t)ch * 8 + bs) * T + t]) * p_mod[bs];
            if (b == 8) v = v * f16[t];
        }
        env[i] = v;
    }
}

// Work buffers of the allocation-free entry point (pach_etc_to_filter_noalloc):
// one fixed array per GPU for the life of the process.  Host code must not
// pass `scratch` itself to a kernel (in host code it names a host-side
// stand-in); noalloc_scratch() returns its device address.
__device__ __align__(256) unsigned char scratch[kNoAllocScratch];

}  // namespace

hipError_t noalloc_scratch(int device, unsigned char **ptr)
{
    static std::mutex mu;
    static unsigned char *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if ((e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = (unsigned char *)p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

hipError_t launch_env_from_prms(hipStream_t st, const double *prms, int n_ch, int T, const double *p_mod,
                                const double *f16, int n_chb, double *env)
{
    const size_t n = (size_t)kFiltBands * n_chb * kFiltCh * T;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_env_from_prms, dim3(blocks), dim3(256), 0, st, prms, n_ch, T, p_mod, f16, n_chb, env);
    return hipGetLastError();
}

hipError_t launch_etc_conv(hipStream_t st, const double *env, const double *car, int T, int n_chb, int n_ch,
                           double *part, double *out)
{
    if (T <= 0 || n_chb <= 0) return hipErrorInvalidValue;
    const int out_len = T + kCarrierLen - 1;
    const dim3 grid((out_len + kTile - 1) / kTile, n_chb, kFiltBands);
    hipLaunchKernelGGL(k_etc_conv, grid, dim3(kThreads), 0, st, env, car, T, n_chb, n_ch, part);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const size_t n = (size_t)n_ch * out_len;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_band_sum, dim3(blocks), dim3(256), 0, st, part, n, out);
    return hipGetLastError();
}

hipError_t launch_receiver_env(hipStream_t st, const double *hist, size_t plane_stride, int R, int T,
                               const double *scale, const double *p_mod, const double *f16, int n_chb, double *env)
{
    const size_t n = (size_t)R * 7 * 8 * T;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_receiver_env, dim3(blocks), dim3(256), 0, st, hist, plane_stride, R, T, scale, p_mod, f16,
                       n_chb, env);
    return hipGetLastError();
}

}  // namespace pach
