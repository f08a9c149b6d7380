// FP64-geometry instantiation of the tracer kernels, compiled with
// -ffp-contract=off so that each operation rounds as the FP64 oracle's does.
#include "trace_impl.h"
namespace pach {
PACH_INSTANTIATE(double)
}

namespace pach {

// Exact accumulation: hist[i] from the four int64 limb sums of value i
// (value = s3 2^96 + s2 2^64 + s1 2^32 + s0, times 2^-s).  The addends of one
// bin share a sign (the energy, pressure and analysis planes are >= 0 and each
// direction plane holds one sign), so the limb sums do too: the magnitude is
// carried up limb by limb and converted as two 64-bit halves (within 2 ulp;
// the same sums give the same double on every run).
__global__ void k_fx_flush(const long long *__restrict__ fx, double *__restrict__ hist, uint64_t stride,
                           double inv_se, double inv_sp)
{
    const uint64_t n = (uint64_t)kPlanes * stride;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        long long s0 = fx[i], s1 = fx[n + i], s2 = fx[2 * n + i], s3 = fx[3 * n + i];
        const bool neg = s0 < 0 || s1 < 0 || s2 < 0 || s3 < 0;
        if (neg) { s0 = -s0; s1 = -s1; s2 = -s2; s3 = -s3; }
        s1 += s0 >> 32;
        s0 &= 0xffffffffLL;
        s2 += s1 >> 32;
        s1 &= 0xffffffffLL;
        s3 += s2 >> 32;
        s2 &= 0xffffffffLL;
        const long long hi = (long long)((unsigned long long)s3 << 32) + s2;
        const unsigned long long lo = ((unsigned long long)s1 << 32) | (unsigned long long)s0;
        const double v = (double)hi * 0x1p64 + (double)lo;
        const double scale = (i / stride == 1) ? inv_sp : inv_se;
        hist[i] = (neg ? -v : v) * scale;
    }
}

hipError_t launch_fx_flush(hipStream_t st, const long long *fx, double *hist, uint64_t stride, double inv_se,
                           double inv_sp)
{
    const uint64_t n = (uint64_t)kPlanes * stride;
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fx_flush, dim3(blocks), dim3(256), 0, st, fx, hist, stride, inv_se, inv_sp);
    return hipGetLastError();
}

}  // namespace pach
