// denoise.hip -- the learned histogram denoiser of include/pach_denoise.h:
// a BF16 MFMA GEMM for gfx950 (v_mfma_f32_32x32x16_bf16, FP32 accumulation),
// the elementwise kernels around it, the network's forward and backward
// passes, and AdamW on FP32 master weights.
//
// GEMM tiling: a 256-thread workgroup computes a 128 x 128 tile of C; each of
// its four wave64s owns a 64 x 64 quarter as 2 x 2 MFMA 32x32 blocks.  K is
// consumed 64 at a time through LDS, where both operands are kept k-contiguous
// (8 chunks of 8 per row, rows padded to 72 elements, chunk positions
// XOR-swizzled by row so neither the writes nor the MFMA operand reads
// conflict on banks); each lane fetches its 8-element MFMA operand with one
// 16-byte read.  An operand stored with k as its row index is read in 8 x 8
// blocks and transposed in registers, so it too is written 16 bytes at a time.
// The global loads of the next K tile are issued before the MFMAs of the
// current one.  Tiles are dealt to workgroups in an XCD-aware order.  Weight
// gradients (K = every patch of the batch) split K over the grid and sum the
// partial tiles in a second kernel, so results are deterministic.
#include <hip/hip_runtime.h>

#include "../../include/pach.h"
#include "../../include/pach_denoise.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 128, kBN = 128, kBK = 64, kLdsK = kBK + 8;
constexpr int kChunks = kBM * kBK / 8 / 256;  // 16-byte chunks per operand tile per thread
constexpr int kThreads = 256;

// f(std::integral_constant<int, 0>{}) ... f(std::integral_constant<int, N - 1>{}):
// a loop whose index is a compile-time constant in each copy, so it can pick
// a register array by index without dynamic indexing
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Workgroup barrier that orders LDS accesses only: __syncthreads() also
// waits for every outstanding global load (s_waitcnt vmcnt(0) before
// s_barrier), which would drain the K tiles in flight at every step.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// v where ok, else zero, word by word (a select of two uint4 values is
// lowered through scratch memory)
__device__ __forceinline__ uint4 keep4(bool ok, const uint4 &v)
{
    const uint32_t m = ok ? ~0u : 0u;
    return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m);
}

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// tanh through one exponential: tanh y = 1 - 2 / (e^(2y) + 1), which saturates
// to +-1 for large |y| (e^(2y) overflows to inf or underflows to 0)
__device__ __forceinline__ float tanh_fast(float y)
{
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * y) + 1.0f);
}

__device__ __forceinline__ float gelu(float x)
{
    const float k = 0.7978845608028654f;  // sqrt(2 / pi)
    return 0.5f * x * (1.0f + tanh_fast(k * (x + 0.044715f * x * x * x)));
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float k = 0.7978845608028654f;
    const float t = tanh_fast(k * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * x * x);
}

struct Epi {
    const float *bias = nullptr;   // [N]
    const float *pos = nullptr;    // [npos, N], added to row m at m % npos
    int npos = 1;
    const float *dgelu = nullptr;  // [M, ldo] pre-activations: result *= gelu'(pre)
    float *out = nullptr;          // FP32 [M, ldo]
    int accumulate = 0;            // out += result
    uint16_t *out_bf = nullptr;    // BF16 [M, ldo] of (gelu ? gelu(result) : result)
    int gelu = 0;
    int ldo = 0;
    float *colpart = nullptr;      // [M tiles][N]: each tile's column sums of the result (after dgelu)
};

// C = op(A) op(B).  TA: A is stored [K, M] (row-major, lda), else [M, K].
// TB: B is stored [N, K], else [K, N].  The contiguous dimension of each
// operand and the leading dimensions are multiples of 8 (16-byte chunks;
// rows are bounds-checked one by one); blockIdx.z takes K range
// [z * kchunk, (z + 1) * kchunk), kchunk a multiple of kBK.  With `partial`
// the raw tile goes to partial[z] and the epilogue is skipped.
//
// Pipeline: an operand copied in k-contiguous chunks keeps kPD K tiles in
// flight in registers (the products with TA false stream their large A
// operand so; their grids hold only about two workgroups per CU, so one tile
// in flight left each K step waiting a full memory latency).  The other
// operand and every operand of the TA products keep one tile in flight.  Per
// step the loads are issued in the order they are consumed, so waiting for
// the oldest never waits for a younger one.
constexpr int kPD = 2;
template <bool TA, bool TB>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(TA || kPD == 1 ? 3 : 2, TA || kPD == 1 ? 3 : 2)))
k_gemm(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B, int ldb, int M, int N, int K,
       int kchunk, Epi epi, float *__restrict__ partial)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[kBM * kLdsK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN * kLdsK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    // XCD-aware tile order.  Workgroups are dealt to the 8 XCDs round-robin
    // in launch order, and each XCD has its own L2; tiles that share operand
    // rows (the N tiles of one M tile, the tiles of one K chunk) are given to
    // workgroups of one XCD, so those rows cross HBM once rather than once
    // per XCD.  XCD x holds launch ids x, x + 8, ... and takes the x-th
    // contiguous range of logical tiles (a bijection for any tile count).
    const int nx = (int)gridDim.x, ny = (int)gridDim.y;
    const int ntile = nx * ny * (int)gridDim.z;
    const int lid = (int)blockIdx.x + nx * ((int)blockIdx.y + ny * (int)blockIdx.z);
    const int q = ntile >> 3, r8 = ntile & 7, xcd = lid & 7;
    const int tile = xcd * q + min(xcd, r8) + (lid >> 3);
    const int tx = tile % nx, ty = (tile / nx) % ny, tz = tile / (nx * ny);
    const int m0 = ty * kBM, n0 = tx * kBN;
    const int kb = tz * kchunk, ke = min(K, kb + kchunk);
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // LDS: row r of an operand tile holds its 64 k values as 8 chunks of 8;
    // chunk c sits at position c ^ ((r / 8) % 8) of the row (rows padded to
    // 72 elements).  With that swizzle the 16-byte writes of a transposed block
    // (8 consecutive rows, one chunk) and the MFMA operand reads (consecutive
    // rows, one chunk) both fall on distinct banks.
    auto lds_off = [](int r, int c) { return r * kLdsK + ((c ^ ((r >> 3) & 7)) << 3); };
    // Operands whose rows run along k (A unless TA, B when TB) are copied in
    // 16-byte chunks: chunk c = tid + pass * 256 is row c / 8, k chunk c % 8.
    // An operand stored with k as its row index (A when TA, B unless TB) is
    // read in 8 x 8 blocks by 128 threads (threads 0-127; when both operands
    // are stored so, threads 128-255 take B): thread t loads k rows
    // 8 (b / 16) .. + 7 at columns 8 (b % 16) .. + 7, b = t % 128, and
    // transposes the block in registers into 8 k-contiguous chunks.
    constexpr bool XA = TA, XB = !TB;
    constexpr int PD = XA ? 1 : kPD;  // A tiles in flight
    constexpr int NDA = XA ? 1 : kChunks, NDB = XB ? 1 : kChunks;
    const bool blk_a = XA && tid < 128;
    const bool blk_b = XB && (XA ? tid >= 128 : tid < 128);
    const int bt = tid & 127, bkg = bt >> 4, bmg = bt & 15;
    uint4 rda[PD][NDA], rdb[NDB], rt[8];
    // In the products with TA false (UNC) loads are unconditional: an element
    // outside the operand (or a tile past the K range, fetched ahead and
    // never used) is read from a clamped address inside it and replaced by
    // zero, so the loop is straight-line code in which the compiler's memory
    // waits count only the loads a step needs, not every load in flight.
    constexpr bool UNC = !XA;
    // A's k-contiguous chunks of tile k0 into register set `set`
    auto load_a = [&](auto set, int k0) __attribute__((always_inline)) {
        if constexpr (!XA) {
#pragma unroll
            for (int pass = 0; pass < kChunks; ++pass) {
                const int c = tid + pass * kThreads, r = c >> 3, kc = (c & 7) * 8;
                const bool ok = m0 + r < M && k0 + kc < ke;
                const uint4 v = *(const uint4 *)(A + (size_t)min(m0 + r, M - 1) * lda + min(k0 + kc, ke - 8));
                rda[set][pass] = keep4(ok, v);
            }
        }
    };
    // the rest of tile k0: B's chunks, or the 8 x 8 blocks of a k-major operand
    auto load = [&](int k0) __attribute__((always_inline)) {
        if constexpr (!XB) {
#pragma unroll
            for (int pass = 0; pass < kChunks; ++pass) {
                const int c = tid + pass * kThreads, r = c >> 3, kc = (c & 7) * 8;
                const bool ok = n0 + r < N && k0 + kc < ke;
                if constexpr (UNC) {
                    const uint4 v = *(const uint4 *)(B + (size_t)min(n0 + r, N - 1) * ldb + min(k0 + kc, ke - 8));
                    rdb[pass] = keep4(ok, v);
                } else {
                    rdb[pass] = ok ? *(const uint4 *)(B + (size_t)(n0 + r) * ldb + k0 + kc) : make_uint4(0, 0, 0, 0);
                }
            }
        }
        if (blk_a || blk_b) {
            const uint16_t *X = blk_a ? A : B;
            const int ld = blk_a ? lda : ldb, col = (blk_a ? m0 : n0) + bmg * 8, ncol = blk_a ? M : N;
#pragma unroll
            for (int rr = 0; rr < 8; ++rr) {
                const int k = k0 + bkg * 8 + rr;
                const bool ok = k < ke && col < ncol;
                if constexpr (UNC) {
                    const uint4 v = *(const uint4 *)(X + (size_t)min(k, ke - 1) * ld + min(col, ncol - 8));
                    rt[rr] = keep4(ok, v);
                } else {
                    rt[rr] = ok ? *(const uint4 *)(X + (size_t)k * ld + col) : make_uint4(0, 0, 0, 0);
                }
            }
        }
    };
    auto word = [](const uint4 &v, int d) __attribute__((always_inline)) { return d == 0 ? v.x : (d == 1 ? v.y : (d == 2 ? v.z : v.w)); };
    auto store_a = [&](auto set) __attribute__((always_inline)) {
        if constexpr (!XA) {
#pragma unroll
            for (int pass = 0; pass < kChunks; ++pass) {
                const int c = tid + pass * kThreads;
                *(uint4 *)(As + lds_off(c >> 3, c & 7)) = rda[set][pass];
            }
        }
    };
    auto store = [&]() __attribute__((always_inline)) {
        if constexpr (!XB) {
#pragma unroll
            for (int pass = 0; pass < kChunks; ++pass) {
                const int c = tid + pass * kThreads;
                *(uint4 *)(Bs + lds_off(c >> 3, c & 7)) = rdb[pass];
            }
        }
        if (blk_a || blk_b) {
            uint16_t *Xs = blk_a ? As : Bs;
            // column mm of the block: word e packs k rows 2e (low half) and 2e + 1
#pragma unroll
            for (int mm = 0; mm < 8; ++mm) {
                uint32_t w[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t lo = word(rt[2 * e], mm >> 1), hi = word(rt[2 * e + 1], mm >> 1);
                    w[e] = (mm & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
                }
                *(uint4 *)(Xs + lds_off(bmg * 8 + mm, bkg)) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
    };
    // the MFMAs of the tile in LDS
    auto mma = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 16) {
            // operand layout of the 32x32x16 MFMA: lane l supplies row (l % 32)
            // of A and column (l % 32) of B at k = 8 (l / 32) .. 8 (l / 32) + 7
            const int kc = (kk >> 3) + (lane >> 5);
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *(const bf16x8 *)(As + lds_off(wm * 64 + i * 32 + (lane & 31), kc));
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *(const bf16x8 *)(Bs + lds_off(wn * 64 + j * 32 + (lane & 31), kc));
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    // prologue: A tiles 0 .. PD - 1 and the first tile of the rest
    const int nsteps = kb < ke ? (ke - kb + kBK - 1) / kBK : 0;
    if (nsteps > 0) {
        load(kb);
        static_for<PD>([&](auto u) __attribute__((always_inline)) { load_a(u, kb + (int)u * kBK); });
    }
    // one K step with A in register set u (compile-time): consume tile k0,
    // then issue the loads of tile k0 + kBK (rest) and k0 + PD kBK (A)
    auto step = [&](auto u, int k0) __attribute__((always_inline)) {
        store_a(u);
        store();
        lds_barrier();
        if (UNC || k0 + kBK < ke) load(k0 + kBK);
        load_a(u, k0 + PD * kBK);
        mma();
        lds_barrier();
    };
    int st = 0;
    for (; st + PD <= nsteps; st += PD)
        static_for<PD>([&](auto u) __attribute__((always_inline)) { step(u, kb + (st + (int)u) * kBK); });
    static_for<PD>([&](auto u) __attribute__((always_inline)) {
        if (st + (int)u < nsteps) step(u, kb + (st + (int)u) * kBK);
    });
    // result layout: lane l, register v holds C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
    // One 32 x 32 block at a time, every input the epilogue reads (position
    // rows, pre-activations, the output it accumulates into) is loaded for the
    // whole block before its first store, so those loads are in flight
    // together rather than one round trip per element.  Which inputs exist is
    // uniform over the grid (scalar branches), a lane's 16 rows are offsets
    // 8 (v / 4) + v % 4 < 32 from its first row, so every address is one
    // 64-bit base plus a 32-bit offset, and the position-table row (row mod
    // npos) is stepped from the first row's instead of divided per element.
    const int ldo = epi.ldo;
    auto roff = [](int v) __attribute__((always_inline)) { return (v >> 2) * 8 + (v & 3); };
    // column sums of this tile (bias gradients): a lane's 16 rows, then the
    // other half-wave's 16 rows of the same column, then the wave's two row
    // blocks, then the two waves along M, in that fixed order
    float csum[2] = {0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn * 64 + j * 32 + (lane & 31);
            const int rb = m0 + wm * 64 + i * 32 + (lane >> 5) * 4;
            const bool cok = col < N;
            if (partial) {
                float *pz = partial + (size_t)tz * M * N + (size_t)rb * N + col;
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (cok && rb + roff(v) < M) pz[roff(v) * N] = acc[i][j][v];
                continue;
            }
            const size_t base = (size_t)rb * ldo + col;
            float r[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) r[v] = acc[i][j][v];
            if (epi.bias) {
                const float bias = cok ? epi.bias[col] : 0.0f;
#pragma unroll
                for (int v = 0; v < 16; ++v) r[v] += bias;
            }
            float pre[16], old[16];
            if (epi.pos) {
                const int np = epi.npos;
                const int p0 = rb % np;
                const float *pc = epi.pos + col;
                float posv[16];
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    int pr = p0 + roff(v);
                    if (np >= 32) pr = pr >= np ? pr - np : pr;
                    else pr %= np;
                    posv[v] = (cok && rb + roff(v) < M) ? pc[(size_t)pr * N] : 0.0f;
                }
#pragma unroll
                for (int v = 0; v < 16; ++v) r[v] += posv[v];
            }
            if (epi.dgelu) {
                const float *pp = epi.dgelu + base;
#pragma unroll
                for (int v = 0; v < 16; ++v) pre[v] = (cok && rb + roff(v) < M) ? pp[roff(v) * ldo] : 0.0f;
#pragma unroll
                for (int v = 0; v < 16; ++v) r[v] *= gelu_grad(pre[v]);
            }
            if (epi.colpart) {
                float cs = 0.0f;
#pragma unroll
                for (int v = 0; v < 16; ++v) cs += rb + roff(v) < M ? r[v] : 0.0f;
                const float other = __shfl_xor(cs, 32);
                csum[j] += lane < 32 ? cs + other : other + cs;
            }
            if (epi.out) {
                float *po = epi.out + base;
                if (epi.accumulate) {
#pragma unroll
                    for (int v = 0; v < 16; ++v) old[v] = (cok && rb + roff(v) < M) ? po[roff(v) * ldo] : 0.0f;
#pragma unroll
                    for (int v = 0; v < 16; ++v) r[v] += old[v];
                }
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (cok && rb + roff(v) < M) po[roff(v) * ldo] = r[v];
            }
            if (epi.out_bf) {
                uint16_t *pb = epi.out_bf + base;
                if (epi.gelu) {
#pragma unroll
                    for (int v = 0; v < 16; ++v) r[v] = gelu(r[v]);
                }
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (cok && rb + roff(v) < M) pb[roff(v) * ldo] = f2bf(r[v]);
            }
        }
    if (epi.colpart) {
        // the wave of the upper row half hands its sums to the lower one
        __shared__ float cs_up[2][64];
        if (wm == 1 && lane < 32) {
            cs_up[wn][lane] = csum[0];
            cs_up[wn][32 + lane] = csum[1];
        }
        lds_barrier();
        if (wm == 0 && lane < 32) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = n0 + wn * 64 + j * 32 + lane;
                if (col < N) epi.colpart[(size_t)ty * N + col] = csum[j] + cs_up[wn][j * 32 + lane];
            }
        }
    }
}

// Sums of `parts` FP32 slices of n values (split-K partial tiles, column-sum
// parts), in a fixed order: four interleaved running sums combined at the end,
// so four loads are in flight per step instead of one.
__device__ __forceinline__ float sum_slices(const float *__restrict__ p, int parts, size_t n, size_t i)
{
    float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int z = 0;
    for (; z + 4 <= parts; z += 4) {
        s0 += p[(size_t)z * n + i];
        s1 += p[(size_t)(z + 1) * n + i];
        s2 += p[(size_t)(z + 2) * n + i];
        s3 += p[(size_t)(z + 3) * n + i];
    }
    for (; z < parts; ++z) s0 += p[(size_t)z * n + i];
    return (s0 + s1) + (s2 + s3);
}

__global__ void k_splitk_sum(const float *__restrict__ partial, int splits, size_t n, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sum_slices(partial, splits, n, i);
}

// Column sums of in[M, N] in two deterministic stages: partial[gridDim.y][N].
__global__ void k_colsum_part(const float *__restrict__ in, int M, int N, float *__restrict__ partial)
{
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + tx;
    // rows blockIdx.y * 4 + ty + j * S (S = 4 gridDim.y), four running sums
    float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (col < N) {
        const int S = gridDim.y * 4;
        int r = blockIdx.y * 4 + ty;
        for (; r + 3 * S < M; r += 4 * S) {
            s0 += in[(size_t)r * N + col];
            s1 += in[(size_t)(r + S) * N + col];
            s2 += in[(size_t)(r + 2 * S) * N + col];
            s3 += in[(size_t)(r + 3 * S) * N + col];
        }
        for (; r < M; r += S) s0 += in[(size_t)r * N + col];
    }
    red[ty][tx] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (ty == 0 && col < N) partial[(size_t)blockIdx.y * N + col] = red[0][tx] + red[1][tx] + red[2][tx] + red[3][tx];
}

__global__ void k_colsum_fin(const float *__restrict__ partial, int parts, int N, float *__restrict__ out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= N) return;
    out[col] = sum_slices(partial, parts, (size_t)N, (size_t)col);
}

// dPos[p][n] = sum over samples b of d[b * P + p][n]
__global__ void k_possum(const float *__restrict__ d, int B, int P, int N, float *__restrict__ out)
{
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + tx, p = blockIdx.y;
    float s = 0;
    if (col < N)
        for (int b = ty; b < B; b += 4) s += d[((size_t)b * P + p) * N + col];
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && col < N) out[(size_t)p * N + col] = red[0][tx] + red[1][tx] + red[2][tx] + red[3][tx];
}

// Feature rows of the patches: [8 * patch band-major bins | descriptors | zeros].
__global__ void k_patchify(const float *__restrict__ x, const float *__restrict__ desc, int B, int P, int patch,
                           int bins, int n_desc, int KE, uint16_t *__restrict__ X0)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)B * P * KE;
    if (i >= total) return;
    const int f = (int)(i % KE);
    const size_t m = i / KE;
    const int b = (int)(m / P), p = (int)(m % P);
    const int F = 8 * patch;
    float v = 0;
    if (f < F) v = x[((size_t)b * 8 + f / patch) * bins + (size_t)p * patch + f % patch];
    else if (f < F + n_desc) v = desc[(size_t)b * n_desc + (f - F)];
    X0[i] = f2bf(v);
}

// Kernel-3 convolution over patches as a product: C[m][tap * D + c] =
// H[m + tap - 1][c] inside the same sample, else 0.  8 channels per thread.
__global__ void k_im2col(const uint16_t *__restrict__ H, int B, int P, int D, uint16_t *__restrict__ C)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int D8 = D / 8;
    const size_t total = (size_t)B * P * 3 * D8;
    if (i >= total) return;
    const int c8 = (int)(i % D8);
    const int tap = (int)((i / D8) % 3);
    const size_t m = i / ((size_t)3 * D8);
    const int p = (int)(m % P) + tap - 1;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (p >= 0 && p < P) v = *(const uint4 *)(H + (m + tap - 1) * D + (size_t)c8 * 8);
    *(uint4 *)(C + m * 3 * D + (size_t)tap * D + (size_t)c8 * 8) = v;
}

// Adjoint of k_im2col, added into dH; refreshes the BF16 copy of dH.
__global__ void k_col2im_add(const float *__restrict__ dC, int B, int P, int D, float *__restrict__ dH,
                             uint16_t *__restrict__ dHb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)B * P * D;
    if (i >= total) return;
    const int c = (int)(i % D);
    const size_t m = i / D;
    const int p = (int)(m % P);
    float s = dH[i] + dC[m * 3 * D + D + c];                         // tap 1 of row m
    if (p + 1 < P) s += dC[(m + 1) * 3 * D + c];                     // tap 0 of row m + 1
    if (p - 1 >= 0) s += dC[(m - 1) * 3 * D + 2 * (size_t)D + c];    // tap 2 of row m - 1
    dH[i] = s;
    dHb[i] = f2bf(s);
}

// out = x + unpatchify(Y); with a target also the loss terms w[t] (out - tgt)^2
// and dL/dY (w: per-bin loss weights, null for 1).
__global__ void k_output(const float *__restrict__ Y, const float *__restrict__ x, const float *__restrict__ tgt,
                         const float *__restrict__ w, int B, int P, int patch, int bins, float *__restrict__ out,
                         float *__restrict__ dY, uint16_t *__restrict__ dYb, double *__restrict__ loss_part)
{
    __shared__ double red[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)B * 8 * bins;
    const int F = 8 * patch;
    double l = 0;
    if (i < total) {
        const int t = (int)(i % bins);
        const int band = (int)((i / bins) % 8);
        const size_t b = i / ((size_t)8 * bins);
        const size_t yi = (b * P + t / patch) * F + (size_t)band * patch + t % patch;
        const float o = x[i] + Y[yi];
        if (out) out[i] = o;
        if (tgt) {
            const float d = o - tgt[i];
            const float wt = w ? w[t] : 1.0f;
            l = (double)wt * d * d;
            const float g = 2.0f * wt * d / (float)total;
            dY[yi] = g;
            dYb[yi] = f2bf(g);
        }
    }
    if (!loss_part) return;
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

__global__ void k_gather(const float *__restrict__ src, const int64_t *__restrict__ idx, int B, size_t row,
                         float *__restrict__ dst)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * row) return;
    const size_t b = i / row;
    dst[i] = src[(size_t)idx[b] * row + i % row];
}

__global__ void k_adamw(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                        float *__restrict__ v, uint16_t *__restrict__ wb, size_t n, float lr, float b1, float b2,
                        float eps, float wd, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * ((mi * c1) / (sqrtf(vi * c2) + eps) + wd * w[i]);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *__restrict__ in, size_t n, uint16_t *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

inline unsigned nblk(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct pach_dn {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    pach_dn_config cfg{};
    int P = 0, F = 0, KE = 0, D = 0, NB = 0, maxM = 0;
    int64_t n_params = 0;
    int64_t oWe = 0, obe = 0, oPos = 0, oWh = 0, obh = 0;
    std::vector<int64_t> oWc, obc, oW1, ob1, oW2, ob2;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wb = nullptr;
    int64_t step = 0;
    // activations (max batch)
    uint16_t *X0b = nullptr, *C1b = nullptr;
    float *H = nullptr, *Y = nullptr;
    std::vector<uint16_t *> Hb, Ub, Vb;
    std::vector<float *> Upre, Vpre;
    float *dY = nullptr, *dH = nullptr, *dC1 = nullptr;
    uint16_t *dYb = nullptr, *dHb = nullptr, *dVb = nullptr, *dUb = nullptr;
    float *partial = nullptr;
    size_t partial_n = 0;
    float *colpart = nullptr;
    float *tilecs = nullptr;  // [M tiles][N] column sums written by GEMM epilogues
    double *lossp = nullptr;
    size_t lossp_n = 0;
    float *x = nullptr, *desc = nullptr, *tgt = nullptr, *out = nullptr;
    float *lw = nullptr;  // per-bin loss weights, null until set
    int64_t *idx = nullptr;
    float *ds_x = nullptr, *ds_desc = nullptr, *ds_t = nullptr;
    int64_t ds_n = 0;
    std::vector<void *> allocs;
};

namespace {

constexpr int kColParts = 128;

int dn_fail(pach_dn *d, int code, const std::string &msg)
{
    if (d) d->err = msg;
    return code;
}

#define DNCHK(d, expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return dn_fail(d, PACH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> hipError_t dn_alloc(pach_dn *d, T *&p, size_t n)
{
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) d->allocs.push_back(p);
    return e;
}

// C = op(A) op(B) with the epilogue; split-K (into `out`, overwritten) when splits > 1.
int gemm(pach_dn *d, bool ta, bool tb, const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K,
         const Epi &epi, int splits = 1)
{
    // operands are read in 8-element chunks along their contiguous dimension
    if ((ta ? M : K) % 8 || (tb ? K : N) % 8 || lda % 8 || ldb % 8)
        return dn_fail(d, PACH_EINVAL, "gemm: contiguous operand dimensions must be multiples of 8");
    splits = std::max(1, std::min(splits, (K + kBK - 1) / kBK));
    int kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + kBK - 1) / kBK * kBK;
    splits = (K + kchunk - 1) / kchunk;
    float *part = nullptr;
    if (splits > 1) {
        if ((size_t)splits * M * N > d->partial_n || !epi.out || epi.accumulate || epi.bias || epi.pos || epi.dgelu ||
            epi.out_bf || epi.colpart || epi.ldo != N)
            return dn_fail(d, PACH_EINVAL, "gemm: split-K takes a plain FP32 output");
        part = d->partial;
    }
    const dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, splits);
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(kThreads), 0, d->st, A, lda, B, ldb, M, N, K, kchunk, epi, part);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(kThreads), 0, d->st, A, lda, B, ldb, M, N, K, kchunk, epi, part);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(kThreads), 0, d->st, A, lda, B, ldb, M, N, K, kchunk, epi, part);
    else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(kThreads), 0, d->st, A, lda, B, ldb, M, N, K, kchunk, epi, part);
    DNCHK(d, hipGetLastError());
    if (splits > 1) {
        const size_t n = (size_t)M * N;
        hipLaunchKernelGGL(k_splitk_sum, dim3(nblk(n)), dim3(256), 0, d->st, (const float *)part, splits, n, epi.out);
        DNCHK(d, hipGetLastError());
    }
    return PACH_OK;
}

int colsum(pach_dn *d, const float *in, int M, int N, float *out);

// out[n] = sum over the M tiles of a GEMM's epilogue column sums: the tile
// sums form a [tiles][N] matrix reduced like any other (fixed order)
int colsum_tiles(pach_dn *d, int M, int N, float *out)
{
    return colsum(d, d->tilecs, (M + kBM - 1) / kBM, N, out);
}

int colsum(pach_dn *d, const float *in, int M, int N, float *out)
{
    hipLaunchKernelGGL(k_colsum_part, dim3((N + 63) / 64, kColParts), dim3(256), 0, d->st, in, M, N, d->colpart);
    hipLaunchKernelGGL(k_colsum_fin, dim3(nblk(N)), dim3(256), 0, d->st, (const float *)d->colpart, kColParts, N, out);
    DNCHK(d, hipGetLastError());
    return PACH_OK;
}

// split count for weight gradients: about 4 workgroups per CU-sized wave of tiles
int wsplits(int M, int N, int K)
{
    const int tiles = ((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN);
    return std::max(1, std::min(64, std::min(1024 / std::max(tiles, 1), K / 512)));
}

int forward(pach_dn *d, int B)
{
    const int P = d->P, D = d->D, M = B * P;
    hipLaunchKernelGGL(k_patchify, dim3(nblk((size_t)M * d->KE)), dim3(256), 0, d->st, (const float *)d->x,
                       (const float *)d->desc, B, P, d->cfg.patch, d->cfg.bins, d->cfg.n_desc, d->KE, d->X0b);
    DNCHK(d, hipGetLastError());
    Epi e;
    e.bias = d->w + d->obe;
    e.pos = d->w + d->oPos;
    e.npos = P;
    e.out = d->H;
    e.out_bf = d->Hb[0];
    e.ldo = D;
    if (int rc = gemm(d, false, false, d->X0b, d->KE, d->wb + d->oWe, D, M, D, d->KE, e)) return rc;
    for (int i = 0; i < d->NB; ++i) {
        hipLaunchKernelGGL(k_im2col, dim3(nblk((size_t)M * 3 * D / 8)), dim3(256), 0, d->st,
                           (const uint16_t *)d->Hb[i], B, P, D, d->C1b);
        DNCHK(d, hipGetLastError());
        Epi u;
        u.bias = d->w + d->obc[i];
        u.out = d->Upre[i];
        u.out_bf = d->Ub[i];
        u.gelu = 1;
        u.ldo = D;
        if (int rc = gemm(d, false, false, d->C1b, 3 * D, d->wb + d->oWc[i], D, M, D, 3 * D, u)) return rc;
        Epi vv;
        vv.bias = d->w + d->ob1[i];
        vv.out = d->Vpre[i];
        vv.out_bf = d->Vb[i];
        vv.gelu = 1;
        vv.ldo = 2 * D;
        if (int rc = gemm(d, false, false, d->Ub[i], D, d->wb + d->oW1[i], 2 * D, M, 2 * D, D, vv)) return rc;
        Epi h;
        h.bias = d->w + d->ob2[i];
        h.out = d->H;
        h.accumulate = 1;
        h.out_bf = d->Hb[i + 1];
        h.ldo = D;
        if (int rc = gemm(d, false, false, d->Vb[i], 2 * D, d->wb + d->oW2[i], D, M, D, 2 * D, h)) return rc;
    }
    Epi y;
    y.bias = d->w + d->obh;
    y.out = d->Y;
    y.ldo = d->F;
    return gemm(d, false, false, d->Hb[d->NB], D, d->wb + d->oWh, d->F, M, d->F, D, y);
}

int backward(pach_dn *d, int B)
{
    const int P = d->P, D = d->D, M = B * P, F = d->F;
    int rc;
    Epi pw;  // plain FP32 product into the gradient vector
    if ((rc = colsum(d, d->dY, M, F, d->g + d->obh))) return rc;
    pw.out = d->g + d->oWh;
    pw.ldo = F;
    if ((rc = gemm(d, true, false, d->Hb[d->NB], D, d->dYb, F, D, F, M, pw, wsplits(D, F, M)))) return rc;
    Epi dh;
    dh.out = d->dH;
    dh.out_bf = d->dHb;
    dh.ldo = D;
    if ((rc = gemm(d, false, true, d->dYb, F, d->wb + d->oWh, F, M, D, F, dh))) return rc;
    for (int i = d->NB - 1; i >= 0; --i) {
        // H_out = H_in + gelu(U W1 + b1) W2 + b2
        if ((rc = colsum(d, d->dH, M, D, d->g + d->ob2[i]))) return rc;
        pw.out = d->g + d->oW2[i];
        pw.ldo = D;
        if ((rc = gemm(d, true, false, d->Vb[i], 2 * D, d->dHb, D, 2 * D, D, M, pw, wsplits(2 * D, D, M)))) return rc;
        // dV = (dH W2^T) gelu'(Vpre): only its BF16 copy and its column sums
        // (the b1 gradient) are used, so no FP32 copy is written
        Epi dv;
        dv.dgelu = d->Vpre[i];
        dv.out_bf = d->dVb;
        dv.colpart = d->tilecs;
        dv.ldo = 2 * D;
        if ((rc = gemm(d, false, true, d->dHb, D, d->wb + d->oW2[i], D, M, 2 * D, D, dv))) return rc;
        if ((rc = colsum_tiles(d, M, 2 * D, d->g + d->ob1[i]))) return rc;
        pw.out = d->g + d->oW1[i];
        pw.ldo = 2 * D;
        if ((rc = gemm(d, true, false, d->Ub[i], D, d->dVb, 2 * D, D, 2 * D, M, pw, wsplits(D, 2 * D, M)))) return rc;
        Epi du;
        du.dgelu = d->Upre[i];
        du.out_bf = d->dUb;
        du.colpart = d->tilecs;
        du.ldo = D;
        if ((rc = gemm(d, false, true, d->dVb, 2 * D, d->wb + d->oW1[i], 2 * D, M, D, 2 * D, du))) return rc;
        if ((rc = colsum_tiles(d, M, D, d->g + d->obc[i]))) return rc;
        hipLaunchKernelGGL(k_im2col, dim3(nblk((size_t)M * 3 * D / 8)), dim3(256), 0, d->st,
                           (const uint16_t *)d->Hb[i], B, P, D, d->C1b);
        DNCHK(d, hipGetLastError());
        pw.out = d->g + d->oWc[i];
        pw.ldo = D;
        if ((rc = gemm(d, true, false, d->C1b, 3 * D, d->dUb, D, 3 * D, D, M, pw, wsplits(3 * D, D, M)))) return rc;
        Epi dc;
        dc.out = d->dC1;
        dc.ldo = 3 * D;
        if ((rc = gemm(d, false, true, d->dUb, D, d->wb + d->oWc[i], D, M, 3 * D, D, dc))) return rc;
        hipLaunchKernelGGL(k_col2im_add, dim3(nblk((size_t)M * D)), dim3(256), 0, d->st, (const float *)d->dC1, B, P,
                           D, d->dH, d->dHb);
        DNCHK(d, hipGetLastError());
    }
    if ((rc = colsum(d, d->dH, M, D, d->g + d->obe))) return rc;
    hipLaunchKernelGGL(k_possum, dim3((D + 63) / 64, P), dim3(256), 0, d->st, (const float *)d->dH, B, P, D,
                       d->g + d->oPos);
    DNCHK(d, hipGetLastError());
    pw.out = d->g + d->oWe;
    pw.ldo = D;
    return gemm(d, true, false, d->X0b, d->KE, d->dHb, D, d->KE, D, M, pw, wsplits(d->KE, D, M));
}

// loss of the staged batch (x, desc, tgt) and, with grad, the full gradient in d->g
int loss_and_grad(pach_dn *d, int B, bool grad, double *loss)
{
    if (int rc = forward(d, B)) return rc;
    const size_t total = (size_t)B * 8 * d->cfg.bins;
    const unsigned nb = nblk(total);
    hipLaunchKernelGGL(k_output, dim3(nb), dim3(256), 0, d->st, (const float *)d->Y, (const float *)d->x,
                       (const float *)d->tgt, (const float *)d->lw, B, d->P, d->cfg.patch, d->cfg.bins, (float *)nullptr, d->dY, d->dYb,
                       d->lossp);
    DNCHK(d, hipGetLastError());
    if (grad)
        if (int rc = backward(d, B)) return rc;
    std::vector<double> lp(nb);
    DNCHK(d, hipMemcpyAsync(lp.data(), d->lossp, nb * sizeof(double), hipMemcpyDeviceToHost, d->st));
    DNCHK(d, hipStreamSynchronize(d->st));
    double s = 0;
    for (double v : lp) s += v;
    if (loss) *loss = s / (double)total;
    return PACH_OK;
}

int check_batch(pach_dn *d, int batch)
{
    if (batch < 1 || batch > d->cfg.max_batch) return dn_fail(d, PACH_EINVAL, "batch outside [1, max_batch]");
    return PACH_OK;
}

}  // namespace

extern "C" int pach_dn_create(pach_dn **out, int device, const pach_dn_config *cfg)
{
    if (!out || !cfg) return PACH_EINVAL;
    *out = nullptr;
    const pach_dn_config c = *cfg;
    if (c.bins <= 0 || c.patch <= 0 || c.bins % c.patch || c.width <= 0 || c.width % 64 || c.blocks < 0 ||
        c.n_desc < 0 || c.max_batch <= 0 || (8 * c.patch) % 8 || !(c.beta1 >= 0 && c.beta1 < 1) ||
        !(c.beta2 >= 0 && c.beta2 < 1) || !(c.eps > 0))
        return PACH_EINVAL;
    pach_dn *d = new pach_dn;
    d->device = device;
    d->cfg = c;
    d->P = c.bins / c.patch;
    d->F = 8 * c.patch;
    d->KE = (d->F + c.n_desc + 31) / 32 * 32;
    d->D = c.width;
    d->NB = c.blocks;
    const int64_t P = d->P, D = d->D, KE = d->KE, F = d->F;
    if ((int64_t)c.max_batch * P > (1 << 24)) { delete d; return PACH_EINVAL; }
    d->maxM = c.max_batch * d->P;
    // parameter offsets (multiples of 8 elements: D % 64 == 0, KE % 32 == 0, F % 8 == 0)
    int64_t o = 0;
    d->oWe = o; o += KE * D;
    d->obe = o; o += D;
    d->oPos = o; o += (P * D + 7) / 8 * 8;
    for (int i = 0; i < d->NB; ++i) {
        d->oWc.push_back(o); o += 3 * D * D;
        d->obc.push_back(o); o += D;
        d->oW1.push_back(o); o += D * 2 * D;
        d->ob1.push_back(o); o += 2 * D;
        d->oW2.push_back(o); o += 2 * D * D;
        d->ob2.push_back(o); o += D;
    }
    d->oWh = o; o += D * F;
    d->obh = o; o += F;
    d->n_params = o;
    auto bad = [&](hipError_t e) {
        if (e == hipSuccess) return false;
        d->err = hipGetErrorString(e);
        return true;
    };
    if (bad(hipSetDevice(device)) || bad(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking))) {
        pach_dn_destroy(d);
        return PACH_EHIP;
    }
    const size_t M = d->maxM;
    d->Hb.resize(d->NB + 1);
    d->Ub.resize(d->NB);
    d->Vb.resize(d->NB);
    d->Upre.resize(d->NB);
    d->Vpre.resize(d->NB);
    size_t maxwn = (size_t)KE * D;
    maxwn = std::max(maxwn, (size_t)3 * D * D);
    maxwn = std::max(maxwn, (size_t)2 * D * D);
    maxwn = std::max(maxwn, (size_t)D * F);
    d->partial_n = maxwn * 64;
    d->lossp_n = nblk((size_t)c.max_batch * 8 * c.bins);
    bool fail = bad(dn_alloc(d, d->w, o)) || bad(dn_alloc(d, d->g, o)) || bad(dn_alloc(d, d->m, o)) ||
                bad(dn_alloc(d, d->v, o)) || bad(dn_alloc(d, d->wb, o)) || bad(dn_alloc(d, d->X0b, M * KE)) ||
                bad(dn_alloc(d, d->C1b, M * 3 * D)) || bad(dn_alloc(d, d->H, M * D)) || bad(dn_alloc(d, d->Y, M * F)) ||
                bad(dn_alloc(d, d->dY, M * F)) || bad(dn_alloc(d, d->dYb, M * F)) || bad(dn_alloc(d, d->dH, M * D)) ||
                bad(dn_alloc(d, d->dHb, M * D)) ||
                bad(dn_alloc(d, d->dVb, M * 2 * D)) || bad(dn_alloc(d, d->dUb, M * D)) ||
                bad(dn_alloc(d, d->dC1, M * 3 * D)) || bad(dn_alloc(d, d->partial, d->partial_n)) ||
                bad(dn_alloc(d, d->colpart, (size_t)kColParts * std::max<int64_t>(3 * D, std::max<int64_t>(KE, F)))) ||
                bad(dn_alloc(d, d->tilecs, ((M + kBM - 1) / kBM) * (size_t)(2 * D))) ||
                bad(dn_alloc(d, d->lossp, d->lossp_n)) || bad(dn_alloc(d, d->x, (size_t)c.max_batch * 8 * c.bins)) ||
                bad(dn_alloc(d, d->tgt, (size_t)c.max_batch * 8 * c.bins)) ||
                bad(dn_alloc(d, d->out, (size_t)c.max_batch * 8 * c.bins)) ||
                bad(dn_alloc(d, d->desc, (size_t)c.max_batch * std::max(c.n_desc, 1))) ||
                bad(dn_alloc(d, d->idx, (size_t)c.max_batch));
    for (int i = 0; !fail && i <= d->NB; ++i) fail = bad(dn_alloc(d, d->Hb[i], M * D));
    for (int i = 0; !fail && i < d->NB; ++i)
        fail = bad(dn_alloc(d, d->Ub[i], M * D)) || bad(dn_alloc(d, d->Vb[i], M * 2 * D)) ||
               bad(dn_alloc(d, d->Upre[i], M * D)) || bad(dn_alloc(d, d->Vpre[i], M * 2 * D));
    if (fail) {
        pach_dn_destroy(d);
        return PACH_ENOMEM;
    }
    // zero parameters and moments (pach_dn_set_params brings the real values)
    if (bad(hipMemset(d->w, 0, o * 4)) || bad(hipMemset(d->m, 0, o * 4)) || bad(hipMemset(d->v, 0, o * 4)) ||
        bad(hipMemset(d->wb, 0, o * 2)) || bad(hipMemset(d->desc, 0, (size_t)c.max_batch * std::max(c.n_desc, 1) * 4))) {
        pach_dn_destroy(d);
        return PACH_EHIP;
    }
    *out = d;
    return PACH_OK;
}

extern "C" void pach_dn_destroy(pach_dn *d)
{
    if (!d) return;
    (void)hipSetDevice(d->device);
    if (d->st) (void)hipStreamSynchronize(d->st);
    for (void *p : d->allocs) (void)hipFree(p);
    if (d->ds_x) (void)hipFree(d->ds_x);
    if (d->ds_t) (void)hipFree(d->ds_t);
    if (d->ds_desc) (void)hipFree(d->ds_desc);
    if (d->lw) (void)hipFree(d->lw);
    if (d->st) (void)hipStreamDestroy(d->st);
    delete d;
}

extern "C" const char *pach_dn_last_error(pach_dn *d) { return d ? d->err.c_str() : "null handle"; }

extern "C" int64_t pach_dn_param_count(pach_dn *d) { return d ? d->n_params : -1; }

extern "C" int pach_dn_set_params(pach_dn *d, const float *params)
{
    if (!d || !params) return PACH_EINVAL;
    (void)hipSetDevice(d->device);
    const size_t n = (size_t)d->n_params;
    DNCHK(d, hipMemcpyAsync(d->w, params, n * 4, hipMemcpyHostToDevice, d->st));
    DNCHK(d, hipMemsetAsync(d->m, 0, n * 4, d->st));
    DNCHK(d, hipMemsetAsync(d->v, 0, n * 4, d->st));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(256), 0, d->st, (const float *)d->w, n, d->wb);
    DNCHK(d, hipGetLastError());
    DNCHK(d, hipStreamSynchronize(d->st));
    d->step = 0;
    return PACH_OK;
}

extern "C" int pach_dn_get_params(pach_dn *d, float *params)
{
    if (!d || !params) return PACH_EINVAL;
    (void)hipSetDevice(d->device);
    DNCHK(d, hipMemcpyAsync(params, d->w, (size_t)d->n_params * 4, hipMemcpyDeviceToHost, d->st));
    DNCHK(d, hipStreamSynchronize(d->st));
    return PACH_OK;
}

static int stage(pach_dn *d, const float *x, const float *desc, const float *tgt, int B)
{
    const size_t n = (size_t)B * 8 * d->cfg.bins;
    DNCHK(d, hipMemcpyAsync(d->x, x, n * 4, hipMemcpyHostToDevice, d->st));
    if (tgt) DNCHK(d, hipMemcpyAsync(d->tgt, tgt, n * 4, hipMemcpyHostToDevice, d->st));
    if (d->cfg.n_desc > 0) {
        if (!desc) return dn_fail(d, PACH_EINVAL, "descriptors missing");
        DNCHK(d, hipMemcpyAsync(d->desc, desc, (size_t)B * d->cfg.n_desc * 4, hipMemcpyHostToDevice, d->st));
    }
    return PACH_OK;
}

extern "C" int pach_dn_forward(pach_dn *d, const float *x, const float *desc, int batch, float *out)
{
    if (!d || !x || !out) return PACH_EINVAL;
    if (int rc = check_batch(d, batch)) return rc;
    (void)hipSetDevice(d->device);
    if (int rc = stage(d, x, desc, nullptr, batch)) return rc;
    if (int rc = forward(d, batch)) return rc;
    const size_t total = (size_t)batch * 8 * d->cfg.bins;
    hipLaunchKernelGGL(k_output, dim3(nblk(total)), dim3(256), 0, d->st, (const float *)d->Y, (const float *)d->x,
                       (const float *)nullptr, (const float *)nullptr, batch, d->P, d->cfg.patch, d->cfg.bins, d->out, (float *)nullptr,
                       (uint16_t *)nullptr, (double *)nullptr);
    DNCHK(d, hipGetLastError());
    DNCHK(d, hipMemcpyAsync(out, d->out, total * 4, hipMemcpyDeviceToHost, d->st));
    DNCHK(d, hipStreamSynchronize(d->st));
    return PACH_OK;
}

extern "C" int pach_dn_loss_grad(pach_dn *d, const float *x, const float *desc, const float *target, int batch,
                                 double *loss, float *grad)
{
    if (!d || !x || !target) return PACH_EINVAL;
    if (int rc = check_batch(d, batch)) return rc;
    (void)hipSetDevice(d->device);
    if (int rc = stage(d, x, desc, target, batch)) return rc;
    if (int rc = loss_and_grad(d, batch, grad != nullptr, loss)) return rc;
    if (grad) {
        DNCHK(d, hipMemcpyAsync(grad, d->g, (size_t)d->n_params * 4, hipMemcpyDeviceToHost, d->st));
        DNCHK(d, hipStreamSynchronize(d->st));
    }
    return PACH_OK;
}

extern "C" int pach_dn_set_loss_weights(pach_dn *d, const float *w)
{
    if (!d) return PACH_EINVAL;
    (void)hipSetDevice(d->device);
    DNCHK(d, hipStreamSynchronize(d->st));
    if (!w) {
        if (d->lw) (void)hipFree(d->lw);
        d->lw = nullptr;
        return PACH_OK;
    }
    for (int t = 0; t < d->cfg.bins; ++t)
        if (!(w[t] >= 0) || !std::isfinite(w[t])) return dn_fail(d, PACH_EINVAL, "loss weights must be finite and >= 0");
    if (!d->lw && hipMalloc((void **)&d->lw, (size_t)d->cfg.bins * 4) != hipSuccess)
        return dn_fail(d, PACH_ENOMEM, "loss weights");
    DNCHK(d, hipMemcpy(d->lw, w, (size_t)d->cfg.bins * 4, hipMemcpyHostToDevice));
    return PACH_OK;
}

extern "C" int pach_dn_dataset(pach_dn *d, const float *x, const float *desc, const float *target, int64_t n)
{
    if (!d || !x || !target || n <= 0 || (d->cfg.n_desc > 0 && !desc)) return PACH_EINVAL;
    (void)hipSetDevice(d->device);
    DNCHK(d, hipStreamSynchronize(d->st));
    for (float **p : {&d->ds_x, &d->ds_t, &d->ds_desc})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    d->ds_n = 0;
    const size_t row = (size_t)8 * d->cfg.bins, nd = (size_t)std::max(d->cfg.n_desc, 1);
    if (hipMalloc((void **)&d->ds_x, n * row * 4) != hipSuccess || hipMalloc((void **)&d->ds_t, n * row * 4) != hipSuccess ||
        hipMalloc((void **)&d->ds_desc, n * nd * 4) != hipSuccess)
        return dn_fail(d, PACH_ENOMEM, "dataset");
    DNCHK(d, hipMemcpy(d->ds_x, x, n * row * 4, hipMemcpyHostToDevice));
    DNCHK(d, hipMemcpy(d->ds_t, target, n * row * 4, hipMemcpyHostToDevice));
    if (d->cfg.n_desc > 0) DNCHK(d, hipMemcpy(d->ds_desc, desc, n * (size_t)d->cfg.n_desc * 4, hipMemcpyHostToDevice));
    d->ds_n = n;
    return PACH_OK;
}

extern "C" int pach_dn_train_step(pach_dn *d, const int64_t *idx, int batch, float lr, double *loss)
{
    if (!d || !idx) return PACH_EINVAL;
    if (int rc = check_batch(d, batch)) return rc;
    if (d->ds_n <= 0) return dn_fail(d, PACH_ESTATE, "no dataset");
    for (int i = 0; i < batch; ++i)
        if (idx[i] < 0 || idx[i] >= d->ds_n) return dn_fail(d, PACH_EINVAL, "dataset index out of range");
    (void)hipSetDevice(d->device);
    DNCHK(d, hipMemcpyAsync(d->idx, idx, (size_t)batch * 8, hipMemcpyHostToDevice, d->st));
    const size_t row = (size_t)8 * d->cfg.bins;
    hipLaunchKernelGGL(k_gather, dim3(nblk(batch * row)), dim3(256), 0, d->st, (const float *)d->ds_x,
                       (const int64_t *)d->idx, batch, row, d->x);
    hipLaunchKernelGGL(k_gather, dim3(nblk(batch * row)), dim3(256), 0, d->st, (const float *)d->ds_t,
                       (const int64_t *)d->idx, batch, row, d->tgt);
    if (d->cfg.n_desc > 0)
        hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)batch * d->cfg.n_desc)), dim3(256), 0, d->st,
                           (const float *)d->ds_desc, (const int64_t *)d->idx, batch, (size_t)d->cfg.n_desc, d->desc);
    DNCHK(d, hipGetLastError());
    if (int rc = loss_and_grad(d, batch, true, loss)) return rc;
    ++d->step;
    const float c1 = 1.0f / (1.0f - std::pow(d->cfg.beta1, (float)d->step));
    const float c2 = 1.0f / (1.0f - std::pow(d->cfg.beta2, (float)d->step));
    const size_t n = (size_t)d->n_params;
    hipLaunchKernelGGL(k_adamw, dim3(nblk(n)), dim3(256), 0, d->st, d->w, (const float *)d->g, d->m, d->v, d->wb, n,
                       lr, d->cfg.beta1, d->cfg.beta2, d->cfg.eps, d->cfg.weight_decay, c1, c2);
    DNCHK(d, hipGetLastError());
    return PACH_OK;
}

extern "C" int pach_dn_gemm(pach_dn *d, const uint16_t *a, const uint16_t *b, int M, int N, int K, int ta, int tb,
                            float *c)
{
    if (!d || !a || !b || !c || M <= 0 || N <= 0 || K <= 0) return PACH_EINVAL;
    (void)hipSetDevice(d->device);
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr;
    int rc = PACH_OK;
    if (hipMalloc((void **)&da, (size_t)M * K * 2) != hipSuccess || hipMalloc((void **)&db, (size_t)K * N * 2) != hipSuccess ||
        hipMalloc((void **)&dc, (size_t)M * N * 4) != hipSuccess) {
        rc = dn_fail(d, PACH_ENOMEM, "gemm test buffers");
    } else if (hipMemcpy(da, a, (size_t)M * K * 2, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(db, b, (size_t)K * N * 2, hipMemcpyHostToDevice) != hipSuccess) {
        rc = dn_fail(d, PACH_EHIP, "gemm test upload");
    } else {
        Epi e;
        e.out = dc;
        e.ldo = N;
        // split K when it is long, as the weight gradients do
        const int splits = (K >= 1024 && (size_t)M * N * 8 <= d->partial_n) ? 4 : 1;
        rc = gemm(d, ta != 0, tb != 0, da, ta ? M : K, db, tb ? K : N, M, N, K, e, splits);
        if (rc == PACH_OK) {
            if (hipMemcpyAsync(c, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost, d->st) != hipSuccess ||
                hipStreamSynchronize(d->st) != hipSuccess)
                rc = dn_fail(d, PACH_EHIP, "gemm test download");
        }
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dc) (void)hipFree(dc);
    return rc;
}
