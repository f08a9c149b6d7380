// image_source.hip -- image-source paths of planar objects (ImageSourceData,
// Classes_ImageSource.cs), validated on the GPU in FP64.  Built with
// -ffp-contract=off: every operation rounds as the CPU oracle's does.
//
// The reference walks the sequences depth-first on CPU threads
// (TraverseOrders :266-348) and, per sequence and receiver, checks the
// reflection path backwards from the receiver (ProcessImages :350-476).  Here
// each thread takes one (sequence, receiver) pair of one order: it mirrors the
// source through the sequence's planes, shoots from the receiver towards each
// image in turn (OcclusionIntersect :1350-1392) with the tracer's FP64 grid
// walk, applies the reference's final checks (both legs on one side of each
// plane, FinalOcclusion :1245-1281) and computes the path energy (planar
// Power_Recursion :535-803, B_Function_NoInterp / H_Function :859-1030,
// Specular_Path :1914-1982).  Valid paths are appended through one atomic
// counter; the host sorts them into the reference's order.
#include "image_source.h"
#include "trace_impl.h"

namespace pach {
namespace {

constexpr int kIsBlock = 256;
// transmissive surfaces a check passes before it gives up (as a miss); every
// pass advances the ray by more than 1e-10, so the limit only bounds the loop
constexpr int kIsPassLimit = 1 << 16;

__device__ __forceinline__ bool transmissive(const DevScene<double> &s, int p)
{
    bool t = false;
    for (int b = 0; b < 8; ++b) t |= s.trans[8 * p + b] > 0;
    return t;
}

// 1 - alpha of polygon p in band b (the material table holds (1 - alpha, s, lo, hi))
__device__ __forceinline__ double one_minus_alpha(const DevScene<double> &s, int p, int b)
{
    return s.mat[8 * p + b].x;
}

// closest hit; X = o + d t
__device__ __forceinline__ int hit(const DevScene<double> &s, const double o[3], const double d[3], double &t,
                                   double X[3])
{
    uint64_t rmask;
    const int tri = shoot<double, false>(s, o[0], o[1], o[2], d[0], d[1], d[2], -1, t, rmask);
    if (tri >= 0)
        for (int j = 0; j < 3; ++j) X[j] = o[j] + d[j] * t;
    return tri;
}

// Can a closest-hit query from o along d end on object px at t >= 1e-6?
// Only if the same intersection test accepts one of px's triangles there with
// no distance limit (the walk accepts a subset of those), so a "no" decides
// the check without walking the grid.
__device__ bool may_hit_object(const DevScene<double> &s, const IsJob &j, const double o[3], const double d[3],
                               int px)
{
    for (int k = j.obj_tri0[px]; k < j.obj_tri0[px + 1]; ++k) {
        double tt;
        if (tri_hit_mt(s, j.obj_tris[k], o[0], o[1], o[2], d[0], d[1], d[2], __builtin_huge_val(), tt) &&
            !(tt < 0.000001))
            return true;
    }
    return false;
}

// Conservative form of may_hit_object (entries [k_lo, k_hi) of the object's
// triangles) for the screening pass, without
// divisions: d = E - O unnormalised (barycentric coordinates do not depend on
// its length), the barycentric bounds widened by 1e-7 and near-parallel rays
// (|det| below 1e-6 of its terms) always kept.  It never rejects a triangle
// that may_hit_object accepts, whose rounding errors are ~1e-15 relative, so
// screening changes which items are validated but never a result.
__device__ bool may_hit_tris_wide(const DevScene<double> &s, const IsJob &j, const double o[3], const double d[3],
                                  int k_lo, int k_hi)
{
    constexpr double kEps = 1e-7;
    for (int k = k_lo; k < k_hi; ++k) {
        const int t = j.obj_tris[k];
        const double4 a = s.v0[t], e1 = s.e1[t], e2 = s.e2[t];
        const double px_ = d[1] * e2.z - d[2] * e2.y;
        const double py_ = d[2] * e2.x - d[0] * e2.z;
        const double pz_ = d[0] * e2.y - d[1] * e2.x;
        const double t0 = e1.x * px_, t1 = e1.y * py_, t2 = e1.z * pz_;
        const double det = t0 + t1 + t2;
        const double ad = fabs(det);
        if (ad <= 1e-6 * (fabs(t0) + fabs(t1) + fabs(t2))) return true;
        const double sg = det > 0 ? 1.0 : -1.0;
        const double sx = o[0] - a.x, sy = o[1] - a.y, sz = o[2] - a.z;
        const double uu = (sx * px_ + sy * py_ + sz * pz_) * sg;
        const double qx = sy * e1.z - sz * e1.y;
        const double qy = sz * e1.x - sx * e1.z;
        const double qz = sx * e1.y - sy * e1.x;
        const double vv = (d[0] * qx + d[1] * qy + d[2] * qz) * sg;
        const double tt = (e2.x * qx + e2.y * qy + e2.z * qz) * sg;
        if (uu >= -kEps * ad && vv >= -kEps * ad && uu + vv <= (1 + kEps) * ad && tt > 0) return true;
    }
    return false;
}

// OcclusionIntersect: from O towards the image E, through transmissive
// polygons of other objects, the first surface must belong to object px.
__device__ bool occl_intersect(const DevScene<double> &s, const IsJob &j, const double O[3], const double E[3],
                               int px, double tm[8], double Xout[3], int &poly)
{
    double D[3] = {E[0] - O[0], E[1] - O[1], E[2] - O[2]};
    const double L = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    D[0] /= L; D[1] /= L; D[2] /= L;
    // without transmissive polygons the first hit decides: it must lie on px
    if (!j.any_trans && !may_hit_object(s, j, O, D, px)) return false;
    double o[3] = {O[0], O[1], O[2]}, X[3], t = 0;
    int tri = -1;
    for (int guard = 0; guard < kIsPassLimit; ++guard) {
        tri = hit(s, o, D, t, X);
        if (tri >= 0) {
            const int p = s.surf[tri];
            if (transmissive(s, p) && px != j.obj[p]) {
                for (int b = 0; b < 8; ++b) tm[b] *= one_minus_alpha(s, p, b) * s.trans[8 * p + b];
                o[0] = X[0]; o[1] = X[1]; o[2] = X[2];
                continue;
            }
        }
        break;
    }
    // a miss is Hare's empty intersection event (t = 0)
    if (tri < 0 || t < 0.000001) return false;
    if (px != j.obj[s.surf[tri]]) return false;
    poly = s.surf[tri];
    Xout[0] = X[0]; Xout[1] = X[1]; Xout[2] = X[2];
    return true;
}

// FinalOcclusion: from the first reflection point back to the source.
__device__ bool final_occluded(const DevScene<double> &s, const IsJob &j, const double S[3], const double P1[3],
                               int px, double tm[8])
{
    double D[3] = {S[0] - P1[0], S[1] - P1[1], S[2] - P1[2]};
    const double L = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    double L2 = 0;
    D[0] /= L; D[1] /= L; D[2] /= L;
    double o[3] = {P1[0], P1[1], P1[2]}, X[3], t = 0;
    int tri = -1;
    for (int guard = 0; guard < kIsPassLimit; ++guard) {
        tri = hit(s, o, D, t, X);
        if (tri >= 0) {
            const int p = s.surf[tri];
            if (transmissive(s, p) && px != j.obj[p] && t < L) {
                for (int b = 0; b < 8; ++b) tm[b] *= one_minus_alpha(s, p, b) * s.trans[8 * p + b];
                o[0] = X[0]; o[1] = X[1]; o[2] = X[2];
                L2 += t;
                continue;
            }
        }
        break;
    }
    if (tri < 0) return false;
    L2 += t;
    return L2 < L;
}

__device__ __forceinline__ int sgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

__device__ __forceinline__ double seg_len(const double *a, const double *b)
{
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return sqrt(x * x + y * y + z * z);
}

// Item lo + l of the launch: receiver and sequence digits (l's digits added
// to those of lo, 32-bit arithmetic).
// x / d for x < 2^31 and 1 <= d < 2^21 through a double reciprocal
// (rinv = 1 / d): the estimate is off by at most one, then corrected.
__device__ __forceinline__ uint32_t div_small(uint32_t x, uint32_t d, double rinv)
{
    int64_t q = (int64_t)((double)x * rinv);
    int64_t r = (int64_t)x - q * (int64_t)d;
    if (r < 0) { --q; r += d; }
    if (r >= (int64_t)d) ++q;
    return (uint32_t)q;
}

template <int NO>
__device__ __forceinline__ int decode_item(const IsJob &j, uint32_t l, int seq[NO])
{
    const uint32_t R = (uint32_t)j.R, N = (uint32_t)j.n_obj;
    uint32_t rest = div_small(l, R, j.inv_R);
    int rec = j.base_rec + (int)(l - rest * R);
    int carry = 0;
    if (rec >= j.R) { rec -= j.R; carry = 1; }
#pragma unroll
    for (int q = NO - 1; q >= 0; --q) {
        const uint32_t next = div_small(rest, N, j.inv_N);
        int v = j.base_seq[q] + (int)(rest - next * N) + carry;
        rest = next;
        carry = 0;
        if (v >= j.n_obj) { v -= j.n_obj; carry = 1; }
        seq[q] = v;
    }
    return rec;
}

// Polygon_Scene.Image: each image mirrors the previous one in the next plane.
template <int NO>
__device__ __forceinline__ void make_images(const IsJob &j, const int seq[NO], double img[NO][3])
{
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const double *prev = q > 0 ? img[q - 1] : j.src;
        const IsPlane &pl = j.plane[seq[q]];
        const double dist = pl.n[0] * (prev[0] - pl.p[0]) + pl.n[1] * (prev[1] - pl.p[1]) +
                            pl.n[2] * (prev[2] - pl.p[2]);
        for (int k = 0; k < 3; ++k) img[q][k] = prev[k] - 2 * pl.n[k] * dist;
    }
}

// Screening pass (scenes without transmissive polygons): an item survives
// only if the ray from its receiver towards the last image can end on the
// last object (may_hit_tris_wide).  Survivors gather in a workgroup buffer in
// LDS and go to the global queue in bulk (one atomic per flush), so the
// validation pass runs on full waves and the queue counter stays cold.
constexpr int kScreenBuf = 4096;

template <int NO>
__global__ void __launch_bounds__(kIsBlock) k_is_screen(DevScene<double> s, IsJob j, uint32_t count,
                                                       uint32_t *queue, unsigned *queued)
{
    constexpr int n = NO;
    __shared__ uint32_t buf[kScreenBuf];
    __shared__ unsigned nbuf, gbase;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) nbuf = 0;
    __syncthreads();
    auto flush = [&]() {
        if (threadIdx.x == 0) gbase = atomicAdd(queued, nbuf);
        __syncthreads();
        for (unsigned i = threadIdx.x; i < nbuf; i += kIsBlock) queue[gbase + i] = buf[i];
        __syncthreads();
        if (threadIdx.x == 0) nbuf = 0;
        __syncthreads();
    };
    for (uint32_t l0 = blockIdx.x * kIsBlock; l0 < count; l0 += gridDim.x * kIsBlock) {
        const uint32_t l = l0 + threadIdx.x;
        bool pass = false;
        if (l < count) {
            int seq[NO];
            const int rec = decode_item<NO>(j, l, seq);
            double img[NO][3];
            make_images<NO>(j, seq, img);
            const double4 rc = j.rec[rec];
            const double D[3] = {img[n - 1][0] - rc.x, img[n - 1][1] - rc.y, img[n - 1][2] - rc.z};
            const double O[3] = {rc.x, rc.y, rc.z};
            pass = may_hit_tris_wide(s, j, O, D, j.obj_tri0[seq[n - 1]], j.obj_tri0[seq[n - 1] + 1]);
        }
        const uint64_t m = __ballot(pass);
        if (m != 0) {
            unsigned base = 0;
            if (lane == 0) base = atomicAdd(&nbuf, (unsigned)__popcll(m));
            base = __shfl(base, 0);
            if (pass) buf[base + __popcll(m & ((1ull << lane) - 1))] = l;
        }
        __syncthreads();
        if (nbuf > kScreenBuf - kIsBlock) flush();
    }
    flush();
}

// Validation pass: ProcessImages for each queued item (every item when queue
// is null).
template <int NO>
__global__ void __launch_bounds__(kIsBlock) k_image_sources(DevScene<double> s, IsJob j, uint32_t count,
                                                           const uint32_t *queue, const unsigned *queued)
{
    constexpr int n = NO;
    const uint32_t total = queue ? *queued : count;
    for (uint32_t qi = blockIdx.x * kIsBlock + threadIdx.x; qi < total; qi += gridDim.x * kIsBlock) {
        const uint32_t l = queue ? queue[qi] : qi;
        int seq[NO];
        const int rec = decode_item<NO>(j, l, seq);
        double img[NO][3];
        make_images<NO>(j, seq, img);
        const double4 rc = j.rec[rec];
        double P[NO + 2][3];
        int polys[NO];
        double tm[8] = {1, 1, 1, 1, 1, 1, 1, 1};
        P[0][0] = j.src[0]; P[0][1] = j.src[1]; P[0][2] = j.src[2];
        P[n + 1][0] = rc.x; P[n + 1][1] = rc.y; P[n + 1][2] = rc.z;
        bool ok = true;
        for (int q = n - 1; q >= 0 && ok; --q) ok = occl_intersect(s, j, P[q + 2], img[q], seq[q], tm, P[q + 1], polys[q]);
        if (!ok) continue;
        // both legs of every reflection on one side of its plane, then the
        // first leg open
        for (int i = 0; i < n && ok; ++i) {
            double d1[3], d2[3];
            for (int k = 0; k < 3; ++k) {
                d1[k] = P[i][k] - P[i + 1][k];
                d2[k] = P[i + 2][k] - P[i + 1][k];
            }
            const double l1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
            d1[0] /= l1; d1[1] /= l1; d1[2] /= l1;
            const double l2 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
            d2[0] /= l2; d2[1] /= l2; d2[2] /= l2;
            const double *nr = j.plane[seq[i]].n;
            if (sgn(nr[0] * d1[0] + nr[1] * d1[1] + nr[2] * d1[2]) != sgn(nr[0] * d2[0] + nr[1] * d2[1] + nr[2] * d2[2]))
                ok = false;
            else if (i == n - 1 && final_occluded(s, j, P[0], P[1], seq[0], tm))
                ok = false;
        }
        if (!ok) continue;
        for (int b = 0; b < 8; ++b) tm[b] *= j.power[b];
        // planar Power_Recursion: curvature 1 / first leg, projected over the rest
        double kv = 1.0 / seg_len(P[1], P[0]);
        const double l0 = seg_len(P[0], P[1]);
        const double t0 = l0 / j.c, t1 = l0 / j.c;
        for (int o = 0; o < n; ++o) {
            const double l = seg_len(P[o + 1], P[o + 2]);
            kv = (kv == __builtin_huge_val()) ? 0 : 1 / kv;
            kv += l;
            kv = (kv == __builtin_huge_val()) ? 0 : 1 / kv;
        }
        const double h = sqrt(fabs((1 / (4 * M_PI)) * (kv * kv - 0.0 * 0.0)));
        const double p = fabs(h);
        double length = 0;
        for (int q = 1; q < n + 2; ++q) {
            const double x = P[q][0] - P[q - 1][0], y = P[q][1] - P[q - 1][1], z = P[q][2] - P[q - 1][2];
            length += sqrt(x * x + y * y + z * z);
        }
        double e[8];
        for (int b = 0; b < 8; ++b) {
            const double mod = sqrt(tm[b] * pow(10.0, -.05 * j.atten[b] * j.c * 0.5 * (t0 + t1)));
            const double B = p * mod;
            const double H = B * sqrt(j.power[b] * 411.6);
            const double Hrc = H * H / j.rho_c;
            e[b] = pow(10.0, -.1 * j.atten[b] * length) * Hrc;
            e[b] *= tm[b];
        }
        for (int q = 0; q < n; ++q) {
            const int pq = polys[q];
            for (int b = 0; b < 8; ++b) {
                const double4 m = s.mat[8 * pq + b];
                e[b] *= m.x * (1 - m.y) * (1 - s.trans[8 * pq + b]);
                if (e[b] == 0) e[b] = 1E-22;
            }
        }
        const unsigned long long slot = atomicAdd(j.count, 1ull);
        if (slot >= j.cap) continue;
        pach_is_path &o = j.out[slot];
        o.receiver = rec;
        o.order = n;
        for (int q = 0; q < PACH_IS_MAX_ORDER; ++q) {
            o.seq[q] = q < n ? seq[q] : 0;
            o.poly[q] = q < n ? polys[q] : 0;
        }
        for (int q = 0; q < PACH_IS_MAX_ORDER + 2; ++q)
            for (int k = 0; k < 3; ++k) o.pts[q][k] = q < n + 2 ? P[q][k] : 0.0;
        o.length = length;
        o.time_s = length / j.c;
        for (int b = 0; b < 8; ++b) o.energy[b] = e[b];
    }
}

}  // namespace

hipError_t launch_image_sources(hipStream_t st, int grid, const DevScene<double> &s, const IsJob &j, uint32_t count,
                                uint32_t *queue, unsigned *queued)
{
    if (count == 0) return hipSuccess;
#define PACH_IS_CASE(NO_)                                                                                   \
    case NO_:                                                                                               \
        if (queue) {                                                                                        \
            hipLaunchKernelGGL(k_is_screen<NO_>, dim3(grid), dim3(kIsBlock), 0, st, s, j, count, queue, queued); \
            if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;                                \
        }                                                                                                   \
        hipLaunchKernelGGL(k_image_sources<NO_>, dim3(grid), dim3(kIsBlock), 0, st, s, j, count,            \
                           (const uint32_t *)queue, (const unsigned *)queued);                              \
        break;
    switch (j.order) {
        PACH_IS_CASE(1) PACH_IS_CASE(2) PACH_IS_CASE(3) PACH_IS_CASE(4)
        PACH_IS_CASE(5) PACH_IS_CASE(6) PACH_IS_CASE(7) PACH_IS_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef PACH_IS_CASE
    return hipGetLastError();
}

}  // namespace pach
