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

// OcclusionIntersect: from O towards the image E, through transmissive
// polygons of other objects, the first surface must belong to object px.
__device__ bool occl_intersect(const DevScene<double> &s, const IsJob &j, const double O[3], const double E[3],
                               int px, double tm[8], double Xout[3], int &poly)
{
    double D[3] = {E[0] - O[0], E[1] - O[1], E[2] - O[2]};
    const double L = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    D[0] /= L; D[1] /= L; D[2] /= L;
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

__global__ void __launch_bounds__(kIsBlock) k_image_sources(DevScene<double> s, IsJob j, int64_t lo, int64_t hi)
{
    const int n = j.order;
    for (int64_t idx = lo + (int64_t)blockIdx.x * kIsBlock + threadIdx.x; idx < hi;
         idx += (int64_t)gridDim.x * kIsBlock) {
        const int rec = (int)(idx % j.R);
        int64_t sq = idx / j.R;
        int seq[PACH_IS_MAX_ORDER];
        for (int q = n - 1; q >= 0; --q) {
            seq[q] = (int)(sq % j.n_obj);
            sq /= j.n_obj;
        }
        // Polygon_Scene.Image: each image mirrors the previous one
        double img[PACH_IS_MAX_ORDER][3];
        for (int q = 0; q < n; ++q) {
            const double *prev = q > 0 ? img[q - 1] : j.src;
            const IsPlane &pl = j.plane[seq[q]];
            const double dist = pl.n[0] * (prev[0] - pl.p[0]) + pl.n[1] * (prev[1] - pl.p[1]) +
                                pl.n[2] * (prev[2] - pl.p[2]);
            for (int k = 0; k < 3; ++k) img[q][k] = prev[k] - 2 * pl.n[k] * dist;
        }
        const double4 rc = j.rec[rec];
        double P[PACH_IS_MAX_ORDER + 2][3];
        int polys[PACH_IS_MAX_ORDER];
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

hipError_t launch_image_sources(hipStream_t st, int grid, const DevScene<double> &s, const IsJob &j, int64_t lo,
                                int64_t hi)
{
    if (hi <= lo) return hipSuccess;
    hipLaunchKernelGGL(k_image_sources, dim3(grid), dim3(kIsBlock), 0, st, s, j, lo, hi);
    return hipGetLastError();
}

}  // namespace pach
