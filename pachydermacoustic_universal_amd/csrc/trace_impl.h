// trace_impl.h -- gfx950 kernels of the stochastic split ray tracer.
//
// Included by trace_f32.hip (FP32 geometry, the throughput path) and
// trace_f64.hip (FP64 geometry, built with -ffp-contract=off, the path the
// deterministic per-bin parity runs on).  Each file instantiates one precision.
//
// Reference behaviour restated here (TaskAll = "Classes_Complete_Trace - TaskAll.cs"):
//   broadband phase        TaskAll:360-424
//   split phase            TaskAll:427-480
//   End_Trace              TaskAll:489-536
//   closest hit            Polygon_Scene.shoot (Classes_Scenes.cs:777-808) over a
//                          voxel DDA (Classes_VoxelGrid_PolyRefractive.cs:84-313)
//   receivers              Spherical_Receiver.CheckRay/CheckBroadbandRay
//                          (Classes_Receivers.cs:803-877), Directional_Histogram.Add (:1813-1825)
//   absorption/scattering  Basic_Material.Absorb (Classes_Materials.cs:155-190),
//                          Lambert_Scattering (Classes_Materials.cs:4463-4702)
//   sources                GeodesicSource/RandomSource.Directions (Classes_Sources.cs:256-403)
//
// Randomness: every ray owns a Philox4x32-10 stream keyed by the trace seed with
// counter (block, uid_lo, uid_hi, source ray).  A spawned ray's uid is derived
// from its parent's uid and the spawning event, so results do not depend on
// which lane or wave traces which ray.
#pragma once

#include "trace_types.h"

#include <type_traits>

namespace pach {

// ---------------------------------------------------------------- RNG
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per pair (a single v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t child_uid(uint64_t parent, uint64_t tag)
{
    return mix64(parent + 0x9E3779B97F4A7C15ull * (tag + 1ull));
}

#define PACH_TAG_EARLY(k, o) ((1ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define PACH_TAG_EARLY_TR(k, o) ((2ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define PACH_TAG_EARLY_TD(k, o) ((3ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define PACH_TAG_SPLIT(o) ((4ull << 40) | (uint64_t)(o))
#define PACH_TAG_LATE(k) ((5ull << 40) | (uint64_t)(k))

template <class R> __device__ __forceinline__ R unif(uint32_t w);
template <> __device__ __forceinline__ double unif<double>(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }
template <> __device__ __forceinline__ float unif<float>(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- math helpers
template <class R> __device__ __forceinline__ R rcp(R x);
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double rcp<double>(double x) { return 1.0 / x; }

// a * b + c: one fused operation in FP32; in FP64 two separately rounded
// operations (the FP64 units are built without contraction, as the oracle is).
template <class R> __device__ __forceinline__ R mad(R a, R b, R c);
template <> __device__ __forceinline__ float mad<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double mad<double>(double a, double b, double c) { return a * b + c; }

// FP64: sqrt and IEEE division, as the oracle; FP32: one reciprocal square root.
template <class R> __device__ __forceinline__ void normalize3(R &x, R &y, R &z)
{
    if constexpr (sizeof(R) == 8) {
        R l = sqrt(x * x + y * y + z * z);
        if (l > R(0)) { x /= l; y /= l; z /= l; }
    } else {
        const float q = x * x + y * y + z * z;
        if (q > 0.0f) {
            const float il = __builtin_amdgcn_rsqf(q);
            x *= il; y *= il; z *= il;
        }
    }
}

// Square root: FP64 IEEE (as the oracle); FP32 the hardware v_sqrt_f32
// (1 ulp), where the FP32 tracer has no bit-level reference to match.
template <class R> __device__ __forceinline__ R fsqrt(R x);
template <> __device__ __forceinline__ double fsqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float fsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// Smallest accepted hit distance.  FP64 keeps Hare's t > 1e-10
// (Classes_VoxelGrid_PolyRefractive.cs:166).  FP32 instead excludes the
// triangle the ray leaves (planar polygons cannot be re-hit) and accepts t > 0.
template <class R> struct Tmin;
template <> struct Tmin<double> { static constexpr double v = 1e-10; static constexpr bool excl = false; };
template <> struct Tmin<float> { static constexpr float v = 0.0f; static constexpr bool excl = true; };

// ---------------------------------------------------------------- closest hit
// Moller-Trumbore with inclusive edges (the FP64 path: the same arithmetic as
// the CPU oracle, so both accept and reject the same candidates).
__device__ __forceinline__ bool tri_hit_mt(const DevScene<double> &s, int t, double ox, double oy, double oz,
                                           double dx, double dy, double dz, double tlim, double &tout)
{
    // Moller-Trumbore without early returns (the rejections are combined at the
    // end), so a wave whose lanes fail at different stages runs one stream;
    // every value that decides is computed by the same operations as in the
    // oracle, so the decisions are bit-identical.
    const double4 a = s.v0[t];
    const double4 e1 = s.e1[t];
    const double4 e2 = s.e2[t];
    const double px = dy * e2.z - dz * e2.y;
    const double py = dz * e2.x - dx * e2.z;
    const double pz = dx * e2.y - dy * e2.x;
    const double det = e1.x * px + e1.y * py + e1.z * pz;
    const double inv = 1.0 / det;  // infinite for det == 0, which is rejected below
    const double sx = ox - a.x, sy = oy - a.y, sz = oz - a.z;
    const double u = (sx * px + sy * py + sz * pz) * inv;
    const double qx = sy * e1.z - sz * e1.y;
    const double qy = sz * e1.x - sx * e1.z;
    const double qz = sx * e1.y - sy * e1.x;
    const double v = (dx * qx + dy * qy + dz * qz) * inv;
    const double tt = (e2.x * qx + e2.y * qy + e2.z * qz) * inv;
    const bool ok = det != 0.0 && !(u < 0.0 || u > 1.0) && !(v < 0.0 || u + v > 1.0) && (tt > 1e-10) &&
                    (tt <= tlim);
    tout = ok ? tt : tout;
    return ok;
}

// FP32 test against the triangle's dual basis (host-precomputed rows r0, r1,
// r2): r2 . x + r2.w is the signed plane distance, r0 / r1 give the barycentric
// coordinates u, v of a point x of the plane.  The rows are read from the
// voxel list's copy at list index j (load_rows); the test has no early return
// and combines its conditions without short-circuiting, so lanes whose
// candidates fail at different stages run one instruction stream and no row
// load is deferred behind a condition.  Edges are widened by 4e-6
// (barycentric) so no ray slips between neighbours.
struct Rows {
    float4 r2, r0, r1;
};

__device__ __forceinline__ Rows load_rows(const DevScene<float> &s, uint32_t j)
{
    const float4 *rows = s.cell_rows + 3 * (size_t)j;
    return Rows{rows[0], rows[1], rows[2]};
}

// LIM false: no limit on t (the caller compares t with its nearest hit
// itself), and tout is set whether or not the ray hits.
template <bool LIM = true>
__device__ __forceinline__ bool tri_hit_rows(const Rows &q, float ox, float oy, float oz, float dx, float dy,
                                             float dz, float tlim, float &tout)
{
    const float4 r2 = q.r2, r0 = q.r0, r1 = q.r1;
    const float den = r2.x * dx + r2.y * dy + r2.z * dz;
    const float num = -(r2.x * ox + r2.y * oy + r2.z * oz + r2.w);
    const float tt = num * __builtin_amdgcn_rcpf(den);
    const float px = ox + dx * tt, py = oy + dy * tt, pz = oz + dz * tt;
    const float u = r0.x * px + r0.y * py + r0.z * pz + r0.w;
    const float v = r1.x * px + r1.y * py + r1.z * pz + r1.w;
    const bool ok = (tt > 0.0f) & (!LIM || tt <= tlim) & !(u < -4e-6f) & !(u > 1.000004f) & !(v < -4e-6f) &
                    !(u + v > 1.000004f);
    if constexpr (LIM) tout = ok ? tt : tout;
    else tout = tt;
    return ok;
}

__device__ __forceinline__ bool tri_hit_dual(const DevScene<float> &s, uint32_t j, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tlim, float &tout)
{
    return tri_hit_rows(load_rows(s, j), ox, oy, oz, dx, dy, dz, tlim, tout);
}

// Triangle t, entry j of the voxel lists.
template <class R>
__device__ __forceinline__ bool tri_hit(const DevScene<R> &s, int t, uint32_t j, R ox, R oy, R oz, R dx, R dy,
                                        R dz, R tlim, R &tout)
{
    if constexpr (sizeof(R) == 8) return tri_hit_mt(s, t, ox, oy, oz, dx, dy, dz, tlim, tout);
    else return tri_hit_dual(s, j, ox, oy, oz, dx, dy, dz, tlim, tout);
}

// Conservative FP32 screen of list entry j for the mixed walk (the FP64
// decisions come from tri_hit_mt).  The screen passes every triangle the FP64
// test could accept with a distance t <= tlim: tt is the FP32 distance, e a
// bound on its error, and the barycentric test is widened by its own error
// bound; both grow with 1 / |n . d| (grazing rays).  Every product is rounded
// once (fused), which the bounds assume.  A ray parallel to the plane in FP32
// (n . d == 0) is flagged in `para`: FP32 cannot place its hit.
__device__ __forceinline__ bool tri_filter(const Rows &q, const TraceConst &tc, float ox, float oy, float oz,
                                           float dx, float dy, float dz, float tlim, float &tt, float &e, bool &para)
{
    const float4 r2 = q.r2, r0 = q.r0, r1 = q.r1;
    const float den = __builtin_fmaf(r2.x, dx, __builtin_fmaf(r2.y, dy, r2.z * dz));
    const float num = -__builtin_fmaf(r2.x, ox, __builtin_fmaf(r2.y, oy, __builtin_fmaf(r2.z, oz, r2.w)));
    const float inv = __builtin_amdgcn_rcpf(den);
    tt = num * inv;
    const float ai = fabsf(inv);
    e = __builtin_fmaf(tc.mx_kt, ai, tc.mx_kt0);
    const float eu = __builtin_fmaf(tc.mx_ku, ai, tc.mx_ku0);
    const float px = __builtin_fmaf(dx, tt, ox), py = __builtin_fmaf(dy, tt, oy), pz = __builtin_fmaf(dz, tt, oz);
    const float u = __builtin_fmaf(r0.x, px, __builtin_fmaf(r0.y, py, __builtin_fmaf(r0.z, pz, r0.w)));
    const float v = __builtin_fmaf(r1.x, px, __builtin_fmaf(r1.y, py, __builtin_fmaf(r1.z, pz, r1.w)));
    para = den == 0.0f;
    // non-short-circuit &, so no row load is sunk behind the distance test
    return (tt + e > 0.0f) & (tt - e <= tlim) & !(u < -eu) & !(v < -eu) & !(u + v > 1.0f + 2.0f * eu);
}

// Voxel distances staged in LDS by k_octave (dynamic shared memory).
extern __shared__ uint8_t pach_lds_dist[];

template <bool LDS>
__device__ __forceinline__ uint8_t cell_dist(const uint8_t *g, int cell)
{
    if constexpr (LDS) return pach_lds_dist[cell];
    else return g[cell];
}

// Resumable closest-hit walk over the uniform grid (Amanatides-Woo), with the
// reference's rule that a hit ends the walk once it lies inside the current
// voxel (Classes_VoxelGrid_PolyRefractive.cs:172).  Empty space: a voxel at
// L-inf distance k >= 2 from every occupied voxel lies inside a (2k-1)^3 empty
// block, crossed in one jump (walk_steps); the flattened walks jump across the
// largest empty cube cornered at the voxel on the ray's side (cell_oct), which
// contains that block's part ahead of the ray and reaches further from walls.
//
// Each triangle is tested once per walk where the reference keeps a per-ray
// mailbox (Classes_VoxelGrid_PolyRefractive.cs:162-164): the flattened walk
// keeps the nearest hit found so far wherever it lies (not only hits inside
// the current voxel), and a voxel entered by one step through face f skips
// the list entries its neighbour across f also lists (DevScene::cell_skip),
// which the walk tested there.  After a jump, and in the first voxel, every
// entry is tested.
template <class R> struct Walk {
    int cx, cy, cz;
    R tmx, tmy, tmz;  // t where the ray leaves the current voxel along each axis
    // exit t of voxel c along axis a: c * sa + qa (sa = cs_a / d_a, one voxel
    // width in t with the direction's sign; sa = 0 and qa = +inf for d_a == 0,
    // so the same fused multiply-add gives the infinite exit)
    R sx, sy, sz;
    R qx, qy, qz;
    int best;
    R tbest;
    uint64_t rmask;   // receivers whose spheres touch a voxel the walk crossed
    uint32_t j, je;   // flattened walk: list window [j, j + 32) of the current voxel, list end je
    uint32_t todo;    // flattened walk: entries of the window still to test (bit i: entry j + i)
    int from;         // face of the current voxel the walk stepped in through (-1: none)
    int k;            // flattened walk: the current move's box edges minus one along x, y, z in
                      // bytes 0-2 (0 in an occupied voxel); -1 before the first voxel
    // mixed walk only (unused, and removed by the compiler, elsewhere)
    R eb;             // error bound of tbest
    R sl, sl2;        // least and second least lower bound t - e of the other candidates (-inf: undecidable)
    int sid;          // the candidate of sl
    int zid;          // a candidate whose interval reaches t <= 0 (a neighbour at the departure point)
    R tend;           // current voxel: the walk ends once tbest + eb <= tend
};

// Slope and offset of the voxel exits along one axis: voxel c is left at
// t = c * sa + qa, through its upper face when d > 0 and its lower one when
// d < 0, never when d == 0 (sa = 0, qa = +inf).
template <class R>
__device__ __forceinline__ void axis_line(R o, R d, const R gmin, const R cs, R &sa, R &qa)
{
    const R INF = R(__builtin_huge_val());
    const R inv = rcp<R>(d);
    sa = d != R(0) ? cs * inv : R(0);
    qa = d != R(0) ? ((d > R(0) ? cs : R(0)) + gmin - o) * inv : INF;
}

// t where the ray leaves voxel c along the axis of (sa, qa)
template <class R> __device__ __forceinline__ R axis_exit(R sa, R qa, int c) { return mad<R>(R(c), sa, qa); }

template <class R>
__device__ __forceinline__ void walk_init(const DevScene<R> &s, Walk<R> &w, R ox, R oy, R oz, R dx, R dy, R dz)
{
    const R INF = R(__builtin_huge_val());
    axis_line<R>(ox, dx, s.gmin[0], s.cs[0], w.sx, w.qx);
    axis_line<R>(oy, dy, s.gmin[1], s.cs[1], w.sy, w.qy);
    axis_line<R>(oz, dz, s.gmin[2], s.cs[2], w.sz, w.qz);
    w.rmask = 0;
    w.cx = min(max((int)floor((ox - s.gmin[0]) * s.inv_cs[0]), 0), s.gx - 1);
    w.cy = min(max((int)floor((oy - s.gmin[1]) * s.inv_cs[1]), 0), s.gy - 1);
    w.cz = min(max((int)floor((oz - s.gmin[2]) * s.inv_cs[2]), 0), s.gz - 1);
    w.tmx = axis_exit<R>(w.sx, w.qx, w.cx);
    w.tmy = axis_exit<R>(w.sy, w.qy, w.cy);
    w.tmz = axis_exit<R>(w.sz, w.qz, w.cz);
    w.best = -1;
    w.tbest = INF;
    w.j = w.je = 0;
    w.todo = 0;
    w.from = -1;
    w.k = -1;  // current voxel not loaded yet
}

// Processes up to `budget` voxels; returns true once the walk has ended
// (w.best >= 0: hit at w.tbest; w.best < 0: the ray left the grid).
template <class R, bool LDS>
__device__ bool walk_steps(const DevScene<R> &s, Walk<R> &w, R ox, R oy, R oz, R dx, R dy, R dz, int excl,
                           int budget, uint32_t &n_vox, uint32_t &n_test)
{
    for (int it = 0; it < budget; ++it) {
        ++n_vox;
        const int cell = (w.cz * s.gy + w.cy) * s.gx + w.cx;
        const int k = cell_dist<LDS>(s.cell_dist, cell);
        const R texit = fmin(w.tmx, fmin(w.tmy, w.tmz));
        if (k == 0) {
            if (s.cell_rmask) w.rmask |= s.cell_rmask[cell];
            const uint2 rg = s.cell_range[cell];
            // accept hits up to just past this voxel's exit; nearer voxels were done
            const R tlim = fmin(w.tbest, texit * R(1.0000001) + R(1e-6));
            n_test += rg.y;
            for (uint32_t j = rg.x; j < rg.x + rg.y; ++j) {
                const int t = (int)s.cell_tris[j];
                R tt;
                if (t != w.best && !(Tmin<R>::excl && t == excl) &&
                    tri_hit<R>(s, t, j, ox, oy, oz, dx, dy, dz, fmin(tlim, w.tbest), tt) &&
                    (tt < w.tbest || (tt == w.tbest && t < w.best))) {
                    w.tbest = tt;
                    w.best = t;
                }
            }
            if (w.best >= 0 && w.tbest <= texit) return true;
        }
        if (k >= 2) {
            // jump across the empty block around the current voxel
            const int h = k - 1;
            const int lx = max(w.cx - h, 0), hx = min(w.cx + h, s.gx - 1);
            const int ly = max(w.cy - h, 0), hy = min(w.cy + h, s.gy - 1);
            const int lz = max(w.cz - h, 0), hz = min(w.cz + h, s.gz - 1);
            const R bx = axis_exit<R>(w.sx, w.qx, dx > R(0) ? hx : lx);
            const R by = axis_exit<R>(w.sy, w.qy, dy > R(0) ? hy : ly);
            const R bz = axis_exit<R>(w.sz, w.qz, dz > R(0) ? hz : lz);
            const int ax = (bx < by) ? (bx < bz ? 0 : 2) : (by < bz ? 1 : 2);
            const R tb = ax == 0 ? bx : (ax == 1 ? by : bz);
            const R px = ox + dx * tb, py = oy + dy * tb, pz = oz + dz * tb;
            const int fx = (int)floor((px - s.gmin[0]) * s.inv_cs[0]);
            const int fy = (int)floor((py - s.gmin[1]) * s.inv_cs[1]);
            const int fz = (int)floor((pz - s.gmin[2]) * s.inv_cs[2]);
            // never move backwards; stay inside the block on the other axes
            int nx = dx > R(0) ? min(max(fx, w.cx), hx) : (dx < R(0) ? max(min(fx, w.cx), lx) : w.cx);
            int ny = dy > R(0) ? min(max(fy, w.cy), hy) : (dy < R(0) ? max(min(fy, w.cy), ly) : w.cy);
            int nz = dz > R(0) ? min(max(fz, w.cz), hz) : (dz < R(0) ? max(min(fz, w.cz), lz) : w.cz);
            if (ax == 0) nx = dx > R(0) ? hx + 1 : lx - 1;
            else if (ax == 1) ny = dy > R(0) ? hy + 1 : ly - 1;
            else nz = dz > R(0) ? hz + 1 : lz - 1;
            if (nx < 0 || nx >= s.gx || ny < 0 || ny >= s.gy || nz < 0 || nz >= s.gz) return true;
            w.cx = nx; w.cy = ny; w.cz = nz;
            w.tmx = axis_exit<R>(w.sx, w.qx, nx);
            w.tmy = axis_exit<R>(w.sy, w.qy, ny);
            w.tmz = axis_exit<R>(w.sz, w.qz, nz);
        } else {
            // one voxel step along the axis whose boundary comes first
            if (w.tmx < w.tmy && w.tmx < w.tmz) {
                w.cx += dx > R(0) ? 1 : -1;
                if (w.cx < 0 || w.cx >= s.gx) return true;
                w.tmx += fabs(w.sx);
            } else if (w.tmy < w.tmz) {
                w.cy += dy > R(0) ? 1 : -1;
                if (w.cy < 0 || w.cy >= s.gy) return true;
                w.tmy += fabs(w.sy);
            } else {
                w.cz += dz > R(0) ? 1 : -1;
                if (w.cz < 0 || w.cz >= s.gz) return true;
                w.tmz += fabs(w.sz);
            }
        }
    }
    return false;
}

// Enter the walk's current voxel: for an occupied voxel its triangle-list
// range and receiver mask, for an empty one the empty box ahead of the ray in
// its octant (the octant cube lengthened along two axes, grid.h octant_boxes).
template <class R, bool LDS>
__device__ __forceinline__ void walk_enter(const DevScene<R> &s, Walk<R> &w, uint32_t &n_vox)
{
    ++n_vox;
    const int cell = (w.cz * s.gy + w.cy) * s.gx + w.cx;
    const CellRec *rec = s.cell_rec + cell;  // every field below lies on this record's cache line
    // every load is issued before any is used, occupied voxel or not, so a
    // voxel costs one memory latency
    const uint4 head = *(const uint4 *)&rec->first;  // first, count, skip[0], skip[1]
    const uint4 tail = *(const uint4 *)&rec->rmask;  // receiver mask (2 words), octant edges (2 words)
    int k;
    if constexpr (LDS) k = pach_lds_dist[cell];
    else k = (int)rec->dist;
    // face mask of the entry face, loaded whether there is one or not (a
    // conditional load would make the wave wait for all of them there)
    const uint32_t seen = rec->skip[max(w.from, 0)] & (w.from >= 0 ? ~0u : 0u);
    // the edge of the empty cube ahead of the ray, in its octant (bit a set
    // when the ray runs down axis a)
    int o;
    if constexpr (sizeof(R) == 4)
        o = (int)((__float_as_uint(w.sx) >> 31) | ((__float_as_uint(w.sy) >> 30) & 2u) |
                  ((__float_as_uint(w.sz) >> 29) & 4u));
    else
        o = (w.sx < R(0) ? 1 : 0) | (w.sy < R(0) ? 2 : 0) | (w.sz < R(0) ? 4 : 0);
    // an empty voxel's box in this octant (grid.h octant_boxes): its first
    // axis length in head.xy, its second axis length in head.zw, the two axes
    // in tail.xy, one byte per octant; the cube's edge in tail.zw
    const bool hi4 = o >= 4;
    const uint32_t sh = 8u * (uint32_t)(o & 3);
    const uint32_t la = ((hi4 ? head.y : head.x) >> sh) & 0xffu;
    const uint32_t lb = ((hi4 ? head.w : head.z) >> sh) & 0xffu;
    const uint32_t axes = ((hi4 ? tail.y : tail.x) >> sh) & 0xffu;
    const uint32_t edge = ((hi4 ? tail.w : tail.z) >> sh) & 0xffu;
    const uint32_t a0 = axes & 3u, a1 = (axes >> 2) & 3u;
    const bool occ = k == 0;
    // the box edges minus one per axis: 0 in an occupied voxel (one step)
    w.k = occ ? 0 : (int)((edge - 1u) * 0x010101u + ((la - edge) << (8u * a0)) + ((lb - edge) << (8u * a1)));
    if (s.cell_rmask) w.rmask |= occ ? (((uint64_t)tail.y << 32) | tail.x) : 0ull;
    w.j = occ ? head.x : 0u;
    w.je = occ ? head.x + head.y : 0u;
    const uint32_t all = head.y >= 32 ? ~0u : (1u << head.y) - 1u;
    w.todo = occ ? all & ~seen : 0u;
    // every word is needed on every path: without this the compiler moves the
    // loads of the words only an empty voxel uses behind the occupancy test,
    // a second round trip
    asm volatile("" ::"v"(head.z), "v"(head.w), "v"(tail.z), "v"(tail.w));
}

// One move of the flattened walk: across the empty box ahead of the ray
// (edges from w.k, cornered at the voxel and extending along the ray's
// octant), else one voxel (the box is the voxel itself), by the same
// straight-line code for both so stepping and jumping lanes do not diverge.
// Returns true when the walk leaves the grid.
template <class R>
__device__ __forceinline__ bool walk_move(const DevScene<R> &s, Walk<R> &w, R ox, R oy, R oz, R dx, R dy, R dz)
{
    // voxels the box reaches past the current one per axis
    const int hx = w.k & 0xff, hy = (w.k >> 8) & 0xff, hz = (w.k >> 16) & 0xff;
    // the box's far faces lie h voxel widths past the current voxel's exits
    // (an axis the ray does not advance along keeps its infinite exit); a box
    // reaching past the grid is not clipped: the ray then leaves the grid
    // inside the box, and the cell reached below lies outside it
    const R bx = mad<R>(R(hx), fabs(w.sx), w.tmx);
    const R by = mad<R>(R(hy), fabs(w.sy), w.tmy);
    const R bz = mad<R>(R(hz), fabs(w.sz), w.tmz);
    const int ax = (bx < by) ? (bx < bz ? 0 : 2) : (by < bz ? 1 : 2);
    const R tb = ax == 0 ? bx : (ax == 1 ? by : bz);
    const R px = mad<R>(dx, tb, ox), py = mad<R>(dy, tb, oy), pz = mad<R>(dz, tb, oz);
    const int fx = (int)floor((px - s.gmin[0]) * s.inv_cs[0]);
    const int fy = (int)floor((py - s.gmin[1]) * s.inv_cs[1]);
    const int fz = (int)floor((pz - s.gmin[2]) * s.inv_cs[2]);
    // never move backwards; stay inside the box on the other axes
    // (clamp to [c, c + h] moving up, [c - h, c] moving down, c when still);
    // across the exit face on the exit axis
    int nx = min(max(fx, dx < R(0) ? w.cx - hx : w.cx), dx > R(0) ? w.cx + hx : w.cx);
    int ny = min(max(fy, dy < R(0) ? w.cy - hy : w.cy), dy > R(0) ? w.cy + hy : w.cy);
    int nz = min(max(fz, dz < R(0) ? w.cz - hz : w.cz), dz > R(0) ? w.cz + hz : w.cz);
    const int sx = dx > R(0) ? hx + 1 : -hx - 1, sy = dy > R(0) ? hy + 1 : -hy - 1, sz = dz > R(0) ? hz + 1 : -hz - 1;
    nx = ax == 0 ? w.cx + sx : nx;
    ny = ax == 1 ? w.cy + sy : ny;
    nz = ax == 2 ? w.cz + sz : nz;
    if ((unsigned)nx >= (unsigned)s.gx || (unsigned)ny >= (unsigned)s.gy || (unsigned)nz >= (unsigned)s.gz)
        return true;
    // a single step enters through the face towards the voxel left (-x when
    // moving in +x); a jump leaves no neighbour whose list was tested
    const R da = ax == 0 ? dx : (ax == 1 ? dy : dz);
    w.from = w.k == 0 ? 2 * ax + (da > R(0) ? 0 : 1) : -1;
    w.cx = nx; w.cy = ny; w.cz = nz;
    w.tmx = axis_exit<R>(w.sx, w.qx, nx);
    w.tmy = axis_exit<R>(w.sy, w.qy, ny);
    w.tmz = axis_exit<R>(w.sz, w.qz, nz);
    return false;
}

// Picks the next three entries of the walk's current voxel to test (the last
// one repeated when fewer are left: a repeated triangle never replaces
// itself), or opens the next 32-entry window of a longer list (never skipped).
struct Three {
    uint32_t j0, j1, j2, n;
};
template <class R> __device__ __forceinline__ Three next_three(Walk<R> &w)
{
    if (w.todo == 0) {
        w.j += 32;
        const uint32_t n = w.je - w.j;
        w.todo = n >= 32 ? ~0u : (1u << n) - 1u;
    }
    const uint32_t m0 = w.todo, m1 = m0 & (m0 - 1u), m2 = m1 & (m1 - 1u);
    const uint32_t b0 = __builtin_ctz(m0);
    const uint32_t b1 = m1 ? __builtin_ctz(m1 | (1u << 31)) : b0;
    const uint32_t b2 = m2 ? __builtin_ctz(m2 | (1u << 31)) : b1;
    w.todo = m2 & (m2 - 1u);
    return Three{w.j + b0, w.j + b1, w.j + b2, 1u + (m1 != 0u) + (m2 != 0u)};
}

// Entries of the current voxel still to test.
template <class R> __device__ __forceinline__ bool walk_pending(const Walk<R> &w)
{
    return (w.todo != 0u) | (w.j + 32u < w.je);
}

// Flattened form of walk_steps: each of the `budget` units either tests the
// next three untested list entries of the current voxel or leaves the voxel
// (one step, or one jump across an empty block, by the same straight-line
// code), so lanes in triangle-dense voxels do not hold the wave on a long
// inner loop.  Every triangle is tested once (Walk), and the nearest hit is
// kept wherever it lies, so the closest hit is that of walk_steps.
// SCHED: the wave takes one kind of unit per iteration -- triangle tests when
// at least half of its walking lanes have one pending, else voxel moves -- so
// an iteration runs one of the two code paths instead of both; lanes of the
// other kind sit that unit out.
template <class R, bool LDS, bool SCHED = false>
__device__ bool walk_units(const DevScene<R> &s, Walk<R> &w, R ox, R oy, R oz, R dx, R dy, R dz, int excl,
                           int budget, uint32_t &n_vox, uint32_t &n_test)
{
    if (w.k < 0) walk_enter<R, LDS>(s, w, n_vox);  // the first voxel of a new walk
    for (int it = 0; it < budget; ++it) {
        const R texit = fmin(w.tmx, fmin(w.tmy, w.tmz));
        const bool pending = walk_pending(w);
        if constexpr (SCHED) {
            const int n_pending = __popcll(__ballot(pending));
            const int n_walking = __popcll(__ballot(true));
            const bool tests_now = 2 * n_pending >= n_walking;
            if (pending != tests_now) continue;  // this lane's kind of unit waits
        }
        if (pending) {
            // any hit nearer than the best so far, inside this voxel or not
            const R tlim = w.tbest;
            // the tests first: their row loads then need not wait for the index loads
            // as selects, not a branch: a branch would defer the id loads behind
            // the tests.  The best triangle tested again gives its own t, which
            // neither order test takes
            auto take = [&](bool hit, int t, R tt) __attribute__((always_inline)) {
                const bool nearer = hit & !(Tmin<R>::excl && t == excl) &
                                    ((tt < w.tbest) | ((tt == w.tbest) & (t < w.best)));
                w.tbest = nearer ? tt : w.tbest;
                w.best = nearer ? t : w.best;
            };
            {
                // three list entries per unit, merged in list order, as
                // single units would (four spill scalar registers and run 7 %
                // slower)
                const Three e = next_three(w);
                const uint32_t j0 = e.j0, j1 = e.j1, j2 = e.j2;
                n_test += e.n;
                const int t0 = (int)s.cell_tris[j0];
                const int t1 = (int)s.cell_tris[j1];
                const int t2 = (int)s.cell_tris[j2];
                R tt0, tt1, tt2;
                bool h0, h1, h2;
                if constexpr (sizeof(R) == 4) {
                    // every row load of the unit first, then the three tests
                    const Rows q0 = load_rows(s, j0), q1 = load_rows(s, j1), q2 = load_rows(s, j2);
                    // unlimited: take() keeps a hit only if it is nearer than the best so far
                    h0 = tri_hit_rows<false>(q0, ox, oy, oz, dx, dy, dz, tlim, tt0);
                    h1 = tri_hit_rows<false>(q1, ox, oy, oz, dx, dy, dz, tlim, tt1);
                    h2 = tri_hit_rows<false>(q2, ox, oy, oz, dx, dy, dz, tlim, tt2);
                } else {
                    h0 = tri_hit<R>(s, t0, j0, ox, oy, oz, dx, dy, dz, tlim, tt0);
                    h1 = tri_hit<R>(s, t1, j1, ox, oy, oz, dx, dy, dz, tlim, tt1);
                    h2 = tri_hit<R>(s, t2, j2, ox, oy, oz, dx, dy, dz, tlim, tt2);
                }
                take(h0, t0, tt0);
                take(h1, t1, tt1);
                take(h2, t2, tt2);
            }
            // the voxel's last entries tested and the nearest hit inside it:
            // the walk ends here rather than in the next move unit
            if (!walk_pending(w) && w.best >= 0 && w.tbest <= texit) return true;
            continue;
        }
        if (w.best >= 0 && w.tbest <= texit) return true;
        if (walk_move<R>(s, w, ox, oy, oz, dx, dy, dz)) return true;
        walk_enter<R, LDS>(s, w, n_vox);
        // nothing to test here and the nearest hit before this voxel's exit
        if (!walk_pending(w) && w.best >= 0 && w.tbest <= fmin(w.tmx, fmin(w.tmy, w.tmz))) return true;
    }
    return false;
}

// ---------------------------------------------------------------- mixed walk
// The mixed walk gives the FP64 closest hit (the Moller-Trumbore test of the
// FP64 path and the oracle, t > 1e-10, lowest triangle on equal t) at the cost
// of an FP32 walk.  The FP32 walk screens candidates conservatively
// (tri_filter) and keeps the one nearest by FP32 distance with its error bound
// e, plus the least lower bounds t - e of the other candidates.  After the
// walk the FP64 test decides the few candidates that can be nearest
// (mix_confirm); when that is not enough to be certain, the bounce is redone
// by the FP64 walk (k_exact, or shoot<double>).  The triangle the ray leaves is
// skipped in FP32: FP64 accepts it again only when the ray leaves at a grazing
// angle, |n . d| < mx_self, and such bounces go to the FP64 walk directly.

// Limit of the voxel just entered.  texit, the FP32 exit distance, is off by
// at most kx / |d_a| + 8 eps texit on its exit axis a.  Every voxel list reaches
// mx_infl past its voxel, so a hit up to mx_infl beyond the exit is tested by
// the time this voxel is done: once the best hit's upper bound is below
// texit + infl - err no untested triangle can be nearer.
__device__ __forceinline__ void mix_limits(const DevScene<float> &s, const TraceConst &tc, Walk<float> &w)
{
    const float texit = fminf(w.tmx, fminf(w.tmy, w.tmz));
    // |1 / d_a| = |s_a| / cs_a.  Select values, not fields: a conditional over
    // two fields is an lvalue, and loading through the selected address keeps
    // the whole walk state in memory
    const float ix = fabsf(w.sx) * s.inv_cs[0], iy = fabsf(w.sy) * s.inv_cs[1], iz = fabsf(w.sz) * s.inv_cs[2];
    const float ia = w.tmx <= w.tmy ? (w.tmx <= w.tmz ? float(ix) : float(iz)) : (w.tmy <= w.tmz ? float(iy) : float(iz));
    const float err = __builtin_fmaf(tc.mx_kx, ia, 4.8e-7f * texit);
    w.tend = texit + tc.mx_infl - err;
}

// Records lower bound lb of candidate id (not the best) among the two least.
__device__ __forceinline__ void mix_lower(Walk<float> &w, float lb, int id)
{
    if (id == w.sid) return;  // listed again in a later voxel
    if (lb < w.sl) {
        w.sl2 = w.sl;
        w.sl = lb;
        w.sid = id;
    } else {
        w.sl2 = fminf(w.sl2, lb);
    }
}

__device__ __forceinline__ void mix_walk_init(const DevScene<float> &s, Walk<float> &w, const float *o, const float *d)
{
    walk_init<float>(s, w, o[0], o[1], o[2], d[0], d[1], d[2]);
    const float INF = __builtin_huge_valf();
    w.eb = 0.0f;
    w.sl = w.sl2 = INF;
    w.sid = w.zid = -1;
    w.tend = 0.0f;
}

// Flattened walk (as walk_units, SCHED: one kind of unit per wave iteration)
// in mixed form; excl: the triangle left.
template <bool LDS, bool SCHED>
__device__ bool walk_units_mix(const DevScene<float> &s, const TraceConst &tc, Walk<float> &w, float ox, float oy,
                               float oz, float dx, float dy, float dz, int excl, int budget, uint32_t &n_vox,
                               uint32_t &n_test)
{
    if (w.k < 0) {
        walk_enter<float, LDS>(s, w, n_vox);
        mix_limits(s, tc, w);
    }
    for (int it = 0; it < budget; ++it) {
        const bool pending = walk_pending(w);
        if constexpr (SCHED) {
            const int n_pending = __popcll(__ballot(pending));
            const int n_walking = __popcll(__ballot(true));
            const bool tests_now = 2 * n_pending >= n_walking;
            if (pending != tests_now) continue;
        }
        if (pending) {
            const Three e = next_three(w);
            const uint32_t j0 = e.j0, j1 = e.j1, j2 = e.j2;
            n_test += e.n;
            const int t0 = (int)s.cell_tris[j0];
            const int t1 = (int)s.cell_tris[j1];
            const int t2 = (int)s.cell_tris[j2];
            // every candidate is recorded (as the best, or by its lower bound)
            // wherever its hit lies: the walk tests each triangle once
            const float lim = __builtin_huge_valf();
            float tt0, tt1, tt2, e0, e1, e2;
            bool p0, p1, p2;
            // every row load of the unit first, then the three screens
            const Rows q0 = load_rows(s, j0), q1 = load_rows(s, j1), q2 = load_rows(s, j2);
            const bool c0 = tri_filter(q0, tc, ox, oy, oz, dx, dy, dz, lim, tt0, e0, p0);
            const bool c1 = tri_filter(q1, tc, ox, oy, oz, dx, dy, dz, lim, tt1, e1, p1);
            const bool c2 = tri_filter(q2, tc, ox, oy, oz, dx, dy, dz, lim, tt2, e2, p2);
            auto take = [&](bool c, bool par, int t, float tt, float e) __attribute__((always_inline)) {
                if (t == w.best || !(c || par)) return;  // not a candidate, or listed again
                if (t != excl) {
                    if (par) {
                        w.sl = -__builtin_huge_valf();  // FP32 cannot place this hit
                    } else if (tt - e <= 0.0f) {
                        // possibly a hit FP64 rejects (t <= 1e-10): decided after the walk
                        if (w.zid < 0 || w.zid == t) w.zid = t;
                        else w.sl = -__builtin_huge_valf();
                    } else if (tt < w.tbest) {
                        if (w.best >= 0) mix_lower(w, w.tbest - w.eb, w.best);
                        w.tbest = tt;
                        w.eb = e;
                        w.best = t;
                    } else {
                        mix_lower(w, tt - e, t);
                    }
                }
            };
            take(c0, p0, t0, tt0, e0);
            take(c1, p1, t1, tt1, e1);
            take(c2, p2, t2, tt2, e2);
            continue;
        }
        if (w.best >= 0 && w.tbest + w.eb <= w.tend) return true;
        if (walk_move<float>(s, w, ox, oy, oz, dx, dy, dz)) return true;
        walk_enter<float, LDS>(s, w, n_vox);
        mix_limits(s, tc, w);
    }
    return false;
}

// After a mixed walk: the FP64 closest hit (hit, t64; hit -1: a miss) when it
// is certain, else false (the bounce needs the FP64 walk).  The FP64 test
// decides the best candidate, the candidate with the least other lower bound
// when the two intervals overlap (two triangles at a shared edge) or the best
// fails, and the candidate near the departure point.  The nearest accepted one
// is certain when every candidate not tested lies beyond it (the walk records
// every candidate it screens: the best, or the lower bound of the others) and
// it lies before the end of the walk.
__device__ __forceinline__ bool mix_confirm(const DevScene<double> &s64, const Walk<float> &w, const double *o,
                                            const double *d, int &hit, double &t64)
{
    if (w.sl == -__builtin_huge_valf()) return false;
    double tm = __builtin_huge_val();
    int im = -1;
    auto test = [&](int t) __attribute__((always_inline)) {
        double tt;
        if (tri_hit_mt(s64, t, o[0], o[1], o[2], d[0], d[1], d[2], __builtin_huge_val(), tt) &&
            (tt < tm || (tt == tm && t < im))) {
            tm = tt;
            im = t;
        }
    };
    if (w.best >= 0) test(w.best);
    float lb = w.sl;  // least lower bound of the candidates not tested in FP64
    if (w.sid >= 0 && (im < 0 || w.sl <= w.tbest + w.eb)) {
        test(w.sid);
        lb = w.sl2;
    }
    if (w.zid >= 0) test(w.zid);
    if (im < 0) {
        hit = -1;
        t64 = 0.0;
        return w.best < 0 && w.zid < 0;  // a miss is certain only without candidates
    }
    if (!(tm < (double)lb) || !(tm <= (double)w.tend)) return false;
    if (w.best >= 0 && !(tm <= (double)(w.tbest + w.eb))) return false;
    hit = im;
    t64 = tm;
    return true;
}

// Complete closest-hit query (the broadband phase uses it directly).
template <class R, bool LDS>
__device__ __forceinline__ int shoot(const DevScene<R> &s, R ox, R oy, R oz, R dx, R dy, R dz, int excl, R &tbest,
                                     uint64_t &rmask)
{
    Walk<R> w;
    walk_init<R>(s, w, ox, oy, oz, dx, dy, dz);
    uint32_t nv = 0, nt = 0;
    while (!walk_steps<R, LDS>(s, w, ox, oy, oz, dx, dy, dz, excl, 1 << 20, nv, nt)) {
    }
    tbest = w.tbest;
    rmask = w.rmask;
    return w.best;
}

// ---------------------------------------------------------------- receivers
// Event counters of a launch.  The voxel and test counts the walk bumps in its
// inner loop stay in registers; every other event goes to its workgroup's
// table in LDS by atomics, so no register holds those counts.
struct LaneCounters {
    uint32_t vox, test;            // per lane
    unsigned long long *sc;        // LDS [C_N]
    double *sl;                    // LDS: lost energy
    __device__ __forceinline__ void add(int c, unsigned n = 1u) const { atomicAdd(sc + c, (unsigned long long)n); }
    __device__ __forceinline__ void lose(double e) const { atomicAdd(sl, e); }
};

// The workgroup's counter table: cleared by its threads, flushed to the
// launch's device counters once every wave of the workgroup is done.
#define PACH_COUNTERS_BEGIN(lc)                                                  \
    __shared__ unsigned long long pach_sc[C_N];                                   \
    __shared__ double pach_sl;                                                     \
    for (int i_ = threadIdx.x; i_ < C_N; i_ += blockDim.x) pach_sc[i_] = 0;       \
    if (threadIdx.x == 0) pach_sl = 0;                                             \
    __syncthreads();                                                               \
    LaneCounters lc{0u, 0u, pach_sc, &pach_sl}

__device__ __forceinline__ void counters_end(const LaneCounters &lc, DevCounters *cnt)
{
    if (lc.vox) atomicAdd(lc.sc + C_VOXEL, (unsigned long long)lc.vox);
    if (lc.test) atomicAdd(lc.sc + C_TEST, (unsigned long long)lc.test);
    __syncthreads();
    if (threadIdx.x < C_N && lc.sc[threadIdx.x]) atomicAdd(&cnt->v[threadIdx.x], lc.sc[threadIdx.x]);
    if (threadIdx.x == 0 && *lc.sl != 0) unsafeAtomicAdd(&cnt->lost, *lc.sl);
}

// p: the pressure sqrt(e rho_c), computed by the caller in the ray's precision
// Exact accumulation of v x scale (a power of two, so the product is exact):
// the rounded magnitude is split into four 32-bit limbs (each split is exact:
// the remainder keeps only low mantissa bits of the value), and each nonzero
// limb, negated for a negative value, is added to its int64 sum.  Sums of up
// to 2^30 deposits per bin cannot overflow.  Returns false (nothing added)
// for a value at or beyond 2^123 after scaling.
__device__ __forceinline__ bool fx_add(long long *fx, uint64_t limb_stride, uint64_t i, double v, double scale)
{
    double t = fabs(v * scale);
    if (!(t < 0x1p123)) return false;
    const double h3 = floor(t * 0x1p-96);
    t -= h3 * 0x1p96;
    const double h2 = floor(t * 0x1p-64);
    t -= h2 * 0x1p64;
    const double h1 = floor(t * 0x1p-32);
    t -= h1 * 0x1p32;
    const double h0 = rint(t);
    const double h[4] = {h0, h1, h2, h3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (h[k] != 0.0) {
            const long long q = (long long)h[k];
            atomicAdd((unsigned long long *)(fx + k * limb_stride + i),
                      (unsigned long long)(v < 0 ? -q : q));
        }
    }
    return true;
}

// The planes of one deposit through exact accumulation (see DevRecv::fx).
__device__ __forceinline__ bool hist_add_fx(long long *fx, uint64_t stride, uint64_t idx, double e, double p,
                                         double ndx, double ndy, double ndz, int ana, double se, double sp)
{
    const uint64_t ls = kPlanes * stride;
    const double vx = (double)(float)(ndx * e), vy = (double)(float)(ndy * e), vz = (double)(float)(ndz * e);
    bool ok = fx_add(fx, ls, idx, e, se);
    ok &= fx_add(fx, ls, stride + idx, p, sp);
    ok &= fx_add(fx, ls, (ndx > 0 ? 2 : 5) * stride + idx, vx, se);
    ok &= fx_add(fx, ls, (ndy > 0 ? 3 : 6) * stride + idx, vy, se);
    ok &= fx_add(fx, ls, (ndz > 0 ? 4 : 7) * stride + idx, vz, se);
    ok &= fx_add(fx, ls, (8 + ana) * stride + idx, e, se);
    return ok;
}

__device__ __forceinline__ void hist_add(const double *base_unused, double *hist, uint64_t stride,
                                         uint64_t idx, double e, double p, double ndx,
                                         double ndy, double ndz, int ana)
{
    unsafeAtomicAdd(hist + idx, e);
    unsafeAtomicAdd(hist + stride + idx, p);
    // signed direction sums through a float cast, as Directional_Histogram.Add does
    const double vx = (double)(float)(ndx * e), vy = (double)(float)(ndy * e), vz = (double)(float)(ndz * e);
    unsafeAtomicAdd(hist + (ndx > 0 ? 2 : 5) * stride + idx, vx);
    unsafeAtomicAdd(hist + (ndy > 0 ? 3 : 6) * stride + idx, vy);
    unsafeAtomicAdd(hist + (ndz > 0 ? 4 : 7) * stride + idx, vz);
    unsafeAtomicAdd(hist + (8 + ana) * stride + idx, e);
}

template <class R> __device__ __forceinline__ double band_gain(const DevRecv<R> &rv, int band, double dist);
// 10^(-0.1 a x) as exp2(k x), k = -0.1 a log2(10) (within ~1e-15 of the
// reference's Math.Pow(10, ...); libm's pow needs far more registers)
template <> __device__ __forceinline__ double band_gain<double>(const DevRecv<double> &rv, int band, double dist)
{
    return exp2(rv.band->kexp2d[band] * dist);
}
template <> __device__ __forceinline__ double band_gain<float>(const DevRecv<float> &rv, int band, double dist)
{
    return (double)__builtin_amdgcn_exp2f(rv.band->kexp2[band] * (float)dist);
}

// Pressure of a deposit, sqrt(e rho_c): FP64 IEEE on the FP64 path, FP32
// hardware square root on the FP32 path.
template <class R> __device__ __forceinline__ double pressure(double e, double rho_c)
{
    if constexpr (sizeof(R) == 8) return sqrt(e * rho_c);
    else return (double)__builtin_amdgcn_sqrtf((float)(e * rho_c));
}

// Receiver_Bank.CheckRay (BROAD false: an octave ray of intensity I in
// `band`) / CheckBroadbandRay (BROAD true: energies E[lo..hi]) for receiver r.
// sqd: squared distance to EndPt.  The octave form takes its intensity by
// value: a pointer into the ray that a runtime band could index would keep the
// whole ray record in private memory.
template <class R, bool BROAD>
__device__ __forceinline__ void receiver_test(const DevRecv<R> &rv, int r, R ox, R oy, R oz, R dx, R dy, R dz,
                                              R sqd, const R *E, R I, int band, int lo, int hi, double tsum, int ana,
                                              LaneCounters &lc)
{
    const auto C = rv.ctr[r];
    const R mx = ox - C.x, my = oy - C.y, mz = oz - C.z;
    const R b = mx * dx + my * dy + mz * dz;
    const R c = (mx * mx + my * my + mz * mz) - rv.r2;
    if (c > R(0) && b > R(0)) return;
    R disc;
    if constexpr (sizeof(R) == 4) {
        // the squared half-chord as r^2 - |m - b d|^2: in FP32, b^2 - c loses
        // ~|m|^2 x 2^-24 (1e-4 m^2 at 40 m) to cancellation against half-chords
        // of at most r^2; the perpendicular offset keeps it to ~|m| x 2^-24 x r
        // (Haines et al., "Precision Improvements for Ray/Sphere Intersection",
        // Ray Tracing Gems, 2019).  FP64 keeps the reference's b^2 - c.
        const R qx = mx - b * dx, qy = my - b * dy, qz = mz - b * dz;
        disc = rv.r2 - (qx * qx + qy * qy + qz * qz);
    } else {
        disc = b * b - c;
    }
    if (disc < R(0)) return;
    const R sq = fsqrt<R>(disc);
    const R t1 = -b - sq, t2 = -b + sq;
    const R t = (t1 + t2) * R(.5);
    if (!(t > R(0) && t * t < sqd)) return;
    const double ts = (double)((t2 - t1) * R(0.5));
    const double rt = (double)t * rv.inv_c + tsum;
    const int sample = (int)(rt * rv.fs);
    if (sample >= rv.sct || sample < 0) return;
    const double dist = rt * rv.c;
    const double ndx = -(double)dx, ndy = -(double)dy, ndz = -(double)dz;
    if constexpr (!BROAD) {
        const uint64_t idx = ((uint64_t)r * kBands + band) * rv.sct + sample;
        const double e = (double)I * band_gain<R>(rv, band, dist) * rv.sizemod * ts;
        if (rv.fx) {
            if (!hist_add_fx(rv.fx, rv.plane_stride, idx, e, pressure<R>(e, rv.rho_c), ndx, ndy, ndz, ana,
                             rv.fx_se, rv.fx_sp))
                lc.add(C_FXOVF);
        } else {
            hist_add(nullptr, rv.hist, rv.plane_stride, idx, e, pressure<R>(e, rv.rho_c), ndx, ndy, ndz, ana);
        }
        lc.add(C_DEP);
    } else {
        for (int o = lo; o <= hi; ++o) {
            const uint64_t idx = ((uint64_t)r * kBands + o) * rv.sct + sample;
            const double e = (double)E[o] * band_gain<R>(rv, o, dist) * rv.sizemod * ts;
            if (rv.fx) {
                if (!hist_add_fx(rv.fx, rv.plane_stride, idx, e, pressure<R>(e, rv.rho_c), ndx, ndy, ndz, 0,
                                 rv.fx_se, rv.fx_sp))
                    lc.add(C_FXOVF);
            } else {
                hist_add(nullptr, rv.hist, rv.plane_stride, idx, e, pressure<R>(e, rv.rho_c), ndx, ndy, ndz, 0);
            }
            lc.add(C_DEP);
        }
    }
}

// Receiver_Bank.CheckRay / CheckBroadbandRay.  sqd: squared distance to EndPt;
// E: intensity (octave) or energies (broadband).  With per-voxel receiver
// masks, only receivers whose spheres touch a voxel the walk crossed are
// tested: a sphere the segment passes through contains the chord midpoint,
// which lies in a crossed voxel, so the deposits are those of the full loop.
template <class R, bool BROAD>
__device__ void check_receivers(const DevRecv<R> &rv, const RoomTab &rm, uint64_t cand, R ox, R oy, R oz, R dx,
                                R dy, R dz, R sqd, const R *E, R I, int band, int lo, int hi, double tsum, int ana,
                                LaneCounters &lc)
{
    // voxel mask bits are receiver indices relative to the room's first
    // receiver.  The room's range is read once: the deposits' atomics would
    // otherwise make every iteration read it again
    const int base = rm.rec_base, count = rm.rec_count;
    if (rv.masked) {
        cand &= count >= 64 ? ~0ull : ((1ull << count) - 1);
        while (cand) {
            const int r = __builtin_ctzll(cand);
            cand &= cand - 1;
            receiver_test<R, BROAD>(rv, base + r, ox, oy, oz, dx, dy, dz, sqd, E, I, band, lo, hi, tsum, ana, lc);
        }
    } else {
        for (int r = base; r < base + count; ++r)
            receiver_test<R, BROAD>(rv, r, ox, oy, oz, dx, dy, dz, sqd, E, I, band, lo, hi, tsum, ana, lc);
    }
}

// ---------------------------------------------------------------- mapping receivers
// PachMapReceiver.CheckRay (Pach_Map_Receiver.cs:1493-1631) and
// CheckBroadbandRay (:1359-1491): a voxel walk over the receiver grid from the
// segment's origin (or its entry into the grid) towards EndPt, each listed
// receiver tested once per segment (CheckReceivers_*, :349-403).  As written
// there, the octave walk stops on entering the end voxel, and the broadband
// walk starts from the ray's own origin, steers by the signs of the origin's
// coordinates and stops once any index reaches its end value.  In FP64 so that
// both precisions deposit what the reference would.
struct MapWalk {
    int X, Y, Z, sx, sy, sz;
    double tx, ty, tz;     // tMax per axis
    double dx, dy, dz;     // tDelta per axis
};

// Voxel index of coordinate v along an axis, clamped to the grid.
__device__ __forceinline__ int map_index(double v, double mn, double vd, int ct)
{
    const int i = (int)floor((v - mn) / vd);
    return i < 0 ? 0 : (i >= ct ? ct - 1 : i);
}

__device__ __forceinline__ bool map_walk_init(const DevMap &m, const double o[3], const double d[3],
                                              const double end[3], bool broadband, MapWalk &w, int Xe[3])
{
    double sx0 = o[0], sy0 = o[1], sz0 = o[2];
    const bool inside = !(o[0] < m.bmin[0] || o[0] > m.bmax[0]) && !(o[1] < m.bmin[1] || o[1] > m.bmax[1]) &&
                        !(o[2] < m.bmin[2] || o[2] > m.bmax[2]);
    if (!inside) {
        // slab entry (t >= 0)
        double t0 = 0, t1 = __builtin_huge_val();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (d[k] == 0) {
                if (o[k] < m.bmin[k] || o[k] > m.bmax[k]) return false;
                continue;
            }
            double a = (m.bmin[k] - o[k]) / d[k], b = (m.bmax[k] - o[k]) / d[k];
            if (a > b) { const double t = a; a = b; b = t; }
            if (a > t0) t0 = a;
            if (b < t1) t1 = b;
        }
        if (t0 > t1) return false;
        sx0 = o[0] + d[0] * t0;
        sy0 = o[1] + d[1] * t0;
        sz0 = o[2] + d[2] * t0;
    }
    // the broadband walk starts from the origin and steers by its coordinates
    const double fx = broadband ? o[0] : sx0, fy = broadband ? o[1] : sy0, fz = broadband ? o[2] : sz0;
    const double gx = broadband ? o[0] : d[0], gy = broadband ? o[1] : d[1], gz = broadband ? o[2] : d[2];
    w.X = map_index(fx, m.bmin[0], m.vd[0], m.ct[0]);
    w.Y = map_index(fy, m.bmin[1], m.vd[1], m.ct[1]);
    w.Z = map_index(fz, m.bmin[2], m.vd[2], m.ct[2]);
    Xe[0] = (int)floor((end[0] - m.bmin[0]) / m.vd[0]);
    Xe[1] = (int)floor((end[1] - m.bmin[1]) / m.vd[1]);
    Xe[2] = (int)floor((end[2] - m.bmin[2]) / m.vd[2]);
    if ((gx < 0 && w.X < Xe[0]) || (gx >= 0 && w.X > Xe[0])) return false;
    if ((gy < 0 && w.Y < Xe[1]) || (gy >= 0 && w.Y > Xe[1])) return false;
    if ((gz < 0 && w.Z < Xe[2]) || (gz >= 0 && w.Z > Xe[2])) return false;
    auto axis = [](int X, double f, double g, double mn, double vd, int &st, double &tm, double &td) {
        const double lo = X * vd + mn, hi = (X + 1) * vd + mn;
        if (g < 0) {
            st = -1;
            tm = (lo - f) / g;
        } else {
            st = 1;
            tm = (hi - f) / g;
        }
        td = vd / g * st;
    };
    axis(w.X, fx, gx, m.bmin[0], m.vd[0], w.sx, w.tx, w.dx);
    axis(w.Y, fy, gy, m.bmin[1], m.vd[1], w.sy, w.ty, w.dy);
    axis(w.Z, fz, gz, m.bmin[2], m.vd[2], w.sz, w.tz, w.dz);
    return true;
}

// one step; false once the walk has ended
__device__ __forceinline__ bool map_walk_step(const DevMap &m, MapWalk &w, const int Xe[3], bool broadband)
{
    if (w.tx < w.ty ? w.tx < w.tz : false) {
        w.X += w.sx;
        if (w.X < 0 || w.X >= m.ct[0]) return false;
        w.tx = w.tx + w.dx;
    } else if (!(w.tx < w.ty) && w.ty < w.tz) {
        w.Y += w.sy;
        if (w.Y < 0 || w.Y >= m.ct[1]) return false;
        w.ty = w.ty + w.dy;
    } else {
        w.Z += w.sz;
        if (w.Z < 0 || w.Z >= m.ct[2]) return false;
        w.tz = w.tz + w.dz;
    }
    return broadband ? (w.X != Xe[0] && w.Y != Xe[1] && w.Z != Xe[2])
                     : (w.X != Xe[0] || w.Y != Xe[1] || w.Z != Xe[2]);
}

__device__ __forceinline__ int map_cell(const DevMap &m, const MapWalk &w)
{
    return (w.X * m.ct[1] + w.Y) * m.ct[2] + w.Z;
}

// is receiver j listed in cell? (lists are sorted)
__device__ __forceinline__ bool map_listed(const DevMap &m, int cell, uint32_t j)
{
    uint32_t lo = m.start[cell], hi = m.start[cell + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = m.list[mid];
        if (v == j) return true;
        if (v < j) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

// Map_Receiver.CheckRay / CheckBroadbandRay (:1780-1826) and
// Histogram.Add / Directional_Histogram.Add at 1000 Hz.  band >= 0: octave ray
// of intensity E[0]; band < 0: broadband energies E[0..7].
template <class R, bool BROAD>
__device__ void map_check(const DevRecv<R> &rv, const R *oR, const R *dR, R ex, R ey, R ez, const R *E, R I, int band,
                          double t_sum, LaneCounters &lc)
{
    const DevMap &m = rv.map;
    constexpr bool broadband = BROAD;
    const double o[3] = {(double)oR[0], (double)oR[1], (double)oR[2]};
    const double d[3] = {(double)dR[0], (double)dR[1], (double)dR[2]};
    const double end[3] = {(double)ex, (double)ey, (double)ez};
    // A deposit needs the point of the segment nearest a centre (the test's
    // t in (0, |segment|)) within r of it, so inside the receivers' box: a
    // segment whose bounding box misses that box (widened by r) deposits nothing.
    if (fmax(oR[0], ex) < (R)m.lo[0] || fmin(oR[0], ex) > (R)m.hi[0] || fmax(oR[1], ey) < (R)m.lo[1] ||
        fmin(oR[1], ey) > (R)m.hi[1] || fmax(oR[2], ez) < (R)m.lo[2] || fmin(oR[2], ez) > (R)m.hi[2])
        return;
    MapWalk w;
    int Xe[3];
    if (!map_walk_init(m, o, d, end, broadband, w, Xe)) return;
    const double qx = end[0] - o[0], qy = end[1] - o[1], qz = end[2] - o[2];
    const double sqd = qx * qx + qy * qy + qz * qz;
    const float dfx = (float)d[0], dfy = (float)d[1], dfz = (float)d[2], r2f = (float)m.r2;
    // FP32 screen of one list entry: the line misses the sphere by more than
    // the FP32 rounding of this test (1e-5 |m|^2 against ~1e-6) and of the
    // entry's FP32 centre (m.scr) can hide, so the FP64 test would reject it too
    auto screen = [&](const float4 &L) __attribute__((always_inline)) {
        float fx, fy, fz;
        if constexpr (std::is_same<R, float>::value) {
            fx = oR[0] - L.x;
            fy = oR[1] - L.y;
            fz = oR[2] - L.z;
        } else {
            fx = (float)(o[0] - (double)L.x);
            fy = (float)(o[1] - (double)L.y);
            fz = (float)(o[2] - (double)L.z);
        }
        const float fb = fx * dfx + fy * dfy + fz * dfz;
        const float fm = fx * fx + fy * fy + fz * fz;
        return !(fb * fb - (fm - r2f) < -1e-5f * fm - 1e-12f - m.scr);
    };
    for (int visit = 0;; ++visit) {
        const int cell = map_cell(m, w);
        const uint32_t k1 = m.start[cell + 1];
        // four entries per step, their loads issued together
        for (uint32_t k = m.start[cell]; k < k1; k += 4) {
            float4 L[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) L[u] = m.lctr[min(k + (uint32_t)u, k1 - 1)];
            uint32_t pass = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (k + (uint32_t)u < k1 && screen(L[u])) pass |= 1u << u;
            while (pass) {
                const int u = __builtin_ctz(pass);
                pass &= pass - 1;
                const uint32_t j = __float_as_uint(u == 0 ? L[0].w : (u == 1 ? L[1].w : (u == 2 ? L[2].w : L[3].w)));
                const double4 C = m.ctr[j];
                const double mx = o[0] - C.x, my = o[1] - C.y, mz = o[2] - C.z;
                const double b = mx * d[0] + my * d[1] + mz * d[2];
                const double c = (mx * mx + my * my + mz * mz) - m.r2;
                if (c > 0 && b > 0) continue;
                const double discr = b * b - c;
                if (discr < 0) continue;
                const double t1 = -b - sqrt(discr), t2 = -b + sqrt(discr);
                const double t = (t1 + t2) / 2;
                if (!(t > 0 && t * t < sqd)) continue;
                // tested once per segment: skip a receiver an earlier voxel of the
                // walk lists (the walk is replayed from its start; hits are rare)
                bool earlier = false;
                if (visit > 0) {
                    MapWalk v;
                    int Xv[3];
                    map_walk_init(m, o, d, end, broadband, v, Xv);
                    for (int q = 0; q < visit && !earlier; ++q) {
                        earlier = map_listed(m, map_cell(m, v), j);
                        map_walk_step(m, v, Xv, broadband);
                    }
                }
                if (earlier) continue;
                const double raydist = t / m.c + t_sum - C.w;
                // Histogram.Add drops samples outside [0, SampleCT); Histogram_1Pt.Add
                // (Classes_Receivers.cs:1393-1396) adds every deposit to sample 0
                const int sample = m.time1pt ? 0 : (int)(raydist * 1000);
                if (!m.time1pt && (sample >= m.sct || sample < 0)) continue;
                const int b_lo = broadband ? 0 : band, b_hi = broadband ? kBands - 1 : band;
                for (int oc = b_lo; oc <= b_hi; ++oc) {
                    double Ei;
                    if constexpr (BROAD) Ei = (double)E[oc];
                    else Ei = (double)I;
                    const double e = Ei * pow(10.0, -.1 * rv.band->atten[oc] * raydist) * m.sizemod;
                    const uint64_t idx = ((uint64_t)j * kBands + oc) * m.sct + sample;
                    unsafeAtomicAdd(m.hist + idx, e);
                    if (m.time1pt) {
                        // the energy plane only
                    } else if (m.directional) {
                        unsafeAtomicAdd(m.hist + m.plane_stride + idx, sqrt(e * m.rho_c));
                        for (int a = 0; a < 3; ++a) {
                            const double dk = -d[a];
                            const double vv = (double)(float)(dk * e);
                            unsafeAtomicAdd(m.hist + (uint64_t)(dk > 0 ? 2 + a : 5 + a) * m.plane_stride + idx, vv);
                        }
                    } else {
                        unsafeAtomicAdd(m.hist + m.plane_stride + idx, sqrt(e));
                    }
                    lc.add(C_DEP);
                }
            }
        }
        if (!map_walk_step(m, w, Xe, broadband)) break;
    }
}

// ---------------------------------------------------------------- reflection
// sin(2 pi u), cos(2 pi u) for u in [0, 1): quarter-turn reduction and the
// Taylor series of sin / cos on [-pi/4, pi/4] to degree 17/18 in a fixed Horner
// order.  FP64 only: +, -, * are IEEE-exact operations, so the FP64 kernel and
// the CPU oracle produce identical bits.
__device__ __forceinline__ void sincos_turn(double u, double &s, double &c)
{
    const int k = (int)floor(4.0 * u + 0.5);
    const double r = u - 0.25 * k;
    const double a = r * 6.283185307179586476925286766559;
    const double z = a * a;
    const double sp = 1.0 + z * (-1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 +
                      z * (-1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
                      z * (1.0 / 355687428096000.0))))))));
    const double cp = 1.0 + z * (-1.0 / 2.0 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 +
                      z * (-1.0 / 3628800.0 + z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 +
                      z * (1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0)))))))));
    const double sa = a * sp, ca = cp;
    switch (k & 3) {
    case 0: s = sa; c = ca; break;
    case 1: s = ca; c = -sa; break;
    case 2: s = -sa; c = -ca; break;
    default: s = -ca; c = sa; break;
    }
}

// Lambert diffuse direction about the incoming-side normal: azimuth 2 pi U,
// cos(zenith) = U (uniform over the hemisphere, Classes_Materials.cs:4519-4549).
// cos(acos U) is U and sin(acos U) is sqrt((1-U)(1+U)).
template <class R>
__device__ __forceinline__ void diffuse_dir(R nx, R ny, R nz, R uaz, R uzen, R &ox, R &oy, R &oz)
{
    R zx = nx, zy = ny, zz = nz;
    R xx = 0, xy = 0, xz = 1;
    const R proj = fabs(zz);
    if (R(0.99) < proj && R(1.01) > proj) { xx = 1; xz = 0; }
    // y = z cross x ; x = y cross z
    R yx = zy * xz - zz * xy, yy = zz * xx - zx * xz, yz = zx * xy - zy * xx;
    xx = yy * zz - yz * zy;
    xy = yz * zx - yx * zz;
    xz = yx * zy - yy * zx;
    normalize3(xx, xy, xz);
    normalize3(yx, yy, yz);
    normalize3(zx, zy, zz);
    R sa, ca, st;
    if constexpr (sizeof(R) == 8) {
        sincos_turn(uaz, sa, ca);
        st = sqrt((1.0 - uzen) * (1.0 + uzen));
    } else {
        // hardware sine / cosine (HIP's __sinf / __cosf: v_sin / v_cos on the
        // angle in turns), accurate to ~1e-6 for angles in [0, 2 pi)
        sa = __sinf(6.2831853071795864f * uaz);
        ca = __cosf(6.2831853071795864f * uaz);
        st = fsqrt<float>((1.0f - uzen) * (1.0f + uzen));
    }
    const R x = ca * st, y = sa * st, z = uzen;
    ox = xx * x + yx * y + zx * z;
    oy = xy * x + yy * y + zy * z;
    oz = xz * x + yz * y + zz * z;
    normalize3(ox, oy, oz);
}

// ---------------------------------------------------------------- phase 1
// src: the global source-ray index (Philox stream and geodesic face k % 80)
template <class R>
__device__ __forceinline__ void source_dir(const TraceConst &tc, uint32_t src, R &dx, R &dy, R &dz)
{
    uint32_t w[4];
    philox4x32_10(0u, 0xFFFFFFFFu, 0xFFFFFFFFu, src, tc.key0, tc.key1, w);
    double x, y, z;
    if (tc.dir_mode == 2 && tc.dirs && tc.n_dirs > 0) {
        const double *p = tc.dirs + 3 * (src % (uint64_t)tc.n_dirs);
        x = p[0]; y = p[1]; z = p[2];
    } else if (tc.dir_mode == 1) {
        double st, ct, sp, cp;  // theta, phi ~ U(0, 2 pi)
        sincos_turn(unif<double>(w[0]), st, ct);
        sincos_turn(unif<double>(w[1]), sp, cp);
        x = st * cp; y = st * sp; z = ct;
    } else {
        const double *f = tc.geo + 9 * (src % 80u);
        const double su = sqrt(unif<double>(w[0])), v = unif<double>(w[1]);
        x = (1 - su) * f[0] + su * (1 - v) * f[3] + su * v * f[6];
        y = (1 - su) * f[1] + su * (1 - v) * f[4] + su * v * f[7];
        z = (1 - su) * f[2] + su * (1 - v) * f[5] + su * v * f[8];
    }
    double l = sqrt(x * x + y * y + z * z);
    if (l > 0) { x /= l; y /= l; z /= l; }
    dx = (R)x; dy = (R)y; dz = (R)z;
}

template <class R>
__device__ __forceinline__ void make_oray(ORay<R> &r, R ox, R oy, R oz, R dx, R dy, R dz, R I, double t,
                                          int band, int tri, uint32_t src, uint64_t uid, uint32_t mask, int room)
{
    r.o[0] = ox; r.o[1] = oy; r.o[2] = oz; r.I = I;
    r.d[0] = dx; r.d[1] = dy; r.d[2] = dz; r.tri = tri;
    r.src = src; r.t = t;
    r.uid_lo = (uint32_t)uid; r.uid_hi = (uint32_t)(uid >> 32);
    r.k = 0;
    r.meta = (uint32_t)band | mask | ((uint32_t)room << M_ROOM_SHIFT);
}

// Buffer-to-buffer copy of a ray record in 16-byte pieces (a struct
// assignment between two buffers goes through a private temporary).
template <class R> __device__ __forceinline__ void copy_ray(ORay<R> *dst, const ORay<R> *src)
{
    static_assert(sizeof(ORay<R>) % 16 == 0, "ray records are whole 16-byte pieces");
    const uint4 *a = reinterpret_cast<const uint4 *>(src);
    uint4 *b = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(ORay<R>) / 16); ++i) b[i] = a[i];
}

template <class R>
__device__ __forceinline__ void store_dead(ORay<R> *out, unsigned from, unsigned to)
{
    for (unsigned q = from; q < to; ++q) out[q].meta = M_DEAD;
}

// cast_ray_distributed's broadband phase and split (TaskAll:345-424) for one
// source ray; may suspend into resume_out when the output buffer is full.
template <class R, bool MAP>
__global__ void __launch_bounds__(kBlock)
k_phase1(const TraceConst tc, const DevScene<R> s, const DevRecv<R> rv, int64_t src_begin,
         int n_fresh, const P1State<R> *__restrict__ resume_in, int n_resume, ORay<R> *out,
         unsigned *out_count, unsigned out_cap, P1State<R> *resume_out, unsigned *resume_count,
         DevCounters *cnt)
{
    PACH_COUNTERS_BEGIN(lc);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = resume_in ? n_resume : n_fresh;
    if (i < n) {
        P1State<R> st;
        if (resume_in) {
            st = resume_in[i];
        } else {
            const int64_t g = src_begin + i;
            // room of source ray g: last room whose ray_base <= g
            int lo = 0, hi = tc.n_rooms - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (tc.rooms[mid].ray_base <= g) lo = mid;
                else hi = mid - 1;
            }
            const RoomTab &rm = tc.rooms[lo];
            st.room = lo;
            st.src = (uint32_t)g;
            st.o[0] = (R)rm.src[0]; st.o[1] = (R)rm.src[1]; st.o[2] = (R)rm.src[2];
            source_dir<R>(tc, st.src, st.d[0], st.d[1], st.d[2]);
            for (int o = 0; o < kBands; ++o) st.E[o] = (R)rm.power[o];
            st.tb = 0; st.order = 0; st.tri = -1; st.k = 0; st.flags = 0;
        }
        const RoomTab &room = tc.rooms[st.room];
        const bool stoch = tc.mode == 0;
        const int excl_on = Tmin<R>::excl ? 1 : 0;
        bool split = true, suspended = false;
        if (!(st.flags & F_SPLIT_PENDING)) {
            for (;;) {
                R leg;
                uint64_t cand;
                const int ht = shoot<R, false>(s, st.o[0], st.o[1], st.o[2], st.d[0], st.d[1], st.d[2],
                                               excl_on ? st.tri : -1, leg, cand);
                if (ht < 0) {
                    lc.add(C_BROAD);
                    if (st.order > tc.is_order) {
                        const R ex = st.o[0] + st.d[0] * R(1000000), ey = st.o[1] + st.d[1] * R(1000000),
                                ez = st.o[2] + st.d[2] * R(1000000);
                        const R qx = ex - st.o[0], qy = ey - st.o[1], qz = ez - st.o[2];
                        if constexpr (MAP) map_check<R, true>(rv, st.o, st.d, ex, ey, ez, st.E, R(0), -1, st.tb, lc);
                        else
                            check_receivers<R, true>(rv, room, cand, st.o[0], st.o[1], st.o[2], st.d[0], st.d[1], st.d[2],
                                               qx * qx + qy * qy + qz * qz, st.E, R(0), -1, tc.oct_lo,
                                               tc.oct_hi, st.tb, 0, lc);
                    }
                    double l = 0;
                    for (int j = 0; j < kBands; ++j) l += (double)st.E[j];
                    lc.lose(l);
                    lc.add(C_MISS);
                    split = false;
                    break;
                }
                const int sf = s.surf[ht];
                unsigned nch = 0;
                if (stoch) {
                    for (int oc = 0; oc < kBands; ++oc) {
                        const R e2 = st.E[oc] * s.mat[sf * kBands + oc].x;
                        if (e2 != R(0)) nch += (s.trans[sf * kBands + oc] > R(0)) ? 3u : 1u;
                    }
                }
                unsigned base = 0;
                if (nch) {
                    base = atomicAdd(out_count, nch);
                    if (base + nch > out_cap) {
                        if (base < out_cap) store_dead<R>(out, base, out_cap);
                        resume_out[atomicAdd(resume_count, 1u)] = st;
                        lc.add(C_RESUME);
                        suspended = true;
                        break;
                    }
                }
                lc.add(C_BROAD);
                const R hx = st.o[0] + st.d[0] * leg, hy = st.o[1] + st.d[1] * leg, hz = st.o[2] + st.d[2] * leg;
                if (st.order > tc.is_order) {
                    const R qx = hx - st.o[0], qy = hy - st.o[1], qz = hz - st.o[2];
                    if constexpr (MAP) map_check<R, true>(rv, st.o, st.d, hx, hy, hz, st.E, R(0), -1, st.tb, lc);
                    else
                        check_receivers<R, true>(rv, room, cand, st.o[0], st.o[1], st.o[2], st.d[0], st.d[1], st.d[2],
                                           qx * qx + qy * qy + qz * qz, st.E, R(0), -1, tc.oct_lo, tc.oct_hi,
                                           st.tb, 0, lc);
                }
                st.tb += (double)leg / rv.c;
                st.o[0] = hx; st.o[1] = hy; st.o[2] = hz;
                st.order++;
                st.tri = ht;
                const auto N = s.nrm[ht];
                const R cos_t = st.d[0] * N.x + st.d[1] * N.y + st.d[2] * N.z;
                for (int o = 0; o < kBands; ++o) st.E[o] *= s.mat[sf * kBands + o].x;
                if (stoch) {
                    // Scatter_Early (Classes_Materials.cs:4498-4570)
                    R fx = N.x, fy = N.y, fz = N.z, cf = cos_t;
                    if (cf > R(0)) { fx = -fx; fy = -fy; fz = -fz; cf = -cf; }
                    unsigned q = base;
                    for (int oc = 0; oc < kBands; ++oc) {
                        if (st.E[oc] == R(0)) continue;
                        uint32_t w[4];
                        philox4x32_10(4u * st.k + (uint32_t)(oc >> 1), 0u, 0u, st.src, tc.key0, tc.key1, w);
                        const R sc = s.mat[sf * kBands + oc].y;
                        R childI = st.E[oc] * sc;
                        st.E[oc] *= (R(1) - sc);
                        R ddx, ddy, ddz;
                        diffuse_dir<R>(fx, fy, fz, unif<R>(w[2 * (oc & 1)]), unif<R>(w[2 * (oc & 1) + 1]),
                                       ddx, ddy, ddz);
                        const R tau = s.trans[sf * kBands + oc];
                        if (tau > R(0)) {
                            ORay<R> tr;
                            make_oray<R>(tr, hx, hy, hz, st.d[0], st.d[1], st.d[2], st.E[oc] * tau, st.tb,
                                         oc, ht, st.src, child_uid(0, PACH_TAG_EARLY_TR(st.k, oc)), 0u, st.room);
                            st.E[oc] *= (R(1) - tau);
                            out[q++] = tr;
                            ORay<R> td;
                            make_oray<R>(td, hx, hy, hz, -ddx, -ddy, -ddz, childI * tau, st.tb, oc, ht,
                                         st.src, child_uid(0, PACH_TAG_EARLY_TD(st.k, oc)), M_SCAT | M_EVER, st.room);
                            childI *= (R(1) - tau);
                            out[q++] = td;
                            lc.add(C_SPAWN, 2);
                            lc.add(C_OCT, 2);
                        }
                        ORay<R> ch;
                        make_oray<R>(ch, hx, hy, hz, ddx, ddy, ddz, childI, st.tb, oc, ht, st.src,
                                     child_uid(0, PACH_TAG_EARLY(st.k, oc)), M_SCAT | M_EVER, st.room);
                        out[q++] = ch;
                        lc.add(C_SPAWN);
                        lc.add(C_OCT);
                    }
                    st.d[0] -= fx * cf * R(2);
                    st.d[1] -= fy * cf * R(2);
                    st.d[2] -= fz * cf * R(2);
                } else {
                    st.d[0] -= N.x * cos_t * R(2);
                    st.d[1] -= N.y * cos_t * R(2);
                    st.d[2] -= N.z * cos_t * R(2);
                }
                st.k++;
                if (!((double)st.E[room.h] > room.T1)) break;
                if (st.order >= tc.p1_max) { lc.add(C_P1CAP); break; }
            }
        }
        if (split && !suspended) {
            const unsigned nb = (unsigned)(tc.oct_hi - tc.oct_lo + 1);
            const unsigned base = atomicAdd(out_count, nb);
            if (base + nb > out_cap) {
                if (base < out_cap) store_dead<R>(out, base, out_cap);
                st.flags |= F_SPLIT_PENDING;
                resume_out[atomicAdd(resume_count, 1u)] = st;
                lc.add(C_RESUME);
            } else {
                // BroadRay.SplitRay(o) per octave (Classes_Rays.cs:331-337)
                for (int oc = tc.oct_lo; oc <= tc.oct_hi; ++oc) {
                    ORay<R> r;
                    make_oray<R>(r, st.o[0], st.o[1], st.o[2], st.d[0], st.d[1], st.d[2], st.E[oc], st.tb,
                                 oc, st.tri, st.src, child_uid(0, PACH_TAG_SPLIT(oc)), 0u, st.room);
                    out[base + (oc - tc.oct_lo)] = r;
                    lc.add(C_OCT);
                }
            }
        }
    }
    counters_end(lc, cnt);
}

// ---------------------------------------------------------------- phases 2, 3
// Everything an octave-ray bounce does after its closest-hit walk: receiver
// checks, Absorb, Scatter_Late / Scatter_VeryLate, the phase hand-off.
// Returns false when the ray terminates.  A specular child Scatter_Late
// spawns is handed to spawn(child) where it is made (so it need not be held
// through the diffuse reflection that follows); in the octave kernel only the
// lanes that spawn execute that call.
template <class R, bool MAP, class Spawn>
__device__ __forceinline__ bool octave_shade(const TraceConst &tc, const DevScene<R> &s, const DevRecv<R> &rv,
                                             ORay<R> &r, int ht, R leg, uint64_t cand, Spawn &&spawn,
                                             LaneCounters &lc)
{
    const int band = (int)(r.meta & M_BAND);
    const bool late3 = (r.meta & M_LATE3) != 0;
    const RoomTab &room = tc.rooms[r.meta >> M_ROOM_SHIFT];
    const int ana = (r.meta & M_SCAT) ? 1 : ((r.meta & M_EVER) ? 2 : 0);
    lc.add(late3 ? C_VLATE : C_LATE);
    const bool miss = ht < 0;
    // the surface's data is read first, so its loads overlap the receiver
    // checks (a miss reads triangle 0's and does not use it)
    const int hs = miss ? 0 : ht;
    const int sf = s.surf[hs];
    const auto N = s.nrm[hs];
    const auto M = s.mat[sf * kBands + band];  // (1 - alpha, s, lo, hi)
    const R tau = tc.has_trans ? s.trans[sf * kBands + band] : R(0);
    // the segment checked against the receivers ends at the hit, or 1e6 m out
    // for a ray that leaves the scene; one call site for both keeps the
    // kernel's code small
    const R lg = miss ? R(1000000) : leg;
    const R hx = r.o[0] + r.d[0] * lg, hy = r.o[1] + r.d[1] * lg, hz = r.o[2] + r.d[2] * lg;
    {
        const R qx = hx - r.o[0], qy = hy - r.o[1], qz = hz - r.o[2];
        if constexpr (MAP) map_check<R, false>(rv, r.o, r.d, hx, hy, hz, nullptr, r.I, band, r.t, lc);
        else
            check_receivers<R, false>(rv, room, cand, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2],
                               qx * qx + qy * qy + qz * qz, nullptr, r.I, band, 0, 0, r.t, ana, lc);
    }
    if (miss) {
        lc.lose((double)r.I);
        lc.add(C_MISS);
        return false;
    }
    // the FP64 path divides as the reference does; the FP32 path's legs carry
    // FP32 rounding already, and a multiply saves the FP64 division
    if constexpr (sizeof(R) == 8) r.t += (double)leg / rv.c;
    else r.t += (double)leg * rv.inv_c;
    r.o[0] = hx; r.o[1] = hy; r.o[2] = hz;
    r.tri = ht;
    const R cos_t = N.x * r.d[0] + N.y * r.d[1] + N.z * r.d[2];
    r.I *= M.x;
    const uint32_t kev = r.k++;
    if (tc.mode == 0) {
        uint32_t w[4];
        philox4x32_10(kev, r.uid_lo, r.uid_hi, r.src, tc.key0, tc.key1, w);
        const bool transmit = tc.has_trans && unif<R>(w[0]) < tau;
        bool diffuse;
        if (!late3) {
            // Scatter_Late (Classes_Materials.cs:4629-4702)
            const R sel = unif<R>(w[1]);
            if (sel > M.w) {
                diffuse = false;
            } else {
                diffuse = true;
                if (sel > M.z) {
                    const R emc = R(1) - M.y;
                    ORay<R> child = r;
                    child.I = r.I * emc;
                    r.I *= (R(1) - emc);
                    child.d[0] -= N.x * cos_t * R(2);
                    child.d[1] -= N.y * cos_t * R(2);
                    child.d[2] -= N.z * cos_t * R(2);
                    if (transmit) { child.d[0] = -child.d[0]; child.d[1] = -child.d[1]; child.d[2] = -child.d[2]; }
                    child.meta &= ~M_SCAT;
                    const uint64_t cu = child_uid(((uint64_t)r.uid_hi << 32) | r.uid_lo, PACH_TAG_LATE(kev));
                    child.uid_lo = (uint32_t)cu;
                    child.uid_hi = (uint32_t)(cu >> 32);
                    child.k = 0;
                    spawn(child);
                    lc.add(C_SPAWN);
                    lc.add(C_OCT);
                }
            }
        } else {
            // Scatter_VeryLate (Classes_Materials.cs:4572-4627)
            diffuse = unif<R>(w[1]) < M.y;
        }
        if (diffuse) {
            R fx = N.x, fy = N.y, fz = N.z;
            if (cos_t > R(0)) { fx = -fx; fy = -fy; fz = -fz; }
            diffuse_dir<R>(fx, fy, fz, unif<R>(w[2]), unif<R>(w[3]), r.d[0], r.d[1], r.d[2]);
            r.meta |= (M_SCAT | M_EVER);
        } else {
            r.d[0] -= N.x * cos_t * R(2);
            r.d[1] -= N.y * cos_t * R(2);
            r.d[2] -= N.z * cos_t * R(2);
            r.meta &= ~M_SCAT;
        }
        if (transmit) { r.d[0] = -r.d[0]; r.d[1] = -r.d[1]; r.d[2] = -r.d[2]; }
    } else {
        r.d[0] -= N.x * cos_t * R(2);
        r.d[1] -= N.y * cos_t * R(2);
        r.d[2] -= N.z * cos_t * R(2);
    }
    if (!late3) {
        if (!(r.t < tc.cotime && (double)r.I > room.T2[band])) r.meta |= M_LATE3;  // hand-off to End_Trace
        return true;
    }
    return r.t < tc.cotime && (double)r.I > room.T3[band];
}

// Exact queue of the mixed kernel (unused by the others).
template <class R> struct ExactQ {
    ORay<R> *rays;
    unsigned *count;
    unsigned cap;
    unsigned drain_at;  // a wave whose push ends past this drains (every wave can push <= 128 more)
};

#ifdef PACH_MIX_UNIT
// One bounce by the FP64 walk for every ray in `in` (the bounces the mixed
// kernel's FP32 screen could not decide): the FP64 closest hit (shoot, t >
// 1e-10), then the bounce's shading; the continuation and any spawned ray go
// to `out` (the next launch's input).  Grid-stride over *n_in rays.
__global__ void __launch_bounds__(kBlock)
k_exact(const TraceConst tc, const DevScene<double> s, const DevRecv<double> rv, const ORay<double> *__restrict__ in,
        const unsigned *n_in, unsigned cap_in, ORay<double> *out, unsigned *out_count, unsigned out_cap,
        DevCounters *cnt)
{
    PACH_COUNTERS_BEGIN(lc);
    const unsigned n = min(*n_in, cap_in);
    auto put = [&](const ORay<double> &ray) __attribute__((always_inline)) {
        const unsigned q = atomicAdd(out_count, 1u);
        if (q < out_cap) out[q] = ray;
        else { lc.add(C_DROP); lc.lose((double)ray.I); }
    };
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        ORay<double> r = in[i];
        double tb = 0;
        uint64_t cand = 0;
        const int ht = shoot<double, false>(s, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2], -1, tb, cand);
        lc.add(C_SHADED);
        if (octave_shade<double, false>(tc, s, rv, r, ht, tb, cand, put, lc)) put(r);
    }
    counters_end(lc, cnt);
}
#endif

// Persistent, wavefront-compacting octave-ray kernel.
//
// Each lane holds one octave ray.  A loop step (a) refills idle lanes (once
// refill_min of them are idle), first from the wave's own LIFO stack of
// spawned rays (an HBM slab per wave), then from the shared input (the
// previous launch's leftovers in0, then the new broadband-phase output in1);
// (b) advances every walking lane's closest-hit walk by walk_steps units;
// (c) shades the lanes whose walk ended, but only once shade_min lanes are
// waiting or none is still walking, so the shading code runs on mostly-full
// waves while long walks continue.
// Once the input is drained and the wave has run min_iters steps it drains:
// every step it moves its stack and the lanes that are between bounces to the
// shared `dump` queue, while walks already under way run to their hit and are
// shaded (a walk is never cut short, so every launch completes at least the
// bounces its waves started and the relaunch sequence always advances).  The
// wave exits when no lane is left; the next launch spreads the leftover work
// over every wave.  No wave ever waits on another.

// BS: workgroup size (waves share the LDS copy of the distance field);
// OCC: 0 = compiler's register allocation, else waves per SIMD to fit
// (a register cap of 512 / OCC VGPRs).
// SQ: shade queue.  A lane whose walk ended writes (ray, hit) to its wave's
// queue (kShadeQ items in HBM/L2) and refills at once; whenever 64 items wait
// (or no lane is walking) every lane shades one, so the shading code runs on
// full waves, and continuations and spawned rays go onto the wave's stack.
// MIX (R = double): mixed walk -- the FP32 walk over sg screens candidates,
// FP64 decides every hit and shades (the results of the FP64 path).
template <class R, bool LDS, int BS, int OCC, bool SQ, bool MAP, bool MIX = false>
__global__ void __launch_bounds__(BS, OCC > 0 ? OCC : 1)
k_octave(const TraceConst tc, const DevScene<R> s, const DevScene<float> sg, const DevRecv<R> rv,
         const ORay<R> *__restrict__ in0, const unsigned *n0_ptr, unsigned cap0,
         const ORay<R> *__restrict__ in1, const unsigned *n1_ptr, unsigned cap1, unsigned *in_head,
         ORay<R> *stacks, int stack_cap, ShadeItem<R> *shq, ORay<R> *dump, unsigned *dump_count,
         unsigned dump_cap, int min_iters, DevCounters *cnt, ExactQ<R> xq)
{
    if constexpr (LDS) {
        // stage the voxel distances (ncell rounded up to 16 bytes) in LDS
        const uint4 *src = (const uint4 *)s.cell_dist;
        uint4 *dst = (uint4 *)pach_lds_dist;
        const int n16 = (s.ncell + 15) >> 4;
        for (int i = threadIdx.x; i < n16; i += BS) dst[i] = src[i];
    }
    PACH_COUNTERS_BEGIN(lc);
    unsigned shades = 0;  // shading passes of this wave
    const int lane = (int)__lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wave = (uint64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
    ORay<R> *ws = stacks + wave * (uint64_t)stack_cap;
    ShadeItem<R> *wq = SQ ? shq + wave * (uint64_t)kShadeQ : nullptr;
    int qn = 0;  // items in the wave's shade queue
    const unsigned n0 = in0 ? min(*n0_ptr, cap0) : 0u;
    const unsigned n1 = in1 ? min(*n1_ptr, cap1) : 0u;
    const unsigned n_in = n0 + n1;
    int stack_n = 0;
    int iters = 0;
    bool input_done = false, draining = false;
    static_assert(!MIX || (sizeof(R) == 8 && !SQ && !MAP), "mixed walk: FP64 rays, no shade queue, no map");
    using G = typename std::conditional<MIX, float, R>::type;
    bool act = false, walking = false, ready = false;
    ORay<R> r;
    Walk<G> w;
    G go[3], gd[3];  // MIX: the ray in FP32 for the walk
    // MIX: a bounce the FP32 screen cannot decide goes to the exact queue (the
    // FP64 walk of k_exact, after this launch); a wave whose push lands past
    // xq.drain_at drains, so the queue never overflows
    auto to_exact = [&](bool give) __attribute__((always_inline)) {
        const uint64_t gm = __ballot(give);
        if (!gm) return;
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(xq.count, (unsigned)__popcll(gm));
        base = __shfl(base, 0);
        if (give) {
            const unsigned q = base + (unsigned)__popcll(gm & lt_mask);
            if (q < xq.cap) xq.rays[q] = r;
            else { lc.add(C_DROP); lc.lose((double)r.I); }
            act = false;
        }
        if (base + (unsigned)__popcll(gm) > xq.drain_at) draining = true;
    };
    for (;;) {
        // ---- (a) refill idle lanes once refill_min of them are idle (or all
        // are): own stack first, then the input queue
        const uint64_t empty = __ballot(!act);
        const int n_empty = __popcll(empty);
        if (!draining && n_empty > 0 && (n_empty >= tc.refill_min || n_empty == 64)) {
            const int rank = __popcll(empty & lt_mask);
            const int from_stack = min(n_empty, stack_n);
            if (!act && rank < from_stack) {
                r = ws[stack_n - 1 - rank];
                act = true;
            }
            stack_n -= from_stack;
            const int need = n_empty - from_stack;
            if (need > 0 && !input_done) {
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(in_head, (unsigned)need);
                base = __shfl(base, 0);
                const int avail = base >= n_in ? 0 : (int)min((unsigned)need, n_in - base);
                if (avail < need) input_done = true;
                const int j = rank - from_stack;
                if (!act && j >= 0 && j < avail) {
                    const unsigned q = base + (unsigned)j;
                    const ORay<R> *src = q < n0 ? in0 + q : in1 + (q - n0);
                    r = *src;
                    act = (r.meta & M_DEAD) == 0;
                }
            }
        }
        if (input_done && iters >= min_iters) draining = true;
        if (draining) {
            // ---- hand the stack and the lanes between bounces to the next launch
            const bool give = act && !walking && !ready;
            const uint64_t gm = __ballot(give);
            const unsigned n_dump = (unsigned)stack_n + (unsigned)__popcll(gm);
            if (n_dump > 0) {
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(dump_count, n_dump);
                base = __shfl(base, 0);
                for (int i = lane; i < stack_n; i += 64) {
                    const unsigned q = base + (unsigned)i;
                    if (q < dump_cap) copy_ray(dump + q, ws + i);
                    else { lc.add(C_DROP); lc.lose((double)ws[i].I); }
                }
                if (give) {
                    const unsigned q = base + (unsigned)stack_n + (unsigned)__popcll(gm & lt_mask);
                    if (q < dump_cap) dump[q] = r;
                    else { lc.add(C_DROP); lc.lose((double)r.I); }
                    act = false;
                }
                if (lane == 0) lc.add(C_SPILL, n_dump);
                stack_n = 0;
            }
            if (!input_done) {
                // drained before the input ran out (the exact queue filled):
                // hand the unclaimed input to the next launch as well
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(in_head, 64u);
                base = __shfl(base, 0);
                const int avail = base >= n_in ? 0 : (int)min(64u, n_in - base);
                if (avail < 64) input_done = true;
                bool mv = false;
                const ORay<R> *src = nullptr;
                if (lane < avail) {
                    const unsigned q = base + (unsigned)lane;
                    src = q < n0 ? in0 + q : in1 + (q - n0);
                    mv = (src->meta & M_DEAD) == 0;
                }
                const uint64_t mm = __ballot(mv);
                if (mm) {
                    unsigned db = 0;
                    if (lane == 0) db = atomicAdd(dump_count, (unsigned)__popcll(mm));
                    db = __shfl(db, 0);
                    if (mv) {
                        const unsigned q = db + (unsigned)__popcll(mm & lt_mask);
                        if (q < dump_cap) copy_ray(dump + q, src);
                        else { lc.add(C_DROP); lc.lose((double)src->I); }
                    }
                    if (lane == 0) lc.add(C_SPILL, (unsigned)__popcll(mm));
                }
            }
        }
        const uint64_t am = __ballot(act);
        if (am == 0 && qn == 0) {
            if (input_done) break;
            continue;  // the claim returned only dead slots: claim again
        }
        // ---- (b) start and advance closest-hit walks
        bool grazing = false;
        if (act && !walking && !ready) {
            if ((int)r.k >= tc.max_bounces) {
                lc.add(C_GUARD);
                act = false;
            } else {
                if constexpr (MIX) {
                    for (int k = 0; k < 3; ++k) {
                        go[k] = (float)r.o[k];
                        gd[k] = (float)r.d[k];
                    }
                    // a grazing departure: FP64 may hit the triangle left again
                    if (r.tri >= 0) {
                        const float4 n = sg.nrm[r.tri];
                        grazing = fabsf(n.x * gd[0] + n.y * gd[1] + n.z * gd[2]) < tc.mx_self;
                    }
                    if (grazing) lc.add(C_EXACT0);
                    else mix_walk_init(sg, w, go, gd);
                } else {
                    walk_init<R>(s, w, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2]);
                }
                walking = !grazing;
            }
        }
        if constexpr (MIX) to_exact(grazing);
        if (walking) {
            const int ex = Tmin<R>::excl ? r.tri : -1;
            bool done;
            if constexpr (MIX)
                done = tc.walk_flat == 2
                           ? walk_units_mix<LDS, true>(sg, tc, w, go[0], go[1], go[2], gd[0], gd[1], gd[2], r.tri,
                                                       tc.walk_steps, lc.vox, lc.test)
                           : walk_units_mix<LDS, false>(sg, tc, w, go[0], go[1], go[2], gd[0], gd[1], gd[2], r.tri,
                                                        tc.walk_steps, lc.vox, lc.test);
            else
            done =
                tc.walk_flat == 2
                    ? walk_units<R, LDS, true>(s, w, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2], ex, tc.walk_steps,
                                               lc.vox, lc.test)
                : tc.walk_flat
                    ? walk_units<R, LDS>(s, w, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2], ex, tc.walk_steps,
                                         lc.vox, lc.test)
                    : walk_steps<R, LDS>(s, w, r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2], ex, tc.walk_steps,
                                         lc.vox, lc.test);
            if (done) {
                walking = false;
                ready = true;
            }
        }
        ++iters;
        // push a ray onto the wave's stack (overflow goes to the dump queue)
        auto push = [&](bool has, const ORay<R> &ray) __attribute__((always_inline)) {
            const uint64_t cm = __ballot(has);
            if (!cm) return;
            if (has) {
                const int pos = stack_n + __popcll(cm & lt_mask);
                if (pos < stack_cap) {
                    ws[pos] = ray;
                } else {
                    const unsigned q = atomicAdd(dump_count, 1u);
                    if (q < dump_cap) { dump[q] = ray; lc.add(C_SPILL); }
                    else { lc.add(C_DROP); lc.lose((double)ray.I); }
                }
            }
            stack_n = min(stack_n + __popcll(cm), stack_cap);
        };
        // a spawned child goes onto the stack where octave_shade makes it: only
        // the spawning lanes run this, so the active mask orders them; the
        // caller then advances stack_n by their count
        bool spawned = false;
        auto spawn = [&](const ORay<R> &ray) __attribute__((always_inline)) {
            spawned = true;
            const int pos = stack_n + __popcll(__ballot(true) & lt_mask);
            if (pos < stack_cap) {
                ws[pos] = ray;
            } else {
                const unsigned q = atomicAdd(dump_count, 1u);
                if (q < dump_cap) { dump[q] = ray; lc.add(C_SPILL); }
                else { lc.add(C_DROP); lc.lose((double)ray.I); }
            }
        };
        auto spawned_done = [&]() __attribute__((always_inline)) {
            const int n_sp = __popcll(__ballot(spawned));
            stack_n = min(stack_n + n_sp, stack_cap);
            spawned = false;
            return n_sp;
        };
        if constexpr (SQ) {
            // ---- (c) queue the finished walks, shade 64 at a time
            const uint64_t fm = __ballot(ready);
            if (fm) {
                if (ready) {
                    ShadeItem<R> it;
                    it.r = r;
                    it.best = w.best;
                    it.tbest = w.tbest;
                    it.rmask = w.rmask;
                    wq[qn + __popcll(fm & lt_mask)] = it;
                    act = false;
                    ready = false;
                }
                qn += __popcll(fm);
                __threadfence_block();
            }
            if (qn >= 64 || (qn > 0 && __ballot(walking) == 0)) {
                const int k = min(qn, 64);
                bool alive = false;
                ORay<R> sr;
                ++shades;
                if (lane < k) {
                    const ShadeItem<R> it = wq[qn - k + lane];
                    sr = it.r;
                    lc.add(C_SHADED);
                    alive = octave_shade<R, MAP>(tc, s, rv, sr, it.best, it.tbest, it.rmask, spawn, lc);
                }
                qn -= k;
                spawned_done();
                push(alive, sr);  // on top: the continuation is walked next
                __threadfence_block();
            }
        } else {
            // ---- (c) shade the lanes whose walk ended
            const int n_ready = __popcll(__ballot(ready));
            if (n_ready > 0 && (n_ready >= tc.shade_min || __ballot(walking) == 0)) {
                ++shades;
                bool unsure = false;
                if (ready) {
                    R tb = (R)w.tbest;
                    int hb = w.best;
                    if constexpr (MIX) {
                        double t64 = 0;
                        unsure = !mix_confirm(s, w, r.o, r.d, hb, t64);
                        tb = t64;
                    }
                    ready = false;
                    if (unsure) {
                        lc.add(C_EXACT);
                    } else {
                        lc.add(C_SHADED);
                        act = octave_shade<R, MAP>(tc, s, rv, r, hb, tb, w.rmask, spawn, lc);
                    }
                }
                if (spawned_done()) __threadfence_block();
                if constexpr (MIX) to_exact(unsure);
            }
        }
    }
    if (lane == 0) {
        lc.add(C_ITER, (unsigned)iters);
        lc.add(C_SHADE, shades);
    }
    counters_end(lc, cnt);
}

// ---------------------------------------------------------------- closest-hit queries
// Polygon_Scene.shoot (Classes_Scenes.cs:757-819) for a batch of rays, one
// thread per ray: the triangle hit (-1: none) and its distance.  FP64: the FP64
// walk; FP32: the octave kernel's FP32 walk; MIX: the mixed walk, confirmed
// in FP64 or walked again in exact mode (counted in *n_exact).
template <class R, bool MIX>
__global__ void __launch_bounds__(kBlock)
k_shoot(const TraceConst tc, const DevScene<R> s, const DevScene<float> sg, const double *__restrict__ o,
        const double *__restrict__ d, int64_t n, int *__restrict__ tri, double *__restrict__ tout, unsigned *n_exact)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double O[3] = {o[3 * i], o[3 * i + 1], o[3 * i + 2]};
    const double D[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    uint32_t nv = 0, nt = 0;
    if constexpr (MIX) {
        const float go[3] = {(float)O[0], (float)O[1], (float)O[2]};
        const float gd[3] = {(float)D[0], (float)D[1], (float)D[2]};
        Walk<float> w;
        double t64 = 0;
        int hit = -1;
        mix_walk_init(sg, w, go, gd);
        while (!walk_units_mix<false, false>(sg, tc, w, go[0], go[1], go[2], gd[0], gd[1], gd[2], -1, 64, nv, nt)) {
        }
        if (!mix_confirm(s, w, O, D, hit, t64)) {
            uint64_t rm = 0;
            hit = shoot<double, false>(s, O[0], O[1], O[2], D[0], D[1], D[2], -1, t64, rm);
            atomicAdd(n_exact, 1u);
        }
        tri[i] = hit;
        tout[i] = hit >= 0 ? t64 : 0.0;
    } else {
        const R ox = (R)O[0], oy = (R)O[1], oz = (R)O[2], dx = (R)D[0], dy = (R)D[1], dz = (R)D[2];
        Walk<R> w;
        walk_init<R>(s, w, ox, oy, oz, dx, dy, dz);
        while (!walk_units<R, false>(s, w, ox, oy, oz, dx, dy, dz, -1, 64, nv, nt)) {
        }
        tri[i] = w.best;
        tout[i] = w.best >= 0 ? (double)w.tbest : 0.0;
    }
}

template <class R>
hipError_t launch_shoot(hipStream_t st, const TraceConst &tc, const DevScene<R> &s, const DevScene<float> &sg,
                        const double *o, const double *d, int64_t n, int *tri, double *t, unsigned *n_exact)
{
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_shoot<R, false>), dim3((unsigned)grid), dim3(kBlock), 0, st, tc, s, sg, o, d, n, tri, t,
                       n_exact);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
// Register cap of the mapping-receiver variant: the compiler's allocation
// (254 VGPRs, two waves per SIMD); a cap at three waves spills 232 bytes per
// lane and runs 8 % slower on the hall's 1,184-receiver map.
constexpr int kMapOcc = 0;
template <class R>
hipError_t launch_phase1(hipStream_t st, const TraceConst &tc, const DevScene<R> &s,
                         const DevRecv<R> &rv, int64_t src_begin, int n_fresh,
                         const P1State<R> *resume_in, int n_resume, ORay<R> *out,
                         unsigned *out_count, unsigned out_cap, P1State<R> *resume_out,
                         unsigned *resume_count, DevCounters *cnt)
{
    const int n = resume_in ? n_resume : n_fresh;
    if (n <= 0) return hipSuccess;
    const int grid = (n + kBlock - 1) / kBlock;
    if (rv.has_map)
        hipLaunchKernelGGL((k_phase1<R, true>), dim3(grid), dim3(kBlock), 0, st, tc, s, rv, src_begin, n_fresh,
                           resume_in, n_resume, out, out_count, out_cap, resume_out, resume_count, cnt);
    else
        hipLaunchKernelGGL((k_phase1<R, false>), dim3(grid), dim3(kBlock), 0, st, tc, s, rv, src_begin, n_fresh,
                           resume_in, n_resume, out, out_count, out_cap, resume_out, resume_count, cnt);
    return hipGetLastError();
}

// A source ray is complete once its broadband phase and every octave ray it
// spawned have ended (TaskAll:274-312 advances the completed count per source
// ray): between launches the live ones are the sources of the rays handed to
// the next launch.  One bit per source ray; each newly set bit counts once.
template <class R>
__global__ void __launch_bounds__(256) k_live_sources(const ORay<R> *__restrict__ rays, const unsigned *n_rays,
                                                      unsigned cap, const P1State<R> *__restrict__ res,
                                                      const unsigned *n_res, unsigned res_cap, uint32_t first,
                                                      uint32_t span, uint32_t *bits, unsigned *live)
{
    const unsigned nr = min(*n_rays, cap), ns = min(*n_res, res_cap);
    const size_t n = (size_t)nr + ns;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        uint32_t src;
        bool alive = true;
        if (i < nr) {
            alive = !(rays[i].meta & M_DEAD);
            src = rays[i].src;
        } else {
            src = res[i - nr].src;
        }
        const uint32_t k = src - first;
        if (alive && k < span) {
            const uint32_t m = 1u << (k & 31u);
            if (!(atomicOr(&bits[k >> 5], m) & m)) atomicAdd(live, 1u);
        }
    }
}

template <class R>
hipError_t launch_live_sources(hipStream_t st, const ORay<R> *rays, const unsigned *n_rays, unsigned cap,
                               const P1State<R> *res, const unsigned *n_res, unsigned res_cap, uint32_t first,
                               uint32_t span, uint32_t *bits, unsigned *live)
{
    hipLaunchKernelGGL(k_live_sources<R>, dim3(1024), dim3(256), 0, st, rays, n_rays, cap, res, n_res, res_cap,
                       first, span, bits, live);
    return hipGetLastError();
}

template <class R, int BS, int OCC, bool SQ, bool MAP = false, bool MIX = false>
static void launch_octave_bs(hipStream_t st, int grid, size_t lds, const TraceConst &tc, const DevScene<R> &s,
                             const DevScene<float> &sg, const DevRecv<R> &rv, const ORay<R> *in0, const unsigned *n0,
                             unsigned cap0, const ORay<R> *in1, const unsigned *n1, unsigned cap1, unsigned *in_head,
                             ORay<R> *stacks, int stack_cap, ShadeItem<R> *shq, ORay<R> *dump,
                             unsigned *dump_count, unsigned dump_cap, int min_iters, DevCounters *cnt,
                             ExactQ<R> xq = ExactQ<R>{})
{
    if (lds)
        hipLaunchKernelGGL((k_octave<R, true, BS, OCC, SQ, MAP, MIX>), dim3(grid), dim3(BS), lds, st, tc, s, sg, rv,
                           in0, n0, cap0, in1, n1, cap1, in_head, stacks, stack_cap, shq, dump, dump_count, dump_cap,
                           min_iters, cnt, xq);
    else
        hipLaunchKernelGGL((k_octave<R, false, BS, OCC, SQ, MAP, MIX>), dim3(grid), dim3(BS), 0, st, tc, s, sg, rv,
                           in0, n0, cap0, in1, n1, cap1, in_head, stacks, stack_cap, shq, dump, dump_count, dump_cap,
                           min_iters, cnt, xq);
}

template <class R>
hipError_t launch_octave(hipStream_t st, int grid, const OctShape &sh, const TraceConst &tc, const DevScene<R> &s,
                         const DevScene<float> &sg, const DevRecv<R> &rv, const ORay<R> *in0, const unsigned *n0,
                         unsigned cap0, const ORay<R> *in1, const unsigned *n1, unsigned cap1, unsigned *in_head,
                         ORay<R> *stacks, int stack_cap, ShadeItem<R> *shq, ORay<R> *dump, unsigned *dump_count,
                         unsigned dump_cap, int min_iters, DevCounters *cnt)
{
    const size_t lds = s.dist_lds ? (size_t)((s.ncell + 15) & ~15) : 0;
#define PACH_OCT(BS_, OCC_, SQ_)                                                                               \
    launch_octave_bs<R, BS_, OCC_, SQ_>(st, grid, lds, tc, s, sg, rv, in0, n0, cap0, in1, n1, cap1, in_head,    \
                                        stacks, stack_cap, shq, dump, dump_count, dump_cap, min_iters, cnt)
#define PACH_OCT_Q(BS_, OCC_) \
    { if (sh.queue) PACH_OCT(BS_, OCC_, true); else PACH_OCT(BS_, OCC_, false); }
#define PACH_OCT_O(BS_)                                  \
    {                                                    \
        if (sh.occ == 4) PACH_OCT_Q(BS_, 4)              \
        else if (sh.occ == 5) PACH_OCT_Q(BS_, 5)         \
        else if (sh.occ == 3) PACH_OCT_Q(BS_, 3)         \
        else PACH_OCT_Q(BS_, 0)                          \
    }
    if (sh.queue && !shq) return hipErrorInvalidValue;
    if (rv.has_map) {
        // mapping receivers: workgroups of 256, no shade queue, kMapOcc (a
        // 128-register cap spills the map walk: 26.0 s against 18.3 s on the
        // hall's 1,184-receiver map); the grid comes from
        // octave_map_blocks_per_cu, this kernel's own residency
        launch_octave_bs<R, 256, kMapOcc, false, true>(st, grid, lds, tc, s, sg, rv, in0, n0, cap0, in1, n1, cap1,
                                                       in_head, stacks, stack_cap, shq, dump, dump_count, dump_cap,
                                                       min_iters, cnt);
        return hipGetLastError();
    }
    if (sh.block == 512) PACH_OCT_O(512)
    else PACH_OCT_O(256)
#undef PACH_OCT_O
#undef PACH_OCT_Q
#undef PACH_OCT
    return hipGetLastError();
}

#ifdef PACH_MIX_UNIT
// Mixed walk (FP64 rays, FP32 screen): workgroup size and register cap as
// OctShape says (occupancy 3, 4 or the compiler's; no shade queue, no mapping
// receivers).
hipError_t launch_octave_mix(hipStream_t st, int grid, const OctShape &sh, const TraceConst &tc,
                             const DevScene<double> &s, const DevScene<float> &sg, const DevRecv<double> &rv,
                             const ORay<double> *in0, const unsigned *n0, unsigned cap0, const ORay<double> *in1,
                             const unsigned *n1, unsigned cap1, unsigned *in_head, ORay<double> *stacks,
                             int stack_cap, ORay<double> *dump, unsigned *dump_count, unsigned dump_cap,
                             int min_iters, DevCounters *cnt, ORay<double> *xq, unsigned *xq_count, unsigned xq_cap)
{
    if (sh.queue || rv.has_map) return hipErrorInvalidValue;
    const size_t lds = s.dist_lds ? (size_t)((s.ncell + 15) & ~15) : 0;
    const unsigned waves = (unsigned)grid * (unsigned)(sh.block / 64);
    if (xq_cap <= 128u * waves) return hipErrorInvalidValue;
    const ExactQ<double> q{xq, xq_count, xq_cap, xq_cap - 128u * waves};
#define PACH_MIX(BS_, OCC_)                                                                                        \
    launch_octave_bs<double, BS_, OCC_, false, false, true>(st, grid, lds, tc, s, sg, rv, in0, n0, cap0, in1, n1,     \
                                                            cap1, in_head, stacks, stack_cap, nullptr, dump,          \
                                                            dump_count, dump_cap, min_iters, cnt, q)
#define PACH_MIX_O(BS_)                                     \
    {                                                       \
        if (sh.occ == 4) PACH_MIX(BS_, 4);                  \
        else if (sh.occ == 3) PACH_MIX(BS_, 3);             \
        else PACH_MIX(BS_, 0);                              \
    }
    if (sh.block == 512) PACH_MIX_O(512)
    else PACH_MIX_O(256)
#undef PACH_MIX_O
#undef PACH_MIX
    return hipGetLastError();
}

template <int BS, int OCC> static int mix_blocks_per_cu_bs(int lds_bytes)
{
    int nb = 0;
    hipError_t e =
        lds_bytes > 0
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<double, true, BS, OCC, false, false, true>, BS,
                                                           lds_bytes)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<double, false, BS, OCC, false, false, true>, BS,
                                                           0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

hipError_t launch_exact(hipStream_t st, int grid, const TraceConst &tc, const DevScene<double> &s,
                        const DevRecv<double> &rv, const ORay<double> *in, const unsigned *n_in, unsigned cap_in,
                        ORay<double> *out, unsigned *out_count, unsigned out_cap, DevCounters *cnt)
{
    hipLaunchKernelGGL(k_exact, dim3(grid), dim3(kBlock), 0, st, tc, s, rv, in, n_in, cap_in, out, out_count, out_cap,
                       cnt);
    return hipGetLastError();
}

hipError_t launch_shoot_mix(hipStream_t st, const TraceConst &tc, const DevScene<double> &s,
                            const DevScene<float> &sg, const double *o, const double *d, int64_t n, int *tri,
                            double *t, unsigned *n_exact)
{
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_shoot<double, true>), dim3((unsigned)grid), dim3(kBlock), 0, st, tc, s, sg, o, d, n, tri,
                       t, n_exact);
    return hipGetLastError();
}

// Resident workgroups per CU of the mixed-walk kernel (its own register count).
int octave_mix_blocks_per_cu(const OctShape &sh, int lds_bytes)
{
    if (sh.block == 512)
        return sh.occ == 4 ? mix_blocks_per_cu_bs<512, 4>(lds_bytes)
                           : (sh.occ == 3 ? mix_blocks_per_cu_bs<512, 3>(lds_bytes) : mix_blocks_per_cu_bs<512, 0>(lds_bytes));
    return sh.occ == 4 ? mix_blocks_per_cu_bs<256, 4>(lds_bytes)
                       : (sh.occ == 3 ? mix_blocks_per_cu_bs<256, 3>(lds_bytes) : mix_blocks_per_cu_bs<256, 0>(lds_bytes));
}

#endif

template <class R, int BS, int OCC, bool SQ> static int blocks_per_cu_bs(int lds_bytes)
{
    int nb = 0;
    hipError_t e =
        lds_bytes > 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<R, true, BS, OCC, SQ, false>, BS, lds_bytes)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<R, false, BS, OCC, SQ, false>, BS, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

template <class R, int BS, bool SQ> static int blocks_per_cu_o(const OctShape &sh, int lds_bytes)
{
    if (sh.occ == 4) return blocks_per_cu_bs<R, BS, 4, SQ>(lds_bytes);
    if (sh.occ == 5) return blocks_per_cu_bs<R, BS, 5, SQ>(lds_bytes);
    if (sh.occ == 3) return blocks_per_cu_bs<R, BS, 3, SQ>(lds_bytes);
    return blocks_per_cu_bs<R, BS, 0, SQ>(lds_bytes);
}

template <class R, bool SQ> static int blocks_per_cu_q(const OctShape &sh, int lds_bytes)
{
    return sh.block == 512 ? blocks_per_cu_o<R, 512, SQ>(sh, lds_bytes) : blocks_per_cu_o<R, 256, SQ>(sh, lds_bytes);
}

template <class R> int octave_blocks_per_cu(const OctShape &sh, int lds_bytes)
{
    return sh.queue ? blocks_per_cu_q<R, true>(sh, lds_bytes) : blocks_per_cu_q<R, false>(sh, lds_bytes);
}

// Resident workgroups per CU of the mapping-receiver variant (its own register
// count, not the plain kernel's: a persistent kernel's grid must be resident).
template <class R> int octave_map_blocks_per_cu(int lds_bytes)
{
    int nb = 0;
    hipError_t e =
        lds_bytes > 0
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<R, true, 256, kMapOcc, false, true>, 256, lds_bytes)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_octave<R, false, 256, kMapOcc, false, true>, 256, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

#define PACH_INSTANTIATE(R)                                                                        \
    template hipError_t launch_phase1<R>(hipStream_t, const TraceConst &, const DevScene<R> &,     \
                                         const DevRecv<R> &, int64_t, int, const P1State<R> *, int, \
                                         ORay<R> *, unsigned *, unsigned, P1State<R> *, unsigned *, \
                                         DevCounters *);                                           \
    template hipError_t launch_octave<R>(hipStream_t, int, const OctShape &, const TraceConst &,   \
                                         const DevScene<R> &, const DevScene<float> &,              \
                                         const DevRecv<R> &, const ORay<R> *,                       \
                                         const unsigned *, unsigned, const ORay<R> *,               \
                                         const unsigned *, unsigned, unsigned *, ORay<R> *, int,    \
                                         ShadeItem<R> *, ORay<R> *, unsigned *, unsigned, int,      \
                                         DevCounters *);                                            \
    template int octave_blocks_per_cu<R>(const OctShape &, int);                                   \
    template hipError_t launch_shoot<R>(hipStream_t, const TraceConst &, const DevScene<R> &,      \
                                        const DevScene<float> &, const double *, const double *,   \
                                        int64_t, int *, double *, unsigned *);                     \
    template int octave_map_blocks_per_cu<R>(int);                                              \
    template hipError_t launch_live_sources<R>(hipStream_t, const ORay<R> *, const unsigned *,     \
                                               unsigned, const P1State<R> *, const unsigned *,    \
                                               unsigned, uint32_t, uint32_t, uint32_t *, unsigned *);

}  // namespace pach
