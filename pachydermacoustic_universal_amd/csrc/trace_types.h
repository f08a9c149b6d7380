// trace_types.h -- plain data shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pach {

constexpr int kBands = 8;
constexpr int kPlanes = 11;
constexpr int kBlock = 256;  // 4 wave64 per workgroup
constexpr int kMaxLdsDist = 32768;  // voxel-distance bytes staged in LDS per workgroup

// Everything a flattened walk reads on entering a voxel, in one 64-byte
// record (one cache line per voxel entry instead of one per array): the list
// range, per entry face the mask of list entries the neighbour across that
// face also lists (face order -x, +x, -y, +y, -z, +z by the face the walk
// came in through: grid.h face_skip), the receivers whose spheres touch the
// voxel, the edge of the largest empty cube cornered at the voxel per ray
// octant (grid.h octant_cubes) and the L-inf distance to the nearest occupied
// voxel (receiver voxels count as occupied when receiver masks are on).  An
// empty voxel has no list, skip masks or receivers: its boxes per octant
// (grid.h octant_boxes; one byte per octant) lie in first and count (first
// axis length), skip[0] and skip[1] (second axis length) and rmask (the two
// axes), words the walk loads for every voxel.
struct alignas(64) CellRec {
    uint32_t first, count;
    uint32_t skip[6];
    uint64_t rmask;
    uint8_t oct[8];
    uint32_t dist;
    uint32_t pad[3];
};
static_assert(sizeof(CellRec) == 64, "one cell record per 64 bytes");

template <class R> struct V4;
template <> struct V4<float> { using T = float4; };
template <> struct V4<double> { using T = double4; };

// Scene in HBM: structure-of-arrays triangle records, per-(polygon, band)
// material records and a uniform voxel grid in CSR form.
// FP64: e1/e2 are the edges v1 - v0, v2 - v0 (Moller-Trumbore), nrm the unit
// normal.  FP32: e1/e2 hold the dual-basis rows r0, r1 and nrm the plane row
// r2 (unit normal, w = -n . v0) of the triangle test in trace_impl.h.
template <class R> struct DevScene {
    const typename V4<R>::T *v0;   // [T] vertex 0 (w unused)
    const typename V4<R>::T *e1;   // [T] v1 - v0 (FP64) or dual row r0 (FP32)
    const typename V4<R>::T *e2;   // [T] v2 - v0 (FP64) or dual row r1 (FP32)
    const typename V4<R>::T *nrm;  // [T] unit normal (Topology.Normal), FP32 w = plane offset
    const int *surf;               // [T] polygon index (Surf_ID)
    const uint2 *cell_range;       // [ncell] (first list index, triangle count)
    const uint32_t *cell_tris;     // triangle ids per cell
    const float4 *cell_rows;       // FP32 rows (r2, r0, r1) of each list entry, in list order
    const uint32_t *cell_skip;     // [ncell][6] entries also listed across each face (grid.h face_skip)
    const uint8_t *cell_dist;      // [ncell] L-inf distance (cells) to the nearest occupied cell, 0 = occupied
    const uint8_t *cell_oct;       // [ncell][8] edge of the empty cube cornered at the cell per ray octant (grid.h)
    const uint64_t *cell_rmask;    // [ncell] receivers touching the voxel (null: test every receiver)
    const CellRec *cell_rec;       // [ncell] the flattened walks' view of the same data
    int gx, gy, gz;
    int ncell;
    int dist_lds;                  // 1: k_octave stages cell_dist in LDS
    R gmin[3], cs[3], inv_cs[3];
    const typename V4<R>::T *mat;  // [S*8] (1 - alpha, s, lo, hi) of Lambert_Scattering
    const R *trans;                // [S*8] transmission
};

// Per-band constants read with a runtime band index (kept in device memory so
// the kernels index them with plain loads instead of copying an argument array).
struct BandTab {
    double atten[kBands];  // air attenuation, dB/m
    float kexp2[kBands];   // -0.1 * atten * log2(10): 10^(-0.1 a x) = exp2(k x)
    double kexp2d[kBands]; // the same in FP64
};

// One room of a trace (a single-scene trace is one room).  Source rays
// [ray_base, ray_base + ray_count) leave this room's source; its receivers are
// [rec_base, rec_base + rec_count) of the receiver list.
struct RoomTab {
    double src[3];
    double T1;             // broadband-phase threshold: power[h] 1e-2 (TaskAll:351)
    double power[kBands];  // source power per band, W
    double T2[kBands];     // split-phase energy threshold (TaskAll:354)
    double T3[kBands];     // End_Trace energy threshold (TaskAll:355)
    int64_t ray_base, ray_count;
    int h;                 // steering band (TaskAll:88-109)
    int rec_base, rec_count;
    int pad;
};

// Mapping receivers (PachMapReceiver, Pach_Map_Receiver.cs): spheres of
// radius r2^(1/2) culled through their own voxel grid (cells x-major:
// (x * ct[1] + y) * ct[2] + z, receiver lists sorted ascending).
struct DevMap {
    const double4 *ctr;      // [n] centre, w = Direct_Time (|source - centre| / c)
    const uint32_t *start;   // [ncell + 1] first list entry of each cell
    const uint32_t *list;    // receiver indices
    const float4 *lctr;      // [list entries] FP32 centre of each entry's receiver, w = its index (bits)
    float scr;               // absolute widening of the FP32 screen for the rounding of lctr
    float lo[3], hi[3];      // the receiver box widened by one radius (segment-box screen)
    double bmin[3], bmax[3], vd[3];
    int ct[3];
    int n;
    double r2, sizemod, c, rho_c;
    int directional, sct;
    int time1pt;             // Histogram_1Pt: every deposit to bin 0 of the energy plane (sct 1)
    double *hist;            // 8 planes (E, P, +x, +y, +z, -x, -y, -z) of [n][8][sct], 1000 Hz
    uint64_t plane_stride;
};

template <class R> struct DevRecv {
    const typename V4<R>::T *ctr;  // [R] centre (w unused)
    int n;
    int masked;                    // 1: test only the receivers of the walk's voxel masks
    R r2;
    double sizemod;   // 1 / (pi r^3)
    double c, inv_c;  // speed of sound and 1/c as Spherical_Receiver holds them
    const BandTab *band;  // per-band constants in HBM (indexed by a runtime band)
    double rho_c;
    int fs, sct;
    double *hist;         // [kPlanes][n][8][sct]
    uint64_t plane_stride;
    // Exact accumulation (nullptr: FP64 atomics into hist): every deposit is
    // scaled by 2^s (fx_se for the energy-class planes, fx_sp for pressure),
    // rounded to an integer of up to 124 bits and added as four signed 32-bit
    // limbs into int64 sums [4][kPlanes][n][8][sct]; k_fx_flush turns the
    // sums into hist.  Integer sums do not depend on the order of the adds.
    long long *fx;
    double fx_se, fx_sp;
    int has_map;          // 1: the receivers are the map (RecMain = PachMapReceiver)
    DevMap map;
};

struct TraceConst {
    int mode, oct_lo, oct_hi, is_order, dir_mode;
    int p1_max, max_bounces;
    int walk_steps;  // voxels a lane's walk advances per loop step of k_octave
    int shade_min;   // lanes that must be waiting before k_octave shades
    int walk_flat;   // 1: flattened walk (walk_steps counts tests + voxels)
    int refill_min;  // idle lanes a wave gathers before it refills them
    double cotime;
    const BandTab *band;
    const RoomTab *rooms;  // [n_rooms], ray_base ascending
    int n_rooms;
    uint32_t key0, key1;
    const double *dirs;   // supplied directions [n_dirs][3]
    int64_t n_dirs;
    const double *geo;    // 80 geodesic faces [80][9]
    int has_trans;        // any transmission coefficient > 0
    // Mixed walk (FP32 screen, FP64 decisions): bounds on the rounding error of
    // the FP32 screen, from the scene's extent and triangle shapes.
    float mx_kt, mx_kt0;  // |t32 - t| <= kt |1 / (n . d)| + kt0 (m)
    float mx_ku, mx_ku0;  // barycentric error <= ku |1 / (n . d)| + ku0
    float mx_kx;          // voxel-exit error <= kx |1 / d_axis| + 8 eps t (m)
    float mx_infl;        // how far past its voxel every voxel list reaches (m)
    float mx_self;        // |n . d| below which FP64 may hit the triangle left again
};

// Octave ray (OctaveRay, Classes_Rays.cs:30-181).  meta bits: 0-2 band,
// 3 currently scattered, 4 ever scattered, 5 End_Trace phase, 6 dead slot,
// 8-31 room.
template <class R> struct alignas(16) ORay {
    R o[3];
    R I;
    R d[3];
    int tri;   // last surface hit (triangle), -1 at the source
    uint32_t src;
    double t;  // t_sum in seconds
    uint32_t uid_lo, uid_hi, k, meta;
};

// Broadband ray between launches (only when the phase-1 output buffer fills).
template <class R> struct alignas(16) P1State {
    R o[3];
    R d[3];
    R E[kBands];
    double tb;
    int order, tri;
    uint32_t k, src;
    int flags;
    int room;
};

enum : uint32_t {
    M_BAND = 7u,
    M_SCAT = 1u << 3,
    M_EVER = 1u << 4,
    M_LATE3 = 1u << 5,
    M_DEAD = 1u << 6,
    M_ROOM_SHIFT = 8,
};

enum : int { F_SPLIT_PENDING = 1 };

// device counters (one set per trace)
enum : int {
    C_BROAD = 0,
    C_LATE,
    C_VLATE,
    C_OCT,
    C_SPAWN,
    C_MISS,
    C_DEP,
    C_P1CAP,
    C_GUARD,
    C_SPILL,
    C_RESUME,
    C_DROP,
    C_ITER,      // persistent-loop steps (per wave)
    C_VOXEL,     // voxels visited by walks (per lane)
    C_TEST,      // triangle tests started (per lane)
    C_SHADE,     // shading passes (per wave)
    C_SHADED,    // lanes shaded (per lane)
    C_EXACT,     // mixed walk: bounces the FP32 screen could not decide (redone by the FP64 walk)
    C_EXACT0,    // mixed walk: grazing departures (walked in FP64 from the start)
    C_FXOVF,     // exact accumulation: deposits beyond the fixed-point range (never expected)
    C_N
};

struct DevCounters {
    unsigned long long v[24];
    double lost;
    double pad;
};

// Exact accumulation (DevRecv::fx): writes hist [kPlanes][stride] from the
// limb sums fx [4][kPlanes][stride], scaled by inv_sp on plane 1 (pressure)
// and inv_se on the others (trace_f64.hip).
hipError_t launch_fx_flush(hipStream_t st, const long long *fx, double *hist, uint64_t stride, double inv_se,
                           double inv_sp);

// launchers (trace_f32.hip / trace_f64.hip)
template <class R>
hipError_t launch_phase1(hipStream_t st, const TraceConst &tc, const DevScene<R> &s,
                         const DevRecv<R> &rv, int64_t src_begin, int n_fresh,
                         const P1State<R> *resume_in, int n_resume, ORay<R> *out,
                         unsigned *out_count, unsigned out_cap, P1State<R> *resume_out,
                         unsigned *resume_count, DevCounters *cnt);

// Progress: marks in bits[] the source rays [first, first + span) that still
// have live rays between launches (the n_rays octave rays of rays, dead slots
// skipped, and the n_res suspended broadband rays of res; counts read on the
// device, at most cap and res_cap) and adds each newly marked source ray to
// *live.  The caller clears
// bits and *live first.
template <class R>
hipError_t launch_live_sources(hipStream_t st, const ORay<R> *rays, const unsigned *n_rays, unsigned cap,
                               const P1State<R> *res, const unsigned *n_res, unsigned res_cap, uint32_t first,
                               uint32_t span, uint32_t *bits, unsigned *live);

// Launch shape of k_octave: workgroup size (256 or 512) and register cap
// (occ 0: the compiler's choice; 4 / 5: at most 128 / 102 VGPRs, so 4 / 5
// waves share a SIMD; -1: 4 for FP32, 0 for FP64, whose 4-wave build spills
// ~370 bytes per lane).
struct OctShape {
    int block = 256;
    int occ = -1;
    int queue = 0;  // 1: finished walks go through a per-wave shade queue
};

// A finished walk waiting in a wave's shade queue (k_octave with a queue):
// the ray, the triangle hit (-1: miss), the hit distance and the receiver
// candidates the walk collected.
template <class R> struct alignas(16) ShadeItem {
    ORay<R> r;
    int best;
    R tbest;
    uint64_t rmask;
};
constexpr int kShadeQ = 128;  // per-wave queue capacity (at most 63 wait + 64 pushed)

template <class R>
hipError_t launch_octave(hipStream_t st, int grid, const OctShape &sh, const TraceConst &tc, const DevScene<R> &s,
                         const DevScene<float> &sg, const DevRecv<R> &rv, const ORay<R> *in0, const unsigned *n0,
                         unsigned cap0,
                         const ORay<R> *in1, const unsigned *n1, unsigned cap1, unsigned *in_head,
                         ORay<R> *stacks, int stack_cap, ShadeItem<R> *shq, ORay<R> *dump, unsigned *dump_count,
                         unsigned dump_cap, int min_iters, DevCounters *cnt);

template <class R> int octave_blocks_per_cu(const OctShape &sh, int lds_bytes);

// mixed walk (trace_mix.hip): FP64 rays and decisions, FP32 screen over sg
hipError_t launch_octave_mix(hipStream_t st, int grid, const OctShape &sh, const TraceConst &tc,
                             const DevScene<double> &s, const DevScene<float> &sg, const DevRecv<double> &rv,
                             const ORay<double> *in0, const unsigned *n0, unsigned cap0, const ORay<double> *in1,
                             const unsigned *n1, unsigned cap1, unsigned *in_head, ORay<double> *stacks,
                             int stack_cap, ORay<double> *dump, unsigned *dump_count, unsigned dump_cap,
                             int min_iters, DevCounters *cnt, ORay<double> *xq, unsigned *xq_count, unsigned xq_cap);
int octave_mix_blocks_per_cu(const OctShape &sh, int lds_bytes);
// the mixed kernel's undecided bounces, by the FP64 walk (k_exact)
hipError_t launch_exact(hipStream_t st, int grid, const TraceConst &tc, const DevScene<double> &s,
                        const DevRecv<double> &rv, const ORay<double> *in, const unsigned *n_in, unsigned cap_in,
                        ORay<double> *out, unsigned *out_count, unsigned out_cap, DevCounters *cnt);

// closest-hit queries (k_shoot): o, d [n][3] in device memory
template <class R>
hipError_t launch_shoot(hipStream_t st, const TraceConst &tc, const DevScene<R> &s, const DevScene<float> &sg,
                        const double *o, const double *d, int64_t n, int *tri, double *t, unsigned *n_exact);
hipError_t launch_shoot_mix(hipStream_t st, const TraceConst &tc, const DevScene<double> &s,
                            const DevScene<float> &sg, const double *o, const double *d, int64_t n, int *tri,
                            double *t, unsigned *n_exact);
template <class R> int octave_map_blocks_per_cu(int lds_bytes);

}  // namespace pach
