// runtime.cpp -- host side of libpach.so: the C ABI of include/pach.h.
//
// Owns the HIP stream, the scene / receiver buffers in HBM, and the trace
// orchestration: the source rays are traced in batches; each batch's
// broadband phase (k_phase1) feeds the persistent octave-ray kernel
// (k_octave) together with the rays the previous launch handed over, until
// every ray of every batch has terminated.
#include "../../include/pach.h"
#include "filters.h"
#include "grid.h"
#include "image_source.h"
#include "is_plan.h"
#include "trace_types.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace pach;

namespace {

template <class R> struct SceneBuf {
    typename V4<R>::T *v0 = nullptr, *e1 = nullptr, *e2 = nullptr, *nrm = nullptr, *mat = nullptr;
    R *trans = nullptr;
};

template <class R> struct WorkBuf {
    ORay<R> *p1_out = nullptr;
    unsigned p1_cap = 0;
    P1State<R> *res[2] = {nullptr, nullptr};
    unsigned res_cap = 0;
    ORay<R> *stacks = nullptr;
    int stack_cap = 0;
    size_t n_waves = 0;
    ShadeItem<R> *shq = nullptr;  // per-wave shade queues (shade_queue on)
    size_t shq_waves = 0;
    ORay<R> *dump[2] = {nullptr, nullptr};
    unsigned dump_cap = 0;
    ORay<R> *xq = nullptr;  // mixed precision: bounces for the FP64 walk (k_exact)
    unsigned xq_cap = 0;
};

}  // namespace

struct pach_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int n_cu = 0;
    std::string err;

    // scene
    bool has_scene = false;
    int T = 0, S = 0;
    std::vector<double> area;                 // [S]
    std::vector<double> alpha, scat, trans;   // [S*8]
    bool any_trans = false;
    int g[3] = {0, 0, 0};
    double gmin[3] = {0, 0, 0}, cs[3] = {1, 1, 1};
    int *d_surf = nullptr;
    uint32_t *d_cell_range = nullptr, *d_cell_tris = nullptr;
    float4 *d_cell_rows = nullptr;  // FP32 test rows (r2, r0, r1) of every voxel-list entry
    uint32_t *d_cell_skip = nullptr;  // [ncell][6] list entries also in each face neighbour
    std::vector<float> tri_rows32;  // [T][12] FP32 rows r2, r0, r1 per triangle
    double row_norm = 0;            // largest |r0|, |r1| (barycentric gradient, 1 / shortest altitude)
    uint8_t *d_cell_dist = nullptr;
    uint8_t *d_cell_oct = nullptr;  // [ncell][8] empty-cube edges per ray octant
    uint64_t *d_cell_rmask = nullptr;
    CellRec *d_cell_rec = nullptr;     // [ncell] packed per-voxel records of the flattened walks
    std::vector<uint32_t> cell_skip;   // host copy of the face skip masks (records are rebuilt with receivers)
    int ncell = 0;
    pach::Grid grid;        // host copy (receiver voxels are marked at trace time)
    std::vector<double> verts;  // host copy of the triangles (grid rebuilds)
    std::vector<int> surf;      // host copy of surf_of_tri
    bool rmask_dirty = true;
    bool rmask_on = false;
    std::vector<double> centers;
    SceneBuf<float> s32;
    SceneBuf<double> s64;

    // source
    bool has_src = false;
    double origin[3] = {0, 0, 0}, power[8] = {0};
    int dir_mode = PACH_DIR_GEODESIC;
    double *d_dirs = nullptr;
    int64_t n_dirs = 0;
    double *d_geo = nullptr;

    // receivers
    bool has_recv = false;
    int R = 0;
    double radius = 1, atten[8] = {0}, c = 343, rho_c = 411, fs = 1000;
    int sct = 0;
    float4 *d_ctr32 = nullptr;
    double4 *d_ctr64 = nullptr;
    double *d_hist = nullptr;
    size_t hist_n = 0;  // per plane
    // exact accumulation (pach_exact_histograms): int64 limb sums [4][kPlanes][hist_n]
    bool hist_exact = false;
    long long *d_fx = nullptr;
    size_t fx_cap = 0;  // values per limb allocated

    // mapping receivers (pach_map_receivers_set)
    bool map_on = false;
    int map_n = 0, map_sct = 0, map_directional = 1, map_time1pt = 0;
    double map_r = 0;
    std::vector<double> map_ctr;  // [M][3]
    double map_bmin[3] = {0, 0, 0}, map_bmax[3] = {0, 0, 0}, map_vd[3] = {1, 1, 1};
    int map_ct[3] = {1, 1, 1};
    double4 *d_map_ctr = nullptr;
    uint32_t *d_map_start = nullptr, *d_map_list = nullptr;
    float4 *d_map_lctr = nullptr;  // per list entry: FP32 centre, w = receiver index (bits)
    float map_scr = 0;             // FP32 screen widening for the rounding of d_map_lctr
    double *d_map_hist = nullptr;
    size_t map_hist_n = 0;  // per plane

    // rooms (pach_rooms_set); empty: one room made of the whole scene
    struct RoomIn {
        double src[3], power[8];
        int64_t rays;
        int surf_count, rec_count;
    };
    std::vector<RoomIn> rooms;
    std::vector<RoomTab> room_tab;  // last trace's room table (ray ranges for scaling)
    RoomTab *d_rooms = nullptr;
    int d_rooms_cap = 0;

    // counters and small device scalars
    DevCounters *d_cnt = nullptr;
    unsigned *d_scal = nullptr;  // [0] p1_count [1] res_count [2] in_head [3..4] dump_count[2]
                                 // [5] exact-queue count [6] source rays with live rays
    uint32_t *d_live = nullptr;  // one bit per source ray of the range being traced
    size_t live_words = 0;
    BandTab *d_band = nullptr;

    WorkBuf<float> w32;
    WorkBuf<double> w64;

    // tuning (pach_tune)
    int tune_dist_lds = 1;     // voxel distances staged in LDS when they fit
    int tune_rmask = 1;        // per-voxel receiver masks (R <= 64)
    int tune_walk_steps = 12;  // walk units (tests + voxel steps) per loop step
    int tune_shade_min = 16;   // lanes waiting before shading
    int tune_walk_flat = 2;    // flattened walk, one kind of unit per wave iteration
    int tune_refill_min = 1;   // idle lanes gathered before a wave refills
    int tune_dump_min = 256;   // loop steps before a drained wave hands over its work
    int64_t tune_xq = (int64_t)32 << 20;  // mixed precision: exact-queue rays beyond the per-wave reserve
    OctShape tune_shape;       // k_octave workgroup size and register cap

    // trace state
    std::thread worker;
    std::atomic<bool> running{false};
    std::atomic<bool> cancel{false};
    std::atomic<int64_t> rays_done{0};
    std::atomic<int64_t> rays_total{0};
    std::atomic<int64_t> oct_rays{0};
    int status = 0;
    pach_stats stats{};
    pach_is_stats is_stats{};
    pach_filter_stats filter_stats{};
    // auralisation-filter buffers, kept between calls (grown on demand)
    double *f_env = nullptr, *f_car = nullptr, *f_part = nullptr, *f_out = nullptr, *f_tab = nullptr;
    size_t f_env_n = 0, f_part_n = 0, f_out_n = 0, f_tab_n = 0;
    int f_car_fs = 0;
    unsigned char *noalloc_scratch = nullptr;  // this device's allocation-free filter scratch
    // closest-hit query buffers (pach_shoot), grown on demand
    double *q_o = nullptr, *q_d = nullptr, *q_t = nullptr;
    int *q_tri = nullptr;
    unsigned *q_exact = nullptr;
    size_t q_o_n = 0, q_d_n = 0, q_t_n = 0, q_tri_n = 0;
    int64_t q_exact_last = 0;
    std::mutex mu;
};

// ------------------------------------------------------------------ helpers
#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return PACH_EHIP;                                                          \
        }                                                                              \
    } while (0)

static int fail(pach_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <class T> static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <class T> static hipError_t dalloc(T *&p, size_t n)
{
    dfree(p);
    if (n == 0) n = 1;
    return hipMalloc((void **)&p, n * sizeof(T));
}

// Device buffer of at least n elements, kept (cap tracks its size).
template <class T> static hipError_t dgrow(T *&p, size_t &cap, size_t n)
{
    if (p && cap >= n) return hipSuccess;
    cap = 0;
    const hipError_t e = dalloc(p, n);
    if (e == hipSuccess) cap = n;
    return e;
}

template <class R> static void free_scene(SceneBuf<R> &s)
{
    dfree(s.v0); dfree(s.e1); dfree(s.e2); dfree(s.nrm); dfree(s.mat); dfree(s.trans);
}

template <class R> static void free_work(WorkBuf<R> &w)
{
    dfree(w.p1_out); dfree(w.res[0]); dfree(w.res[1]); dfree(w.stacks); dfree(w.dump[0]); dfree(w.dump[1]);
    dfree(w.shq);
    dfree(w.xq);
    w.xq_cap = 0;
    w.p1_cap = w.res_cap = w.dump_cap = 0;
    w.stack_cap = 0;
    w.n_waves = 0;
    w.shq_waves = 0;
}

// ------------------------------------------------------------------ medium
extern "C" int pach_uniform_medium(int air_choice, double temp_c, double hr, double pa_hpa,
                                   int edge_correction, double atten[8], double *c, double *rho)
{
    // Calculate_Attenuation (Classes_Medium_Props.cs:52-190) with the pressure
    // Uniform_Medium passes (Pa/10, :289) and C# integer-division exponents
    // (-1/2 == 0, -1/3 == 0, -5/2 == -2, 1/2 == 0).
    if (!atten) return PACH_EINVAL;
    static const double F[8] = {62.5, 125, 250, 500, 1000, 2000, 4000, 8000};
    const double Tk = temp_c + 273.15, Pa = pa_hpa / 10.0;
    for (int i = 0; i < 8; ++i) atten[i] = 0;
    if (air_choice == 0) {
        const double h = hr * std::pow(10.0, -6.8346 * std::pow(273.16 / Tk, 1.261) + 4.6151) / (Pa / 101.325);
        for (int o = 0; o < 8; ++o) {
            const double frO = (Pa / 101.325) * (24 + 4.04e4 * h * ((0.02 + h) / (0.391 + h)));
            const double frN = (Pa / 101.325) * (9 + 280 * h);
            const double z = 0.1068 * std::exp(-3352 / Tk) / (frN + F[o] * F[o] / frN);
            const double y = std::pow(Tk / 293.15, -2.0) *
                             (0.01275 * std::exp(-2239.1 / Tk) / (frO + F[o] * F[o] / frO) + z);
            atten[o] = 8.686 * F[o] * F[o] * ((1.84e-11 / (Pa / 101.325)) + y);
        }
    } else if (air_choice == 1) {
        for (int q = 0; q < 8; ++q) {
            double a = ((85 / hr) * ((F[q] / 1000) * (F[q] / 1000) * 0.0001)) / 8.7;
            a -= 0.04 * (Tk - 273.15 - 20) * a;
            atten[q] = a * 0.50248756218905477;
        }
    } else if (air_choice == 2) {
        static const double A[6][8] = {
            {0.0001, 0.0001, 0.0003, 0.0006, 0.001, 0.0019, 0.0058, 0.0203},
            {0.0001, 0.0001, 0.0003, 0.0006, 0.001, 0.0017, 0.0041, 0.0135},
            {0.0001, 0.0001, 0.0003, 0.0006, 0.0011, 0.0017, 0.0035, 0.0106},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0011, 0.0027, 0.0094, 0.029},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0008, 0.0018, 0.0059, 0.0211},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0007, 0.0014, 0.0044, 0.0158}};
        const int base = (Tk - 273.15 > 15) ? 0 : 3;
        int row = -1;
        if (hr < 50) row = base;
        else if (hr < 70) row = base + 1;
        else if (hr > 70) row = base + 2;
        if (row >= 0)
            for (int q = 0; q < 8; ++q) atten[q] = A[row][q];
        if (edge_correction)
            for (int q = 0; q < 8; ++q) atten[q] /= std::sqrt(2.0);
    } else {
        return PACH_EINVAL;
    }
    if (c) *c = 331 + 0.6 * temp_c;  // AcousticalMath.SoundSpeed (Classes_Math.cs:272-275)
    if (rho) {                         // Calculate_Density(Tk, hPa*100, hr) (:42-49)
        const double P = pa_hpa * 100.0, T = temp_c;
        const double Pv = hr * 6.1078 * std::pow(10.0, (7.5 * T) / (237.3 + T));
        *rho = (P / (287.05 * Tk)) * (1 - 0.378 * Pv / P);
    }
    return PACH_OK;
}

// ------------------------------------------------------------------ lifetime
extern "C" int pach_abi_version(void) { return PACH_ABI_VERSION; }

extern "C" int pach_create(pach_ctx **out, int device)
{
    if (!out) return PACH_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PACH_EHIP;
    if (device < 0 || device >= n) return PACH_EINVAL;
    pach_ctx *c = new pach_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PACH_EHIP;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    // device address of the allocation-free filters' scratch array (once per device)
    if (hipError_t e = noalloc_scratch(device, &c->noalloc_scratch); e != hipSuccess) {
        pach_destroy(c);
        return PACH_EHIP;
    }
    if (c->n_cu <= 0) c->n_cu = 256;
    if (dalloc(c->d_cnt, 1) != hipSuccess || dalloc(c->d_scal, 8) != hipSuccess || dalloc(c->d_band, 1) != hipSuccess || dalloc(c->d_geo, 80 * 9) != hipSuccess) {
        pach_destroy(c);
        return PACH_ENOMEM;
    }
    // GeodesicSource.GeoSphere(1): icosahedron subdivided once (Classes_Sources.cs:290-370)
    std::vector<double> faces;
    {
        const double sqr5 = std::sqrt(5.0), phi = (1.0 + sqr5) * 0.5;
        const double ratio = std::sqrt(10.0 + 2.0 * sqr5) / (4.0 * phi);
        const double a = (.25 / ratio) * 0.5, b = (.25 / ratio) / (2.0 * phi);
        double P[12][3] = {{0, b, -a}, {b, a, 0}, {-b, a, 0}, {0, b, a}, {0, -b, a}, {-a, 0, b},
                           {0, -b, -a}, {a, 0, -b}, {a, 0, b}, {-a, 0, -b}, {b, -a, 0}, {-b, -a, 0}};
        auto nrm = [](double *v) {
            const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (l > 0) { v[0] /= l; v[1] /= l; v[2] /= l; }
        };
        for (auto &p : P) nrm(p);
        static const int F[20][3] = {{0, 1, 2}, {3, 2, 1}, {3, 4, 5}, {3, 8, 4}, {0, 6, 7},
                                     {0, 9, 6}, {4, 10, 11}, {6, 11, 10}, {2, 5, 9}, {11, 9, 5},
                                     {1, 7, 8}, {10, 8, 7}, {3, 5, 2}, {3, 1, 8}, {0, 2, 9},
                                     {0, 7, 1}, {6, 9, 11}, {6, 10, 7}, {4, 11, 5}, {4, 8, 10}};
        for (auto &f : F) {
            const double *p0 = P[f[0]], *p1 = P[f[1]], *p2 = P[f[2]];
            double m01[3], m12[3], m20[3];
            for (int k = 0; k < 3; ++k) {
                m01[k] = (p0[k] + p1[k]) / 2;
                m12[k] = (p1[k] + p2[k]) / 2;
                m20[k] = (p2[k] + p0[k]) / 2;
            }
            nrm(m01); nrm(m12); nrm(m20);
            const double *tri[4][3] = {{p0, m01, m20}, {p1, m12, m01}, {p2, m20, m12}, {m01, m12, m20}};
            for (auto &t : tri)
                for (auto *v : t) faces.insert(faces.end(), v, v + 3);
        }
    }
    if (hipMemcpy(c->d_geo, faces.data(), faces.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        pach_destroy(c);
        return PACH_EHIP;
    }
    *out = c;
    return PACH_OK;
}

extern "C" void pach_destroy(pach_ctx *c)
{
    if (!c) return;
    if (c->worker.joinable()) {
        c->cancel = true;
        c->worker.join();
    }
    (void)hipSetDevice(c->device);
    free_scene(c->s32);
    free_scene(c->s64);
    free_work(c->w32);
    free_work(c->w64);
    dfree(c->q_o); dfree(c->q_d); dfree(c->q_t); dfree(c->q_tri); dfree(c->q_exact);
    dfree(c->d_surf); dfree(c->d_cell_range); dfree(c->d_cell_tris); dfree(c->d_cell_dist); dfree(c->d_cell_rows);
    dfree(c->d_cell_skip); dfree(c->d_cell_oct);
    dfree(c->d_cell_rmask);
    dfree(c->d_cell_rec);
    dfree(c->d_dirs); dfree(c->d_geo);
    dfree(c->d_ctr32); dfree(c->d_ctr64); dfree(c->d_hist); dfree(c->d_fx);
    dfree(c->f_env); dfree(c->f_car); dfree(c->f_part); dfree(c->f_out); dfree(c->f_tab);
    dfree(c->d_map_ctr); dfree(c->d_map_start); dfree(c->d_map_list); dfree(c->d_map_lctr); dfree(c->d_map_hist);
    dfree(c->d_cnt); dfree(c->d_scal); dfree(c->d_band); dfree(c->d_rooms); dfree(c->d_live);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *pach_last_error(const pach_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ------------------------------------------------------------------ scene + grid
namespace {

template <class R, class V>
hipError_t upload_vec4(V *&dst, const std::vector<double> &src4)
{
    const size_t n = src4.size() / 4;
    std::vector<V> tmp(n);
    for (size_t i = 0; i < n; ++i) {
        tmp[i].x = (R)src4[4 * i];
        tmp[i].y = (R)src4[4 * i + 1];
        tmp[i].z = (R)src4[4 * i + 2];
        tmp[i].w = (R)src4[4 * i + 3];
    }
    hipError_t e = dalloc(dst, n);
    if (e != hipSuccess) return e;
    return hipMemcpy(dst, tmp.data(), n * sizeof(V), hipMemcpyHostToDevice);
}

template <class R> hipError_t upload_scalars(R *&dst, const std::vector<double> &src)
{
    std::vector<R> tmp(src.begin(), src.end());
    hipError_t e = dalloc(dst, tmp.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(R), hipMemcpyHostToDevice);
}

// Builds the voxel grid of the stored triangles (dims 0: automatic) and
// uploads its voxel lists and distance field.
int upload_grid(pach_ctx *c, const int dims[3])
{
    pach::Grid grid;
    std::string gerr;
    if (pach::build_grid(c->verts.data(), (int)(c->verts.size() / 9), dims, grid, gerr) != 0)
        return fail(c, PACH_EINVAL, gerr);
    // FP32 rows copied into list order: a triangle test reads its id and its
    // rows at the same list index, so the two loads do not wait on each other
    std::vector<float4> rows(std::max<size_t>(grid.tris.size(), 1) * 3);
    for (size_t j = 0; j < grid.tris.size(); ++j) {
        const float *q = c->tri_rows32.data() + 12 * (size_t)grid.tris[j];
        for (int k = 0; k < 3; ++k) rows[3 * j + k] = make_float4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]);
    }
    std::vector<uint32_t> skip;
    pach::face_skip(grid, skip);
    std::vector<uint8_t> oct;  // receiver voxels join the occupancy when receivers are set
    pach::octant_cubes(grid.dims, grid.dist, oct);
    std::vector<uint8_t> ext;
    pach::octant_boxes(grid.dims, grid.dist, oct, ext);
    std::vector<uint32_t> rec;
    pach::cell_records(grid, skip, {}, grid.dist, oct, ext, rec);
    hipError_t e;
    if ((e = dalloc(c->d_cell_rec, rec.size() / 16)) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(c, PACH_EHIP, std::string("grid upload: ") + hipGetErrorString(e));
    if ((e = dalloc(c->d_cell_oct, oct.size())) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_oct, oct.data(), oct.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dalloc(c->d_cell_skip, skip.size())) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_skip, skip.data(), skip.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dalloc(c->d_cell_rows, rows.size())) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_rows, rows.data(), rows.size() * sizeof(float4), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dalloc(c->d_cell_range, grid.ranges.size())) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_range, grid.ranges.data(), grid.ranges.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dalloc(c->d_cell_tris, std::max<size_t>(grid.tris.size(), 1))) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_tris, grid.tris.data(), grid.tris.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dalloc(c->d_cell_dist, grid.dist.size())) != hipSuccess ||
        (e = hipMemcpy(c->d_cell_dist, grid.dist.data(), grid.dist.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(c, PACH_EHIP, std::string("grid upload: ") + hipGetErrorString(e));
    for (int k = 0; k < 3; ++k) {
        c->g[k] = grid.dims[k];
        c->gmin[k] = grid.origin[k];
        c->cs[k] = grid.cell[k];
    }
    c->ncell = grid.dims[0] * grid.dims[1] * grid.dims[2];
    c->cell_skip = std::move(skip);
    c->grid = std::move(grid);
    c->rmask_dirty = true;
    return PACH_OK;
}

}  // namespace

extern "C" int pach_scene_upload(pach_ctx *c, const double *verts, int T, const int *surf, int S,
                                 const double *alpha, const double *scat, const double *trans,
                                 int gx, int gy, int gz)
{
    if (!c) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (!verts || T <= 0 || !surf || S <= 0 || !alpha || !scat)
        return fail(c, PACH_EINVAL, "scene: null array or empty scene");
    for (int t = 0; t < T; ++t)
        if (surf[t] < 0 || surf[t] >= S) return fail(c, PACH_EINVAL, "scene: surf_of_tri out of range");
    for (int i = 0; i < 8 * S; ++i) {
        if (!(alpha[i] >= 0 && alpha[i] <= 1) || !(scat[i] >= 0 && scat[i] <= 1) ||
            (trans && !(trans[i] >= 0 && trans[i] <= 1)))
            return fail(c, PACH_EINVAL, "scene: coefficients must lie in [0, 1]");
    }
    (void)hipSetDevice(c->device);
    c->has_scene = false;
    c->T = T;
    c->S = S;
    c->alpha.assign(alpha, alpha + 8 * S);
    c->scat.assign(scat, scat + 8 * S);
    c->trans.assign(8 * S, 0.0);
    c->any_trans = false;
    if (trans) {
        c->trans.assign(trans, trans + 8 * S);
        for (double t : c->trans) c->any_trans |= t > 0;
    }
    c->area.assign(S, 0.0);
    // FP64 triangles: vertex 0 and edges for Moller-Trumbore.  FP32 triangles: the
    // dual basis of (e1, e2, n): rows r0 = (e2 x n)/|n|^2, r1 = (n x e1)/|n|^2 give
    // the barycentric u, v of a plane point x as r . (x - v0), row r2 = n/|n| its
    // signed distance to the plane; each row stores -r . v0 in w.
    std::vector<double> v0(4 * T), e1(4 * T), e2(4 * T), nr(4 * T), d0(4 * T), d1(4 * T), d2(4 * T);
    for (int t = 0; t < T; ++t) {
        const double *p = verts + 9 * t;
        double a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(p[k]) || !std::isfinite(p[3 + k]) || !std::isfinite(p[6 + k]))
                return fail(c, PACH_EINVAL, "scene: non-finite vertex");
            v0[4 * t + k] = p[k];
            a[k] = e1[4 * t + k] = p[3 + k] - p[k];
            b[k] = e2[4 * t + k] = p[6 + k] - p[k];
        }
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        const double l = std::sqrt(nn);
        c->area[surf[t]] += 0.5 * l;
        for (int k = 0; k < 3; ++k) nr[4 * t + k] = l > 0 ? n[k] / l : 0.0;
        v0[4 * t + 3] = e1[4 * t + 3] = e2[4 * t + 3] = nr[4 * t + 3] = 0;
        const double bn[3] = {b[1] * n[2] - b[2] * n[1], b[2] * n[0] - b[0] * n[2], b[0] * n[1] - b[1] * n[0]};
        const double na[3] = {n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]};
        double w0 = 0, w1 = 0, w2 = 0;
        for (int k = 0; k < 3; ++k) {
            d0[4 * t + k] = nn > 0 ? bn[k] / nn : 0.0;
            d1[4 * t + k] = nn > 0 ? na[k] / nn : 0.0;
            d2[4 * t + k] = nr[4 * t + k];
            w0 -= d0[4 * t + k] * p[k];
            w1 -= d1[4 * t + k] * p[k];
            w2 -= d2[4 * t + k] * p[k];
        }
        d0[4 * t + 3] = w0;
        d1[4 * t + 3] = w1;
        d2[4 * t + 3] = w2;
    }
    c->row_norm = 0;
    for (int t = 0; t < T; ++t)
        for (const std::vector<double> *dv : {&d0, &d1}) {
            const double *q = dv->data() + 4 * t;
            c->row_norm = std::max(c->row_norm, std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
        }
    c->verts.assign(verts, verts + 9 * (size_t)T);
    c->surf.assign(surf, surf + T);
    c->tri_rows32.resize(12 * (size_t)T);
    for (size_t t = 0; t < (size_t)T; ++t)
        for (int k = 0; k < 4; ++k) {
            c->tri_rows32[12 * t + k] = (float)d2[4 * t + k];
            c->tri_rows32[12 * t + 4 + k] = (float)d0[4 * t + k];
            c->tri_rows32[12 * t + 8 + k] = (float)d1[4 * t + k];
        }
    const int dims[3] = {gx, gy, gz};
    if (int rc = upload_grid(c, dims)) return rc;
    // materials: (1 - alpha, s, lo, hi) per (polygon, band), Lambert_Scattering
    // split bounds (Classes_Materials.cs:4466-4484)
    std::vector<double> mat(4 * 8 * (size_t)S);
    for (int i = 0; i < 8 * S; ++i) {
        const double s = scat[i];
        double lo, hi;
        if (s < 0.25) { lo = 0.1; hi = 0.4; }
        else { const double m = std::fabs(1 - s) / 5; lo = s - m; hi = s + m; }
        mat[4 * i] = 1 - alpha[i];
        mat[4 * i + 1] = s;
        mat[4 * i + 2] = lo;
        mat[4 * i + 3] = hi;
    }
    std::vector<double> surfd(surf, surf + T);
    hipError_t e;
    if ((e = upload_vec4<float>(c->s32.v0, v0)) != hipSuccess || (e = upload_vec4<float>(c->s32.e1, d0)) != hipSuccess ||
        (e = upload_vec4<float>(c->s32.e2, d1)) != hipSuccess || (e = upload_vec4<float>(c->s32.nrm, d2)) != hipSuccess ||
        (e = upload_vec4<float>(c->s32.mat, mat)) != hipSuccess || (e = upload_scalars<float>(c->s32.trans, c->trans)) != hipSuccess ||
        (e = upload_vec4<double>(c->s64.v0, v0)) != hipSuccess || (e = upload_vec4<double>(c->s64.e1, e1)) != hipSuccess ||
        (e = upload_vec4<double>(c->s64.e2, e2)) != hipSuccess || (e = upload_vec4<double>(c->s64.nrm, nr)) != hipSuccess ||
        (e = upload_vec4<double>(c->s64.mat, mat)) != hipSuccess || (e = upload_scalars<double>(c->s64.trans, c->trans)) != hipSuccess)
        return fail(c, PACH_EHIP, std::string("scene upload: ") + hipGetErrorString(e));
    if ((e = dalloc(c->d_surf, T)) != hipSuccess || (e = hipMemcpy(c->d_surf, surf, T * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(c, PACH_EHIP, std::string("scene upload: ") + hipGetErrorString(e));
    c->has_scene = true;
    return PACH_OK;
}

extern "C" int pach_source_set(pach_ctx *c, const double origin[3], const double power[8], int dir_mode,
                               const double *dirs, int64_t n_dirs)
{
    if (!c) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (!origin || !power) return fail(c, PACH_EINVAL, "source: null origin or power");
    if (dir_mode < 0 || dir_mode > 2) return fail(c, PACH_EINVAL, "source: bad direction mode");
    if (dir_mode == PACH_DIR_SUPPLIED && (!dirs || n_dirs <= 0))
        return fail(c, PACH_EINVAL, "source: supplied mode needs directions");
    for (int o = 0; o < 8; ++o)
        if (!(power[o] >= 0) || !std::isfinite(power[o])) return fail(c, PACH_EINVAL, "source: bad power");
    (void)hipSetDevice(c->device);
    std::memcpy(c->origin, origin, sizeof(c->origin));
    std::memcpy(c->power, power, sizeof(c->power));
    c->dir_mode = dir_mode;
    c->n_dirs = 0;
    dfree(c->d_dirs);
    if (dir_mode == PACH_DIR_SUPPLIED) {
        HIPCHK(c, dalloc(c->d_dirs, 3 * (size_t)n_dirs));
        HIPCHK(c, hipMemcpy(c->d_dirs, dirs, 3 * (size_t)n_dirs * sizeof(double), hipMemcpyHostToDevice));
        c->n_dirs = n_dirs;
    }
    c->has_src = true;
    return PACH_OK;
}

extern "C" int pach_receivers_set(pach_ctx *c, const double *centers, int R, double radius,
                                  const double atten[8], double cs, double rho_c, int fs, int sct)
{
    if (!c) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (!centers || R <= 0 || !atten) return fail(c, PACH_EINVAL, "receivers: null array or none");
    if (!(radius > 0) || !(cs > 0) || !(rho_c > 0) || fs <= 0 || sct <= 0)
        return fail(c, PACH_EINVAL, "receivers: radius, c, rho_c, sample rate and count must be positive");
    (void)hipSetDevice(c->device);
    c->R = R;
    c->radius = radius;
    std::memcpy(c->atten, atten, sizeof(c->atten));
    c->c = cs;
    c->rho_c = rho_c;
    c->fs = fs;
    c->sct = sct;
    std::vector<float4> c32(R);
    std::vector<double4> c64(R);
    for (int r = 0; r < R; ++r) {
        c32[r] = make_float4((float)centers[3 * r], (float)centers[3 * r + 1], (float)centers[3 * r + 2], 0.f);
        c64[r] = make_double4(centers[3 * r], centers[3 * r + 1], centers[3 * r + 2], 0.0);
    }
    HIPCHK(c, dalloc(c->d_ctr32, R));
    HIPCHK(c, dalloc(c->d_ctr64, R));
    HIPCHK(c, hipMemcpy(c->d_ctr32, c32.data(), R * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ctr64, c64.data(), R * sizeof(double4), hipMemcpyHostToDevice));
    c->centers.assign(centers, centers + 3 * (size_t)R);
    c->rmask_dirty = true;
    c->hist_n = (size_t)R * 8 * sct;
    if (dalloc(c->d_hist, c->hist_n * kPlanes) != hipSuccess) return fail(c, PACH_ENOMEM, "histogram allocation");
    HIPCHK(c, hipMemset(c->d_hist, 0, c->hist_n * kPlanes * sizeof(double)));
    c->has_recv = true;
    return PACH_OK;
}

// ------------------------------------------------------------------ trace
namespace {

int h_oct(const pach_ctx *c, const double *power, int s0, int s1, int lo, int hi)
{
    // SplitRayTracer ctor (TaskAll:88-109): band with the least total
    // absorption (over the room's polygons s0..s1-1) among bands carrying
    // >= 1e-6 of the strongest band's power.
    double total[8] = {0};
    for (int s = s0; s < s1; ++s)
        for (int o = lo; o <= hi; ++o) total[o] += c->area[s] * c->alpha[8 * s + o];
    double mn = HUGE_VAL, op = 0;
    int h = 0;
    for (int o = lo; o <= hi; ++o) op = std::max(op, power[o]);
    for (int o = lo; o <= hi; ++o) {
        if (power[o] < op * 1e-6) continue;
        if (mn > total[o]) { mn = total[o]; h = o; }
    }
    return h;
}

// Room table of a trace: the user's rooms, or one room of the whole scene.
int build_rooms(pach_ctx *c, const pach_params &p, std::vector<RoomTab> &tab)
{
    std::vector<pach_ctx::RoomIn> in = c->rooms;
    if (in.empty()) {
        pach_ctx::RoomIn one{};
        std::memcpy(one.src, c->origin, sizeof(one.src));
        std::memcpy(one.power, c->power, sizeof(one.power));
        one.rays = p.ray_count;
        one.surf_count = c->S;
        one.rec_count = c->R;
        in.push_back(one);
    }
    tab.assign(in.size(), RoomTab{});
    int64_t rb = 0;
    int sb = 0, cb = 0;
    for (size_t i = 0; i < in.size(); ++i) {
        RoomTab &t = tab[i];
        std::memcpy(t.src, in[i].src, sizeof(t.src));
        std::memcpy(t.power, in[i].power, sizeof(t.power));
        t.h = h_oct(c, in[i].power, sb, sb + in[i].surf_count, p.oct_lo, p.oct_hi);
        t.T1 = in[i].power[t.h] * 1e-2;  // TaskAll:351
        for (int o = 0; o < 8; ++o) {
            const bool on = o >= p.oct_lo && o <= p.oct_hi;
            t.T2[o] = on ? in[i].power[o] * std::pow(10, -4.5) : 0.0;  // TaskAll:354
            t.T3[o] = on ? in[i].power[o] * 1e-9 : 0.0;                // TaskAll:355
        }
        t.ray_base = rb;
        t.ray_count = in[i].rays;
        t.rec_base = cb;
        t.rec_count = in[i].rec_count;
        rb += in[i].rays;
        sb += in[i].surf_count;
        cb += in[i].rec_count;
    }
    if (sb != c->S || cb != c->R) return fail(c, PACH_EINVAL, "rooms: surface or receiver counts do not match the scene");
    return PACH_OK;
}

template <class R> SceneBuf<R> &scene_of(pach_ctx *c);
template <> SceneBuf<float> &scene_of<float>(pach_ctx *c) { return c->s32; }
template <> SceneBuf<double> &scene_of<double>(pach_ctx *c) { return c->s64; }
template <class R> WorkBuf<R> &work_of(pach_ctx *c);
template <> WorkBuf<float> &work_of<float>(pach_ctx *c) { return c->w32; }
template <> WorkBuf<double> &work_of<double>(pach_ctx *c) { return c->w64; }
template <class R> const typename V4<R>::T *ctr_of(pach_ctx *c);
template <> const float4 *ctr_of<float>(pach_ctx *c) { return c->d_ctr32; }
template <> const double4 *ctr_of<double>(pach_ctx *c) { return c->d_ctr64; }

template <class R> int ensure_work(pach_ctx *c, int64_t batch, size_t waves, bool queue, bool exact_q = false)
{
    WorkBuf<R> &w = work_of<R>(c);
    const unsigned p1_cap = (unsigned)std::min<int64_t>(batch * 200 + 4096, (int64_t)1 << 31);
    const unsigned res_cap = (unsigned)batch;
    const int stack_cap = 4096;
    // hand-over queue: every wave's stack and lanes, a batch's broadband output
    // beside a backlog, and the continuations of the bounces the mixed kernel
    // leaves to the FP64 walk
    const unsigned dump_cap = (unsigned)std::min<size_t>(
        waves * (stack_cap + 64) + (size_t)p1_cap + (exact_q ? ((size_t)16 << 20) : (1u << 20)), (size_t)1 << 31);
    hipError_t e = hipSuccess;
    if (w.p1_cap < p1_cap) {
        if ((e = dalloc(w.p1_out, p1_cap)) != hipSuccess) return fail(c, PACH_ENOMEM, "phase-1 buffer");
        w.p1_cap = p1_cap;
    }
    if (w.res_cap < res_cap) {
        if ((e = dalloc(w.res[0], res_cap)) != hipSuccess || (e = dalloc(w.res[1], res_cap)) != hipSuccess)
            return fail(c, PACH_ENOMEM, "resume buffers");
        w.res_cap = res_cap;
    }
    if (w.n_waves < waves || w.stack_cap != stack_cap) {
        if ((e = dalloc(w.stacks, waves * (size_t)stack_cap)) != hipSuccess) return fail(c, PACH_ENOMEM, "wave stacks");
        w.n_waves = waves;
        w.stack_cap = stack_cap;
    }
    if (queue && w.shq_waves < waves) {
        if ((e = dalloc(w.shq, waves * (size_t)kShadeQ)) != hipSuccess) return fail(c, PACH_ENOMEM, "shade queues");
        w.shq_waves = waves;
    }
    if (w.dump_cap < dump_cap) {
        if ((e = dalloc(w.dump[0], dump_cap)) != hipSuccess || (e = dalloc(w.dump[1], dump_cap)) != hipSuccess)
            return fail(c, PACH_ENOMEM, "dump buffers");
        w.dump_cap = dump_cap;
    }
    // exact queue: a wave drains once its push lands past cap - 128 x waves
    const unsigned xq_cap = (unsigned)std::min<size_t>(waves * 128 + (size_t)c->tune_xq, (size_t)1 << 31);
    if (exact_q && w.xq_cap != xq_cap) {
        if ((e = dalloc(w.xq, xq_cap)) != hipSuccess) return fail(c, PACH_ENOMEM, "exact queue");
        w.xq_cap = xq_cap;
    }
    return PACH_OK;
}

// Receiver voxel masks and the receiver-aware distance field (rebuilt when the
// scene, the receivers or the setting change).
int refresh_receiver_cells(pach_ctx *c)
{
    if (!c->rmask_dirty) return PACH_OK;
    // mask bits are receiver indices relative to each room's first receiver
    std::vector<int> bit(c->R);
    bool fits = true;
    if (c->rooms.empty()) {
        for (int r = 0; r < c->R; ++r) bit[r] = r;
        fits = c->R <= 64;
    } else {
        int base = 0;
        for (const auto &rm : c->rooms) {
            fits = fits && rm.rec_count <= 64;
            for (int k = 0; k < rm.rec_count && base + k < c->R; ++k) bit[base + k] = k;
            base += rm.rec_count;
        }
    }
    c->rmask_on = c->tune_rmask && fits;
    std::vector<uint64_t> mask;
    std::vector<uint8_t> dist;
    if (c->rmask_on) {
        pach::receiver_cells(c->grid, c->centers.data(), c->R, bit.data(), c->radius, mask, dist);
        HIPCHK(c, dalloc(c->d_cell_rmask, mask.size()));
        HIPCHK(c, hipMemcpy(c->d_cell_rmask, mask.data(), mask.size() * 8, hipMemcpyHostToDevice));
    } else {
        dist = c->grid.dist;
    }
    HIPCHK(c, hipMemcpy(c->d_cell_dist, dist.data(), c->ncell, hipMemcpyHostToDevice));
    // the walk's moves across empty space, from the same occupancy
    std::vector<uint8_t> oct;
    pach::octant_cubes(c->g, dist, oct);
    HIPCHK(c, dalloc(c->d_cell_oct, oct.size()));
    HIPCHK(c, hipMemcpy(c->d_cell_oct, oct.data(), oct.size(), hipMemcpyHostToDevice));
    std::vector<uint8_t> ext;
    pach::octant_boxes(c->g, dist, oct, ext);
    std::vector<uint32_t> rec;
    pach::cell_records(c->grid, c->cell_skip, c->rmask_on ? mask : std::vector<uint64_t>{}, dist, oct, ext, rec);
    HIPCHK(c, dalloc(c->d_cell_rec, rec.size() / 16));
    HIPCHK(c, hipMemcpy(c->d_cell_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
    c->rmask_dirty = false;
    return PACH_OK;
}

// ------------------------------------------------------------------ convergence mode
// TaskAll:111-176 and :288-329: with rayCount < 1 the tracer runs revolutions
// of Source.Revolution_Period() = 80 source rays (Classes_Sources.cs:197-200)
// until the convergence checks of two receivers pass.

// Closest hit of the stored triangles (FP64 Moller-Trumbore, t > 1e-10):
// Check_Validity's Room.shoot (TaskAll:178-199).
bool host_shoot(const std::vector<double> &v, const double o[3], const double d[3], double &tbest)
{
    tbest = 1e300;
    bool hit = false;
    const size_t T = v.size() / 9;
    for (size_t t = 0; t < T; ++t) {
        const double *a = &v[9 * t];
        const double e1[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
        const double e2[3] = {a[6] - a[0], a[7] - a[1], a[8] - a[2]};
        const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
        const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
        if (det == 0.0) continue;
        const double inv = 1.0 / det;
        const double sv[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
        const double u = (sv[0] * pv[0] + sv[1] * pv[1] + sv[2] * pv[2]) * inv;
        if (u < 0.0 || u > 1.0) continue;
        const double q[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
        const double w = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
        if (w < 0.0 || u + w > 1.0) continue;
        const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
        if (tt > 1e-10 && tt < tbest) {
            tbest = tt;
            hit = true;
        }
    }
    return hit;
}

// AcousticalMath.Envelope (Classes_Math.cs:327-383): attack / release
// smoothing clipped to a sliding maximum over localMaxWindow seconds.
std::vector<double> envelope(const double *x, int n, double fs, double attack, double release, double window)
{
    std::vector<double> y(n);
    if (n == 0) return y;
    const int w = std::max(1, (int)std::lround(window * fs));
    const double a = 1.0 - std::exp(-1.0 / (attack * fs));
    const double r = 1.0 - std::exp(-1.0 / (release * fs));
    auto lmax = [&](int i) {
        double m = x[i];
        for (int k = std::max(0, i - w + 1); k < i; ++k) m = std::max(m, x[k]);
        return m;
    };
    double prev = x[0];
    y[0] = std::min(prev, lmax(0));
    prev = y[0];
    for (int i = 1; i < n; ++i) {
        const double coeff = x[i] > prev ? a : r;
        const double yn = std::min(prev + coeff * (x[i] - prev), lmax(i));
        y[i] = yn;
        prev = yn;
    }
    return y;
}

// AcousticalMath.SPL_Intensity (Classes_Math.cs:97-102), floored at 0 dB.
double spl_intensity(double I)
{
    const double v = 10 * std::log10(I / 1e-12);
    return v < 0 ? 0 : v;
}

int64_t conv_state_size(int kind, int n, int fs)
{
    if (kind == PACH_CONV_MINIMUM) return 4;
    const int step = std::max(1, fs / 20);
    return 2 + n / step;
}

// One revolution of Minimum_Convergence_Check.Check (TaskAll:1063-1112) or
// SpecMin_Convergence_Check.Check (TaskAll:870-956) on the running (unscaled)
// energy histogram of one receiver in the steering band.
bool conv_step(int kind, const double *E, const double *spec, const double *diff, const double *rogue, int n, int fs,
               double direct_s, double *st)
{
    if (kind == PACH_CONV_MINIMUM) {
        // state: snapshot50, snapshot80, snapshotinf, count
        const int s0 = (int)std::floor(direct_s * fs);
        const int s50 = (int)std::floor(50.0 * fs / 1000) + s0;
        const int s80 = (int)std::floor(80.0 * fs / 1000) + s0;
        double sn50 = 0, sn80 = 0, sninf = 0;
        for (int i = s0; i < s50 && i < n; ++i) sn50 += E[i];
        sn80 += sn50;
        for (int i = s50; i < s80 && i < n; ++i) sn80 += E[i];
        for (int i = s80; i < n; ++i) sninf += E[i];
        const double conv1 = std::fabs(sn50 - st[0]) / st[0];
        const double conv2 = std::fabs(sn80 - st[1]) / st[1];
        const double convinf = std::fabs(sninf - st[2]) / st[2];
        if (sn50 == 0 || conv1 > 0.02 || sn80 == 0 || conv2 > 0.02 || sninf == 0 || convinf > 0.1) st[3] = 0;
        else st[3] += 1;
        st[0] = sn50;
        st[1] = sn80;
        st[2] = sninf;
        return st[3] > 5;
    }
    // state: ray count, count, snapshot[binct]
    const int step = std::max(1, fs / 20);
    const int binct = n / step;
    double *snap = st + 2;
    st[0] += 80;
    std::vector<double> sn(binct, 0.0), sn_spec(binct, 0.0), sn_diff(binct, 0.0);
    for (int i = 0; i < binct; ++i)
        for (int j = 0; j < step; ++j) {
            sn[i] += E[i * step + j];
            sn_spec[i] += spec[i * step + j];
            sn_diff[i] += diff[i * step + j];
        }
    std::vector<double> es = envelope(spec, n, fs, .001, 0.01, 0.02);
    std::vector<double> ed = envelope(diff, n, fs, .001, 0.01, 0.02);
    std::vector<double> er = envelope(rogue, n, fs, .001, 0.01, 0.02);
    std::vector<double> si_spec(n), si_comb(n);
    for (int i = 0; i < n; ++i) {
        si_comb[i] = spl_intensity(ed[i] + er[i]) - 110;
        si_spec[i] = spl_intensity(es[i]) - 110;
    }
    const double inf60diff = binct > 0 ? sn_diff[0] * 1e-6 : 0.0;
    int stop = binct - 1;
    while (stop > 1 && sn_spec[stop] < inf60diff) --stop;
    double conv = 0;
    for (int i = 0; i < binct; ++i) {
        if (sn[i] == 0 || snap[i] == 0) continue;
        const double d = std::fabs(sn[i] - snap[i]) / snap[i];
        if (i < stop) conv = std::max(conv, d);
    }
    int spec_cross = 0;
    for (int i = n - 1; i > 0; --i) {
        spec_cross = i;
        if (si_spec[i] > si_comb[i] + 4) break;
    }
    const double nm = 343.0 * spec_cross / n;
    const double nmin = std::max(5000.0, std::ceil(4 * nm * nm));
    for (int i = 0; i < binct; ++i) snap[i] = sn[i];
    if (conv == 0) return false;
    if (conv > 0.1) st[1] = 0;
    st[1] += 1;
    return st[1] > 10 && st[0] > nmin;
}

// Check_Validity over every receiver (TaskAll:132-162): the farthest receiver
// the source sees and the farthest one it does not.
void conv_receivers(const pach_ctx *c, int &vis, int &occ)
{
    vis = occ = -1;
    double max_v = 0, max_o = 0;
    for (int i = 0; i < c->R; ++i) {
        const double *C = &c->centers[3 * i];
        double d[3] = {C[0] - c->origin[0], C[1] - c->origin[1], C[2] - c->origin[2]};
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (double &x : d) x /= dist;
        double t;
        const bool visible = host_shoot(c->verts, c->origin, d, t) ? t >= dist : true;
        if (visible) {
            if (max_v < dist) { vis = i; max_v = dist; }
        } else if (max_o < dist) {
            occ = i;
            max_o = dist;
        }
    }
}

// Device view of the uploaded scene (no receiver masks, distances read from HBM).
template <class R> DevScene<R> dev_scene(pach_ctx *c)
{
    SceneBuf<R> &sb = scene_of<R>(c);
    DevScene<R> s{};
    s.v0 = sb.v0; s.e1 = sb.e1; s.e2 = sb.e2; s.nrm = sb.nrm; s.surf = c->d_surf;
    s.cell_range = (const uint2 *)c->d_cell_range; s.cell_tris = c->d_cell_tris;
    s.cell_rows = c->d_cell_rows;
    s.cell_skip = c->d_cell_skip;
    s.cell_oct = c->d_cell_oct;
    s.cell_dist = c->d_cell_dist;
    s.cell_rmask = nullptr;
    s.cell_rec = c->d_cell_rec;
    s.gx = c->g[0]; s.gy = c->g[1]; s.gz = c->g[2];
    s.ncell = c->ncell;
    s.dist_lds = 0;
    for (int k = 0; k < 3; ++k) {
        s.gmin[k] = (R)c->gmin[k];
        s.cs[k] = (R)c->cs[k];
        s.inv_cs[k] = (R)(1.0 / c->cs[k]);
    }
    s.mat = sb.mat;
    s.trans = sb.trans;
    return s;
}

// Error bounds of the mixed walk's FP32 screen (trace_impl.h, tri_filter and
// mix_limits) for the current scene and grid.  eps = 2^-24; L: the largest
// distance of a grid point from the coordinate origin, D: the grid diagonal,
// G: the largest barycentric row norm.  The screen's plane distance is off by
// at most 12 eps L + 8 eps |t| over |n . d| plus 3 eps |t| (rounded inputs,
// three fused products, the reciprocal), its hit point by that plus
// eps (2 L + D), its barycentric coordinates by G times that plus 11 eps L G;
// each bound is doubled.  A voxel exit, c s + q with s = cs / d and
// q = (cs [d > 0] + gmin - o) / d (the walk's axis lines), is off by at most
// (7 D + 2 L) eps / |d_axis| + eps t <= 16 eps L / |d_axis| + eps t; the bound
// used is 32 eps L / |d_axis| + 8 eps t.  FP64 can hit the triangle a ray leaves again only when the
// departure point's FP64 distance from its plane, ~2.2e-16 (L + D), exceeds
// 1e-10 |n . d|.
void mix_bounds(const pach_ctx *c, TraceConst &tc)
{
    const double eps = std::ldexp(1.0, -24);
    double L = 0, D2 = 0;
    for (int k = 0; k < 3; ++k) {
        const double lo = c->gmin[k], hi = c->gmin[k] + c->g[k] * c->cs[k];
        const double m = std::max(std::fabs(lo), std::fabs(hi));
        L += m * m;
        D2 += (hi - lo) * (hi - lo);
    }
    L = std::sqrt(L);
    const double D = std::sqrt(D2), G = c->row_norm;
    const double kt = 2 * eps * (12 * L + 8 * D), kt0 = 2 * 3 * eps * D;
    tc.mx_kt = (float)kt;
    tc.mx_kt0 = (float)kt0;
    tc.mx_ku = (float)(2 * G * kt);
    tc.mx_ku0 = (float)(2 * G * (kt0 + eps * (2 * L + D) + 11 * eps * L));
    tc.mx_kx = (float)(32 * eps * L);
    // grid.cpp lists a triangle in every voxel its box, grown by 1e-3 of a voxel per side, touches
    tc.mx_infl = (float)(0.9e-3 * std::min({c->cs[0], c->cs[1], c->cs[2]}));
    tc.mx_self = (float)std::max(1e-4, 16 * 1.1102230246251565e-16 * (L + D) / 1e-10);
}

template <class R> int run_trace(pach_ctx *c, const pach_params p)
{
    (void)hipSetDevice(c->device);
    if (int rc = refresh_receiver_cells(c)) return rc;
    auto t0 = std::chrono::steady_clock::now();
    TraceConst tc{};
    tc.mode = p.mode;
    tc.oct_lo = p.oct_lo;
    tc.oct_hi = p.oct_hi;
    tc.is_order = p.is_order;
    tc.dir_mode = c->dir_mode;
    tc.p1_max = p.p1_max_bounces > 0 ? p.p1_max_bounces : 128;
    tc.max_bounces = p.max_bounces > 0 ? p.max_bounces : 100000;
    tc.walk_steps = c->tune_walk_steps;
    tc.shade_min = c->tune_shade_min;
    tc.walk_flat = c->tune_walk_flat;
    tc.refill_min = c->tune_refill_min;
    tc.cotime = p.cutoff_m / c->c;  // TaskAll:83
    BandTab bt{};
    for (int o = 0; o < 8; ++o) {
        bt.atten[o] = c->atten[o];
        bt.kexp2[o] = (float)(-0.1 * c->atten[o] * std::log2(10.0));
        bt.kexp2d[o] = -0.1 * c->atten[o] * std::log2(10.0);
    }
    tc.band = c->d_band;
    // convergence mode (rayCount < 1): revolutions of 80 source rays until the checks pass
    const bool conv_mode = p.ray_count <= 0 && c->rooms.empty();
    pach_params pr = p;
    if (conv_mode) pr.ray_count = (int64_t)1 << 40;
    std::vector<RoomTab> rooms;
    if (int rc = build_rooms(c, pr, rooms)) return rc;
    if ((int)rooms.size() > c->d_rooms_cap) {
        HIPCHK(c, dalloc(c->d_rooms, rooms.size()));
        c->d_rooms_cap = (int)rooms.size();
    }
    HIPCHK(c, hipMemcpy(c->d_rooms, rooms.data(), rooms.size() * sizeof(RoomTab), hipMemcpyHostToDevice));
    tc.rooms = c->d_rooms;
    tc.n_rooms = (int)rooms.size();
    c->room_tab = rooms;
    tc.key0 = (uint32_t)p.seed;
    tc.key1 = (uint32_t)(p.seed >> 32);
    tc.dirs = c->d_dirs;
    tc.n_dirs = c->n_dirs;
    tc.geo = c->d_geo;
    tc.has_trans = c->any_trans ? 1 : 0;

    DevScene<R> s = dev_scene<R>(c);
    s.cell_rmask = c->rmask_on ? c->d_cell_rmask : nullptr;
    const int lds_bytes = (c->ncell + 15) & ~15;
    s.dist_lds = (lds_bytes <= kMaxLdsDist && c->tune_dist_lds) ? 1 : 0;
    // mixed precision: FP64 rays and decisions, the FP32 grid and rows screen the walk
    const bool mix = sizeof(R) == 8 && p.precision == PACH_PREC_MIXED;
    DevScene<float> sg = dev_scene<float>(c);
    sg.cell_rmask = s.cell_rmask;
    sg.dist_lds = s.dist_lds;
    mix_bounds(c, tc);

    DevRecv<R> rv{};
    rv.ctr = ctr_of<R>(c);
    rv.n = c->R;
    rv.masked = c->rmask_on ? 1 : 0;
    rv.r2 = (R)(c->radius * c->radius);
    rv.sizemod = 1.0 / (M_PI * c->radius * c->radius * c->radius);
    rv.c = c->c;
    rv.inv_c = 1 / c->c;
    rv.band = c->d_band;
    rv.rho_c = c->rho_c;
    rv.fs = (int)c->fs;
    rv.sct = c->sct;
    rv.hist = c->d_hist;
    rv.plane_stride = c->hist_n;
    rv.has_map = 0;
    double fx_inv_se = 0, fx_inv_sp = 0;
    if (c->hist_exact) {
        if (c->map_on) return fail(c, PACH_EINVAL, "exact histograms cover the spherical receivers, not a map");
        const size_t need = (size_t)kPlanes * c->hist_n;
        if (c->fx_cap < need) {
            dfree(c->d_fx);
            c->fx_cap = 0;
            if (dalloc(c->d_fx, 4 * need) != hipSuccess) return fail(c, PACH_ENOMEM, "exact histogram allocation");
            c->fx_cap = need;
        }
        // Fixed-point scales from the scene alone (never the ray count, so a
        // deposit becomes the same integer in every trace of the scene): a
        // deposit is at most u = P / (pi r^2) (intensity <= the source power P,
        // half-chord <= r, SizeMod 1 / (pi r^3)), doubled for FP32 rounding,
        // and a bin takes at most 16 deposits from each of up to 2^40 source
        // rays.  2^s maps that bound, 2^44 u, to 2^120 (the flush needs the
        // sums below 2^127), so one unit is 2^-76 u.
        double pmax = 0;
        for (const RoomTab &t : rooms)
            for (int o = 0; o < kBands; ++o) pmax = std::max(pmax, t.power[o]);
        const double ue = 2.0 * std::max(pmax, 1e-300) / (M_PI * c->radius * c->radius);
        const double up = std::sqrt(ue * c->rho_c);
        const int se = 120 - 44 - (int)std::ceil(std::log2(ue));
        const int sp = 120 - 44 - (int)std::ceil(std::log2(up));
        if (se < -1000 || se > 1000 || sp < -1000 || sp > 1000)
            return fail(c, PACH_EINVAL, "source power outside the exact histograms' range");
        rv.fx = c->d_fx;
        rv.fx_se = std::ldexp(1.0, se);
        rv.fx_sp = std::ldexp(1.0, sp);
        fx_inv_se = std::ldexp(1.0, -se);
        fx_inv_sp = std::ldexp(1.0, -sp);
    }
    // exact accumulation: the limb sums into the FP64 planes
    auto fx_flush = [&]() -> hipError_t {
        return rv.fx ? launch_fx_flush(c->stream, c->d_fx, c->d_hist, c->hist_n, fx_inv_se, fx_inv_sp) : hipSuccess;
    };
    if (c->map_on) {
        if (rooms.size() != 1) return fail(c, PACH_EINVAL, "mapping receivers trace one room");
        // Direct_Time of each map receiver from the current source (Map_Receiver ctor)
        std::vector<double4> ctr(c->map_n);
        for (int i = 0; i < c->map_n; ++i) {
            const double *P = &c->map_ctr[3 * i];
            const double lx = c->origin[0] - P[0], ly = c->origin[1] - P[1], lz = c->origin[2] - P[2];
            ctr[i] = make_double4(P[0], P[1], P[2], std::sqrt(lx * lx + ly * ly + lz * lz) / c->c);
        }
        HIPCHK(c, hipMemcpy(c->d_map_ctr, ctr.data(), c->map_n * sizeof(double4), hipMemcpyHostToDevice));
        DevMap &m = rv.map;
        m.ctr = c->d_map_ctr;
        m.start = c->d_map_start;
        m.list = c->d_map_list;
        m.lctr = c->d_map_lctr;
        m.scr = c->map_scr;
        for (int k = 0; k < 3; ++k) {
            // a deposit needs a point of the segment within r of a centre, so
            // inside [bmin, bmax]; widened by r and by 1e-5 of the box's
            // coordinates against the rounding of segment ends and of FP32
            const double pad = c->map_r + 1e-5 * (1 + std::max(std::fabs(c->map_bmin[k]), std::fabs(c->map_bmax[k])));
            m.lo[k] = (float)(c->map_bmin[k] - pad);
            m.hi[k] = (float)(c->map_bmax[k] + pad);
            m.bmin[k] = c->map_bmin[k];
            m.bmax[k] = c->map_bmax[k];
            m.vd[k] = c->map_vd[k];
            m.ct[k] = c->map_ct[k];
        }
        m.n = c->map_n;
        m.r2 = c->map_r * c->map_r;
        m.sizemod = 1 / (M_PI * m.r2);
        m.c = c->c;
        m.rho_c = c->rho_c;
        m.directional = c->map_directional;
        m.sct = c->map_sct;
        m.time1pt = c->map_time1pt;
        m.hist = c->d_map_hist;
        m.plane_stride = c->map_hist_n;
        rv.has_map = 1;
    }

    const int64_t N = rooms.back().ray_base + rooms.back().ray_count;
    // 131,072 source rays per broadband batch: each octave launch then runs
    // long enough that the drain at its end costs under 1 % (6 launches for 1e6
    // hall rays instead of 33 at 32,768: 1 % faster)
    int64_t batch = p.batch_rays > 0 ? p.batch_rays : std::min<int64_t>(N, 131072);
    batch = std::max<int64_t>(1, std::min<int64_t>(batch, 1 << 22));
    OctShape shape = c->tune_shape;
    if (shape.occ < 0) shape.occ = sizeof(R) == 4 ? 4 : 3;
    if (mix) shape.queue = 0;  // the mixed walk shades in place
    if (c->map_on) shape = OctShape{256, 0, 0};  // the map variant's launch shape
    const int lds_use = s.dist_lds ? lds_bytes : 0;
    const int grid = c->n_cu * (c->map_on ? octave_map_blocks_per_cu<R>(lds_use)
                                : mix    ? octave_mix_blocks_per_cu(shape, lds_use)
                                         : octave_blocks_per_cu<R>(shape, lds_use));
    int rc = ensure_work<R>(c, batch, (size_t)grid * (shape.block / 64), shape.queue != 0, mix);
    if (rc) return rc;
    WorkBuf<R> &w = work_of<R>(c);
    unsigned *p1_count = c->d_scal + 0, *res_count = c->d_scal + 1, *in_head = c->d_scal + 2;
    unsigned *dump_count = c->d_scal + 3;  // [2]
    unsigned *xq_count = c->d_scal + 5;
    unsigned *live_count = c->d_scal + 6;
    hipStream_t st = c->stream;
    HIPCHK(c, hipMemcpyAsync(c->d_band, &bt, sizeof(bt), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemsetAsync(c->d_hist, 0, c->hist_n * kPlanes * sizeof(double), st));
    if (rv.fx) HIPCHK(c, hipMemsetAsync(c->d_fx, 0, 4 * c->hist_n * kPlanes * sizeof(long long), st));
    if (c->map_on) HIPCHK(c, hipMemsetAsync(c->d_map_hist, 0, c->map_hist_n * 8 * sizeof(double), st));
    HIPCHK(c, hipMemsetAsync(c->d_cnt, 0, sizeof(DevCounters), st));
    HIPCHK(c, hipMemsetAsync(c->d_scal, 0, 8 * sizeof(unsigned), st));

    double ms_p1 = 0, ms_oct = 0, ms_exact = 0;
    int64_t n_p1 = 0, n_oct = 0;
    int64_t started = 0;
    const int min_iters = c->tune_dump_min;
    // Source rays [first, first + count): broadband batches feed the octave
    // kernel together with the previous launch's leftovers until every ray
    // they spawned has ended.
    auto trace_range = [&](int64_t first, int64_t count) -> int {
        int cur = 0;            // dump buffer holding the previous launch's leftovers
        unsigned pending = 0;   // rays in dump[cur]
        unsigned n_resume = 0;  // broadband rays waiting in res[res_cur]
        int res_cur = 0;
        int64_t next = first;   // next fresh source ray
        const int64_t end = first + count;
        const size_t words = (size_t)(count + 31) / 32;
        if (words > c->live_words) {
            HIPCHK(c, hipStreamSynchronize(st));
            HIPCHK(c, dalloc(c->d_live, words));
            c->live_words = words;
        }
        HIPCHK(c, hipMemsetAsync(c->d_scal, 0, 8 * sizeof(unsigned), st));
        for (;;) {
            const bool stop_new = c->cancel.load();
            bool p1_launched = false;
            HIPCHK(c, hipMemsetAsync(p1_count, 0, 2 * sizeof(unsigned), st));  // p1_count, res_count
            if (n_resume > 0) {
                HIPCHK(c, hipEventRecord(c->ev[0], st));
                HIPCHK(c, launch_phase1<R>(st, tc, s, rv, 0, 0, w.res[res_cur], (int)n_resume, w.p1_out, p1_count,
                                           w.p1_cap, w.res[1 - res_cur], res_count, c->d_cnt));
                HIPCHK(c, hipEventRecord(c->ev[1], st));
                res_cur = 1 - res_cur;
                p1_launched = true;
            } else if (next < end && !stop_new && (size_t)pending + w.p1_cap <= w.dump_cap) {
                // (a new batch only while the hand-over queue can take its output
                // beside the backlog: a launch that drained early leaves work behind)
                const int nb = (int)std::min<int64_t>(batch, end - next);
                HIPCHK(c, hipEventRecord(c->ev[0], st));
                HIPCHK(c, launch_phase1<R>(st, tc, s, rv, next, nb, nullptr, 0, w.p1_out, p1_count, w.p1_cap,
                                           w.res[1 - res_cur], res_count, c->d_cnt));
                HIPCHK(c, hipEventRecord(c->ev[1], st));
                res_cur = 1 - res_cur;
                next += nb;
                started += nb;
                p1_launched = true;
            }
            if (!p1_launched && pending == 0) break;
            HIPCHK(c, hipMemsetAsync(in_head, 0, sizeof(unsigned), st));
            HIPCHK(c, hipMemsetAsync(dump_count + (1 - cur), 0, sizeof(unsigned), st));
            if (mix) HIPCHK(c, hipMemsetAsync(xq_count, 0, sizeof(unsigned), st));
            HIPCHK(c, hipEventRecord(c->ev[2], st));
            if constexpr (sizeof(R) == 8) {
                if (mix) {
                    HIPCHK(c, launch_octave_mix(st, grid, shape, tc, s, sg, rv, pending ? w.dump[cur] : nullptr,
                                                dump_count + cur, w.dump_cap, p1_launched ? w.p1_out : nullptr,
                                                p1_count, w.p1_cap, in_head, w.stacks, w.stack_cap, w.dump[1 - cur],
                                                dump_count + (1 - cur), w.dump_cap, min_iters, c->d_cnt, w.xq,
                                                xq_count, w.xq_cap));
                }
            }
            if (!mix)
                HIPCHK(c, launch_octave<R>(st, grid, shape, tc, s, sg, rv, pending ? w.dump[cur] : nullptr,
                                           dump_count + cur, w.dump_cap, p1_launched ? w.p1_out : nullptr, p1_count,
                                           w.p1_cap, in_head, w.stacks, w.stack_cap,
                                           shape.queue ? w.shq : nullptr, w.dump[1 - cur], dump_count + (1 - cur),
                                           w.dump_cap, min_iters, c->d_cnt));
            HIPCHK(c, hipEventRecord(c->ev[3], st));
            if constexpr (sizeof(R) == 8) {
                if (mix) {
                    // the bounces the FP32 screen left undecided, by the FP64 walk, into the next input
                    HIPCHK(c, launch_exact(st, c->n_cu * 4, tc, s, rv, w.xq, xq_count, w.xq_cap, w.dump[1 - cur],
                                           dump_count + (1 - cur), w.dump_cap, c->d_cnt));
                    HIPCHK(c, hipEventRecord(c->ev[4], st));
                }
            }
            // progress: the source rays started in this range that still have
            // rays in the next launch's input (octave or suspended broadband)
            HIPCHK(c, hipMemsetAsync(c->d_live, 0, (size_t)(next - first + 31) / 32 * sizeof(uint32_t), st));
            HIPCHK(c, hipMemsetAsync(live_count, 0, sizeof(unsigned), st));
            HIPCHK(c, launch_live_sources<R>(st, w.dump[1 - cur], dump_count + (1 - cur), w.dump_cap,
                                             w.res[res_cur], res_count, (unsigned)w.res_cap, (uint32_t)first,
                                             (uint32_t)count,
                                             c->d_live, live_count));
            unsigned h[8];
            HIPCHK(c, hipMemcpyAsync(h, c->d_scal, sizeof(h), hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
            float ms = 0;
            if (p1_launched && hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) { ms_p1 += ms; ++n_p1; }
            if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) { ms_oct += ms; ++n_oct; }
            if (mix && hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) ms_exact += ms;
            n_resume = h[1];
            cur = 1 - cur;
            pending = std::min(h[3 + cur], w.dump_cap);
            if (n_resume > w.res_cap) return fail(c, PACH_EHIP, "resume buffer overflow");
            // every launch completes the bounces its waves start, so this only trips on a defect
            if (n_oct > ((int64_t)1 << 20)) return fail(c, PACH_EHIP, "octave launches exceed 2^20: no progress");
            // a source ray is complete once its broadband phase and every ray it
            // spawned have ended (earlier ranges are complete when this one starts)
            c->rays_done = started - (int64_t)h[6];
        }
        return PACH_OK;
    };
    int64_t revolutions = 0;
    int conv_vis = -1, conv_occ = -1;
    bool converged = false;
    if (!conv_mode) {
        if (int e = trace_range(0, N)) return e;
    } else {
        const int kind = p.ray_count < 0 ? PACH_CONV_MINIMUM : PACH_CONV_SPECMIN;
        const int hb = rooms[0].h;
        conv_receivers(c, conv_vis, conv_occ);
        const int recs[2] = {conv_vis, conv_occ};
        const int64_t ss = conv_state_size(kind, c->sct, (int)c->fs);
        std::vector<double> state(2 * ss, 0.0), rows(4 * (size_t)c->sct);
        const int period = 80;  // Source.Revolution_Period()
        for (;;) {
            if (int e = trace_range(revolutions * period, period)) return e;
            ++revolutions;
            if (c->cancel.load()) break;
            bool conv = true;
            HIPCHK(c, fx_flush());
            for (int k = 0; k < 2; ++k) {
                if (recs[k] < 0) continue;
                const size_t off = ((size_t)recs[k] * kBands + hb) * c->sct;
                const int planes[4] = {0, 8, 9, 10};  // energy; specular, diffuse, specular-scattered
                for (int q = 0; q < 4; ++q)
                    HIPCHK(c, hipMemcpyAsync(rows.data() + (size_t)q * c->sct, c->d_hist + planes[q] * c->hist_n + off,
                                             c->sct * sizeof(double), hipMemcpyDeviceToHost, st));
                HIPCHK(c, hipStreamSynchronize(st));
                const double *C = &c->centers[3 * recs[k]];
                const double dist = std::sqrt((C[0] - c->origin[0]) * (C[0] - c->origin[0]) +
                                              (C[1] - c->origin[1]) * (C[1] - c->origin[1]) +
                                              (C[2] - c->origin[2]) * (C[2] - c->origin[2]));
                conv &= conv_step(kind, rows.data(), rows.data() + c->sct, rows.data() + 2 * (size_t)c->sct,
                                  rows.data() + 3 * (size_t)c->sct, c->sct, (int)c->fs, dist / c->c,
                                  state.data() + k * ss);
            }
            if (conv) { converged = true; break; }
            if (revolutions >= ((int64_t)1 << 20)) break;  // 84 M source rays without convergence
        }
        // the room's ray range ends with the last revolution (scaling, e_init)
        rooms[0].ray_count = started;
        c->room_tab = rooms;
    }
    c->rays_done = started;
    HIPCHK(c, fx_flush());
    DevCounters hc{};
    HIPCHK(c, hipMemcpy(&hc, c->d_cnt, sizeof(hc), hipMemcpyDeviceToHost));
    pach_stats &S = c->stats;
    S = pach_stats{};
    S.bounces_broad = (int64_t)hc.v[C_BROAD];
    S.bounces_late = (int64_t)hc.v[C_LATE];
    S.bounces_verylate = (int64_t)hc.v[C_VLATE];
    S.octave_rays = (int64_t)hc.v[C_OCT];
    S.spawned = (int64_t)hc.v[C_SPAWN];
    S.misses = (int64_t)hc.v[C_MISS];
    S.deposits = (int64_t)hc.v[C_DEP];
    S.p1_cap_hits = (int64_t)hc.v[C_P1CAP];
    S.guard_hits = (int64_t)hc.v[C_GUARD];
    S.spills = (int64_t)hc.v[C_SPILL];
    S.resumes = (int64_t)hc.v[C_RESUME];
    S.loop_steps = (int64_t)hc.v[C_ITER];
    S.voxels = (int64_t)hc.v[C_VOXEL];
    S.tri_tests = (int64_t)hc.v[C_TEST];
    S.shade_passes = (int64_t)hc.v[C_SHADE];
    S.lanes_shaded = (int64_t)hc.v[C_SHADED];
    S.exact_rewalks = (int64_t)hc.v[C_EXACT];
    S.exact_walks = (int64_t)hc.v[C_EXACT0];
    S.ms_exact = ms_exact;
    S.lost = hc.lost;
    S.e_init = 0;
    for (const RoomTab &t : rooms) {
        double psum = 0;
        for (int o = 0; o < 8; ++o) psum += t.power[o];
        S.e_init += psum * (double)std::min(std::max<int64_t>(started - t.ray_base, 0), t.ray_count);
    }
    S.h_oct = rooms[0].h;
    S.rays_done = started;
    S.ms_phase1 = ms_p1;
    S.ms_octave = ms_oct;
    S.launches_phase1 = n_p1;
    S.launches_octave = n_oct;
    S.ms_trace = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    S.revolutions = revolutions;
    S.conv_receiver_visible = conv_vis;
    S.conv_receiver_occluded = conv_occ;
    S.converged = converged ? 1 : 0;
    c->oct_rays = S.octave_rays;
    if (hc.v[C_DROP]) return fail(c, PACH_ENOMEM, "octave-ray queue overflow: rays dropped");
    if (hc.v[C_FXOVF]) return fail(c, PACH_ERANGE, "deposit beyond the exact histograms' range");
    if (conv_mode) return c->cancel.load() && !converged ? PACH_ECANCEL : PACH_OK;
    if (c->cancel.load() && started < N) return PACH_ECANCEL;
    return PACH_OK;
}

}  // namespace

extern "C" int pach_rooms_set(pach_ctx *c, int n, const double *src, const double *power, const int64_t *rays,
                              const int *surf_count, const int *rec_count)
{
    if (!c) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (n == 0) {
        c->rooms.clear();
        c->rmask_dirty = true;
        return PACH_OK;
    }
    if (n < 0 || n > (1 << 24) || !src || !power || !rays || !surf_count || !rec_count)
        return fail(c, PACH_EINVAL, "rooms: bad count or null array");
    std::vector<pach_ctx::RoomIn> v(n);
    for (int i = 0; i < n; ++i) {
        std::memcpy(v[i].src, src + 3 * i, sizeof(v[i].src));
        std::memcpy(v[i].power, power + 8 * i, sizeof(v[i].power));
        v[i].rays = rays[i];
        v[i].surf_count = surf_count[i];
        v[i].rec_count = rec_count[i];
        if (rays[i] < 0 || surf_count[i] < 0 || rec_count[i] < 0) return fail(c, PACH_EINVAL, "rooms: negative count");
        for (int o = 0; o < 8; ++o)
            if (!(v[i].power[o] >= 0)) return fail(c, PACH_EINVAL, "rooms: bad power");
    }
    c->rooms = std::move(v);
    c->rmask_dirty = true;
    return PACH_OK;
}

// PachMapReceiver.Partition (Pach_Map_Receiver.cs:99-303): spheres of
// diameter `increment`, their bounding box, voxels of 4 x 4 x 2.5 increments,
// each listing (ascending) the receivers whose sphere reaches into it.
extern "C" int pach_map_receivers_set(pach_ctx *c, const double *centers, int M, double increment, int directional,
                                      int sample_ct, int time1pt)
{
    if (!c || !centers || M <= 0 || !(increment > 0) || sample_ct <= 0) return PACH_EINVAL;
    if (time1pt != 0 && (time1pt != 1 || sample_ct != 1)) return PACH_EINVAL;  // SampleCT = Time1Pt ? 1 : ...
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    for (int i = 0; i < 3 * M; ++i)
        if (!std::isfinite(centers[i])) return fail(c, PACH_EINVAL, "map: non-finite centre");
    (void)hipSetDevice(c->device);
    const double r = increment * 0.5, r2 = r * r;
    double bmin[3], bmax[3], vd[3], inv[3];
    int ct[3];
    for (int k = 0; k < 3; ++k) { bmin[k] = INFINITY; bmax[k] = -INFINITY; }
    for (int i = 0; i < M; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = centers[3 * i + k];
            if (v + increment * .5 > bmax[k]) bmax[k] = v + increment * .5;
            if (v - increment * .5 < bmin[k]) bmin[k] = v - increment * .5;
        }
    ct[0] = (int)std::ceil((bmax[0] - bmin[0]) / (increment * 4));
    ct[1] = (int)std::ceil((bmax[1] - bmin[1]) / (increment * 4));
    ct[2] = (int)std::ceil((bmax[2] - bmin[2]) / (increment * 2.5));
    const int64_t ncell = (int64_t)ct[0] * ct[1] * ct[2];
    if (ncell <= 0 || ncell > ((int64_t)1 << 26)) return fail(c, PACH_EINVAL, "map: receiver grid too large");
    for (int k = 0; k < 3; ++k) {
        vd[k] = (bmax[k] - bmin[k]) / ct[k];
        inv[k] = 1 / vd[k];
    }
    std::vector<std::vector<uint32_t>> cells(ncell);
    for (int i = 0; i < M; ++i) {
        const double *o = centers + 3 * i;
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::max(0, (int)std::floor((o[k] - r - bmin[k]) * inv[k]));
            hi[k] = std::min(ct[k] - 1, (int)std::floor((o[k] + r - bmin[k]) * inv[k]));
        }
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    const double mn[3] = {x * vd[0] + bmin[0], y * vd[1] + bmin[1], z * vd[2] + bmin[2]};
                    const double mx[3] = {(x + 1) * vd[0] + bmin[0], (y + 1) * vd[1] + bmin[1], (z + 1) * vd[2] + bmin[2]};
                    double dd[3];
                    for (int k = 0; k < 3; ++k) dd[k] = std::max(mn[k] - o[k], std::max(0.0, o[k] - mx[k]));
                    if (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < r2)
                        cells[((int64_t)x * ct[1] + y) * ct[2] + z].push_back((uint32_t)i);
                }
    }
    std::vector<uint32_t> start(ncell + 1, 0), list;
    for (int64_t q = 0; q < ncell; ++q) {
        start[q + 1] = start[q] + (uint32_t)cells[q].size();
        list.insert(list.end(), cells[q].begin(), cells[q].end());
    }
    if (list.empty()) list.push_back(0);
    // The walk screens each listed receiver in FP32 from a copy of its centre
    // stored with the entry (one load per entry).  Rounding a centre C to FP32
    // moves it by e <= sqrt(3) 2^-24 max|C|, which moves the line's distance
    // from it by at most e; for a sphere the line meets (distance <= r) the
    // screened discriminant r^2 - distance^2 then falls by at most
    // 2 r e + e^2, widened here by a factor of 2.
    std::vector<float4> lctr(list.size());
    double cmax = 0;
    for (int i = 0; i < 3 * M; ++i) cmax = std::max(cmax, std::fabs(centers[i]));
    for (size_t q = 0; q < list.size(); ++q) {
        const double *P = centers + 3 * (size_t)list[q];
        uint32_t j = list[q];
        float w;
        std::memcpy(&w, &j, 4);
        lctr[q] = make_float4((float)P[0], (float)P[1], (float)P[2], w);
    }
    const double e = std::sqrt(3.0) * std::ldexp(cmax, -24);
    HIPCHK(c, dalloc(c->d_map_lctr, lctr.size()));
    HIPCHK(c, hipMemcpy(c->d_map_lctr, lctr.data(), lctr.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, dalloc(c->d_map_ctr, M));
    HIPCHK(c, dalloc(c->d_map_start, (size_t)ncell + 1));
    HIPCHK(c, dalloc(c->d_map_list, list.size()));
    HIPCHK(c, hipMemcpy(c->d_map_start, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_map_list, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    const size_t per_plane = (size_t)M * kBands * sample_ct;
    HIPCHK(c, dalloc(c->d_map_hist, per_plane * 8));
    HIPCHK(c, hipMemset(c->d_map_hist, 0, per_plane * 8 * sizeof(double)));
    c->map_ctr.assign(centers, centers + 3 * M);
    c->map_n = M;
    c->map_r = r;
    c->map_scr = (float)(2 * (2 * r * e + e * e));
    c->map_sct = sample_ct;
    c->map_directional = directional ? 1 : 0;
    c->map_time1pt = time1pt;
    c->map_hist_n = per_plane;
    for (int k = 0; k < 3; ++k) {
        c->map_bmin[k] = bmin[k];
        c->map_bmax[k] = bmax[k];
        c->map_vd[k] = vd[k];
        c->map_ct[k] = ct[k];
    }
    c->map_on = true;
    return PACH_OK;
}

extern "C" int pach_map_receivers_clear(pach_ctx *c)
{
    if (!c) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    c->map_on = false;
    return PACH_OK;
}

extern "C" int pach_map_histogram(pach_ctx *c, int plane, double *out, int scaled)
{
    if (!c || !out) return PACH_EINVAL;
    if (plane < 0 || plane >= 8) return fail(c, PACH_EINVAL, "bad map plane");
    if (!c->map_on) return fail(c, PACH_ESTATE, "no mapping receivers");
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_map_hist + (size_t)plane * c->map_hist_n, c->map_hist_n * sizeof(double),
                        hipMemcpyDeviceToHost));
    if (scaled) {
        const double n = (double)c->rays_done.load();
        if (n > 0) {
            const double f = plane == PACH_PLANE_PRESSURE ? 1.0 / std::sqrt(n) : 1.0 / n;
            for (size_t i = 0; i < c->map_hist_n; ++i) out[i] *= f;
        }
    }
    return PACH_OK;
}

extern "C" int pach_exact_histograms(pach_ctx *c, int on)
{
    if (!c || (on != 0 && on != 1)) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    c->hist_exact = on != 0;
    return PACH_OK;
}

extern "C" int pach_tune(pach_ctx *c, const char *key, int64_t value)
{
    if (!c || !key) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (std::strcmp(key, "dist_lds") == 0) {
        c->tune_dist_lds = value != 0;
        return PACH_OK;
    }
    if (std::strcmp(key, "receiver_masks") == 0) {
        c->tune_rmask = value != 0;
        c->rmask_dirty = true;
        return PACH_OK;
    }
    if (std::strcmp(key, "walk_steps") == 0 && value >= 1 && value <= 256) {
        c->tune_walk_steps = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "walk_flat") == 0 && value >= 0 && value <= 2) {
        c->tune_walk_flat = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "shade_min") == 0 && value >= 1 && value <= 64) {
        c->tune_shade_min = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "refill_min") == 0 && value >= 1 && value <= 64) {
        c->tune_refill_min = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "dump_min") == 0 && value >= 1 && value <= (1 << 20)) {
        c->tune_dump_min = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "block") == 0 && (value == 256 || value == 512)) {
        c->tune_shape.block = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "exact_queue") == 0 && value >= 1 && value <= ((int64_t)1 << 30)) {
        c->tune_xq = value;
        return PACH_OK;
    }
    if (std::strcmp(key, "occupancy") == 0 && (value == -1 || value == 0 || value == 3 || value == 4 || value == 5)) {
        c->tune_shape.occ = (int)value;
        return PACH_OK;
    }
    if (std::strcmp(key, "shade_queue") == 0 && (value == 0 || value == 1)) {
        c->tune_shape.queue = (int)value;
        return PACH_OK;
    }
    return fail(c, PACH_EINVAL, std::string("unknown tuning key ") + key);
}

extern "C" int pach_trace(pach_ctx *c, const pach_params *p)
{
    if (!c || !p) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace already running");
    if (c->worker.joinable()) c->worker.join();
    if (!c->has_scene || !c->has_src || !c->has_recv)
        return fail(c, PACH_ESTATE, "scene, source and receivers must be set before tracing");
    int64_t total = p->ray_count;
    if (!c->rooms.empty()) {
        total = 0;
        for (const auto &rm : c->rooms) total += rm.rays;
        if (p->ray_count != 0 && p->ray_count != total)
            return fail(c, PACH_EINVAL, "ray_count must be 0 or the sum of the rooms' ray counts");
    }
    const bool conv_mode = c->rooms.empty() && p->ray_count <= 0;  // TaskAll:111, :236-239
    if (!conv_mode && (total <= 0 || total > ((int64_t)1 << 32)))
        return fail(c, PACH_EINVAL, "ray_count must lie in [1, 2^32] (or be < 1 for convergence mode)");
    if (conv_mode) total = 0;
    if (p->oct_lo < 0 || p->oct_hi > 7 || p->oct_lo > p->oct_hi) return fail(c, PACH_EINVAL, "bad octave range");
    if (!(p->cutoff_m > 0)) return fail(c, PACH_EINVAL, "cutoff must be positive");
    if (p->mode != PACH_MODE_STOCHASTIC && p->mode != PACH_MODE_SPECULAR) return fail(c, PACH_EINVAL, "bad mode");
    if (p->precision != PACH_PREC_FP32 && p->precision != PACH_PREC_FP64 && p->precision != PACH_PREC_MIXED)
        return fail(c, PACH_EINVAL, "bad precision");
    if (p->precision == PACH_PREC_MIXED && c->map_on)
        return fail(c, PACH_EINVAL, "mixed precision does not trace mapping receivers (use FP32 or FP64)");
    c->cancel = false;
    c->rays_done = 0;
    c->rays_total = total;
    c->status = 0;
    c->running = true;
    const pach_params pp = *p;
    auto body = [c, pp]() {
        const int rc = pp.precision == PACH_PREC_FP32 ? run_trace<float>(c, pp) : run_trace<double>(c, pp);
        c->status = rc;
        c->running = false;
        return rc;
    };
    if (!p->async) return body();
    c->worker = std::thread(body);
    return PACH_OK;
}

// ------------------------------------------------------------------ image sources
namespace {

// Reference order of the valid paths: by receiver, then the depth-first order
// of TraverseOrders (a sequence precedes its extensions).
bool is_path_before(const pach_is_path &a, const pach_is_path &b)
{
    if (a.receiver != b.receiver) return a.receiver < b.receiver;
    const int n = std::min(a.order, b.order);
    for (int q = 0; q < n; ++q)
        if (a.seq[q] != b.seq[q]) return a.seq[q] < b.seq[q];
    return a.order < b.order;
}

}  // namespace

extern "C" int pach_image_sources(pach_ctx *c, const int *object_of_surface, int n_objects, int max_order,
                                  pach_is_path *out, int64_t cap, int64_t *n_paths)
{
    if (!c || !n_paths || cap < 0 || (cap > 0 && !out)) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (c->worker.joinable()) c->worker.join();
    if (!c->has_scene || !c->has_src || !c->has_recv)
        return fail(c, PACH_ESTATE, "scene, source and receivers must be set before image sources");
    if (max_order < 1 || max_order > PACH_IS_MAX_ORDER)
        return fail(c, PACH_EINVAL, "max_order must lie in [1, PACH_IS_MAX_ORDER]");
    const int S = c->S, T = c->T;
    std::vector<int> obj(S);
    if (!object_of_surface) {
        if (n_objects != 0 && n_objects != S) return fail(c, PACH_EINVAL, "n_objects must be 0 or S without a map");
        n_objects = S;
        for (int p = 0; p < S; ++p) obj[p] = p;
    } else {
        if (n_objects < 1) return fail(c, PACH_EINVAL, "n_objects must be positive");
        for (int p = 0; p < S; ++p) {
            if (object_of_surface[p] < 0 || object_of_surface[p] >= n_objects)
                return fail(c, PACH_EINVAL, "object_of_surface out of range");
            obj[p] = object_of_surface[p];
        }
    }
    if (n_objects >= (1 << 21) || c->R >= (1 << 21))
        return fail(c, PACH_EINVAL, "image sources: at most 2^21 - 1 objects and receivers");
    // threads per order: n_objects^order x R, kept below 2^62
    {
        double tot = c->R;
        for (int o = 0; o < max_order; ++o) tot *= n_objects;
        if (tot > 4.0e18) return fail(c, PACH_EINVAL, "too many reflection sequences (n_objects^max_order x R)");
    }
    // each object's plane: its lowest polygon (ObjectMembers[o][0]), that
    // polygon's first triangle (unit normal as uploaded, vertex 0)
    std::vector<int> first_poly(n_objects, -1), first_tri(n_objects, -1);
    for (int p = 0; p < S; ++p)
        if (first_poly[obj[p]] < 0) first_poly[obj[p]] = p;
    for (int t = 0; t < T; ++t) {
        const int o = obj[c->surf[t]];
        if (first_poly[o] == c->surf[t] && first_tri[o] < 0) first_tri[o] = t;
    }
    // triangles of each object (CSR), for the walk-free first test
    std::vector<int> obj_tri0(n_objects + 1, 0), obj_tris(T);
    for (int t = 0; t < T; ++t) obj_tri0[obj[c->surf[t]] + 1]++;
    for (int o = 0; o < n_objects; ++o) obj_tri0[o + 1] += obj_tri0[o];
    {
        std::vector<int> fill(obj_tri0.begin(), obj_tri0.end() - 1);
        for (int t = 0; t < T; ++t) obj_tris[fill[obj[c->surf[t]]]++] = t;
    }
    std::vector<IsPlane> planes(n_objects);
    for (int o = 0; o < n_objects; ++o) {
        IsPlane &pl = planes[o];
        std::memset(&pl, 0, sizeof(pl));
        const int t = first_tri[o];
        if (t < 0) continue;  // an object without triangles is never hit
        const double *v = &c->verts[9 * (size_t)t];
        const double a[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const double b[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
            pl.n[k] = l > 0 ? n[k] / l : 0.0;
            pl.p[k] = v[k];
        }
    }
    (void)hipSetDevice(c->device);
    const auto t_start = std::chrono::steady_clock::now();
    c->is_stats = pach_is_stats{};
    constexpr unsigned long long kChunk = 1ull << 16;  // device result records per launch
    int *d_obj = nullptr, *d_obj_tri0 = nullptr, *d_obj_tris = nullptr;
    IsPlane *d_plane = nullptr;
    pach_is_path *d_out = nullptr;
    unsigned long long *d_count = nullptr;
    uint32_t *d_queue = nullptr;
    unsigned *d_queued = nullptr;
    constexpr int64_t kSpan = (int64_t)1 << 26;  // items per launch at most
    std::vector<pach_is_path> found;
    auto body = [&]() -> int {
        HIPCHK(c, dalloc(d_obj, S));
        HIPCHK(c, dalloc(d_obj_tri0, n_objects + 1));
        HIPCHK(c, dalloc(d_obj_tris, T));
        HIPCHK(c, hipMemcpy(d_obj_tri0, obj_tri0.data(), (n_objects + 1) * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d_obj_tris, obj_tris.data(), T * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, dalloc(d_plane, n_objects));
        HIPCHK(c, dalloc(d_out, kChunk));
        HIPCHK(c, dalloc(d_count, 1));
        if (!c->any_trans) {
            // screening queue: without transmissive polygons the first hit decides
            HIPCHK(c, dalloc(d_queue, (size_t)kSpan));
            HIPCHK(c, dalloc(d_queued, 1));
        }
        HIPCHK(c, hipMemcpy(d_obj, obj.data(), S * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d_plane, planes.data(), n_objects * sizeof(IsPlane), hipMemcpyHostToDevice));
        IsJob j{};
        j.obj = d_obj;
        j.plane = d_plane;
        j.rec = c->d_ctr64;
        j.obj_tri0 = d_obj_tri0;
        j.obj_tris = d_obj_tris;
        j.any_trans = c->any_trans ? 1 : 0;
        j.n_obj = n_objects;
        j.inv_R = 1.0 / c->R;
        j.inv_N = 1.0 / n_objects;
        j.R = c->R;
        for (int k = 0; k < 3; ++k) j.src[k] = c->origin[k];
        for (int b = 0; b < 8; ++b) {
            j.power[b] = c->power[b];
            j.atten[b] = c->atten[b];
        }
        j.c = c->c;
        j.rho_c = c->rho_c;
        j.out = d_out;
        j.count = d_count;
        j.cap = kChunk;
        const DevScene<double> s = dev_scene<double>(c);
        const int grid = std::max(1, c->n_cu) * 8;
        hipStream_t st = c->stream;
        for (int order = 1; order <= max_order; ++order) {
            j.order = order;
            int64_t total = c->R;
            for (int o = 0; o < order; ++o) total *= n_objects;
            // launches hold whole sequences: spans are multiples of R
            const int64_t span_max = kSpan / c->R * c->R;
            int64_t span = std::min<int64_t>(total, span_max);
            for (int64_t lo = 0; lo < total;) {
                const int64_t hi = std::min(total, lo + span);
                unsigned long long n = 0;
                HIPCHK(c, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), st));
                if (d_queued) HIPCHK(c, hipMemsetAsync(d_queued, 0, sizeof(unsigned), st));
                HIPCHK(c, hipEventRecord(c->ev[0], st));
                // receiver and sequence digits of item lo
                int64_t rest = lo / c->R;
                j.base_rec = (int)(lo % c->R);
                for (int q = order - 1; q >= 0; --q) {
                    j.base_seq[q] = (int)(rest % n_objects);
                    rest /= n_objects;
                }
                HIPCHK(c, launch_image_sources(st, grid, s, j, (uint32_t)(hi - lo), d_queue, d_queued));
                HIPCHK(c, hipEventRecord(c->ev[1], st));
                HIPCHK(c, hipMemcpyAsync(&n, d_count, sizeof(n), hipMemcpyDeviceToHost, st));
                HIPCHK(c, hipStreamSynchronize(st));
                float ms = 0;
                HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                c->is_stats.ms_kernel += ms;
                c->is_stats.launches++;
                // more valid paths than the buffer holds: redo the range smaller,
                // or with a larger buffer when it is a single sequence
                const IsSpanStep step = is_after_launch(span, lo, hi, n, j.cap, c->R, span_max);
                span = step.span;
                if (step.grow_to) {
                    HIPCHK(c, dalloc(d_out, (size_t)step.grow_to));
                    j.out = d_out;
                    j.cap = step.grow_to;
                }
                if (step.redo) continue;
                c->is_stats.checks += hi - lo;
                const size_t old = found.size();
                found.resize(old + n);
                if (n) HIPCHK(c, hipMemcpy(found.data() + old, d_out, n * sizeof(pach_is_path), hipMemcpyDeviceToHost));
                lo = hi;
            }
        }
        return PACH_OK;
    };
    const int rc = body();
    dfree(d_obj); dfree(d_obj_tri0); dfree(d_obj_tris); dfree(d_plane); dfree(d_out); dfree(d_count);
    dfree(d_queue); dfree(d_queued);
    if (rc) return rc;
    std::stable_sort(found.begin(), found.end(), is_path_before);
    *n_paths = (int64_t)found.size();
    c->is_stats.paths = (int64_t)found.size();
    c->is_stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    const size_t w = std::min<size_t>(found.size(), (size_t)cap);
    if (w) std::memcpy(out, found.data(), w * sizeof(pach_is_path));
    return (int64_t)found.size() > cap ? fail(c, PACH_ERANGE, "image sources: output buffer too small") : PACH_OK;
}

extern "C" int pach_image_sources_stats(pach_ctx *c, pach_is_stats *out)
{
    if (!c || !out) return PACH_EINVAL;
    *out = c->is_stats;
    return PACH_OK;
}

extern "C" int pach_autotune(pach_ctx *c, const pach_params *p, int64_t sample_rays, pach_tuning *out)
{
    if (!c || !p || !out) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (c->worker.joinable()) c->worker.join();
    if (!c->has_scene || !c->has_src || !c->has_recv)
        return fail(c, PACH_ESTATE, "scene, source and receivers must be set before tuning");
    pach_params pp = *p;
    pp.async = 0;
    const std::vector<pach_ctx::RoomIn> saved = c->rooms;
    int64_t total = p->ray_count;
    if (!saved.empty()) {
        total = 0;
        for (const auto &rm : saved) total += rm.rays;
    }
    int64_t sample = sample_rays > 0 ? std::min(sample_rays, total) : std::min<int64_t>(total, 8192);
    if (sample <= 0) return fail(c, PACH_EINVAL, "autotune: no rays to sample");
    auto set_sample = [&](int64_t smp) {
        if (!saved.empty()) {
            // every room keeps its share of the sample (at least one ray if it has any)
            int64_t n = 0;
            c->rooms = saved;
            for (auto &rm : c->rooms) {
                if (rm.rays > 0)
                    rm.rays = std::max<int64_t>(1, (int64_t)std::llround((double)rm.rays * (double)smp / (double)total));
                n += rm.rays;
            }
            pp.ray_count = n;
        } else {
            pp.ray_count = smp;
        }
    };
    set_sample(sample);
    struct Knobs {
        int grid[3], block, occ, queue, walk_flat, walk_steps, shade_min, refill_min, dump_min;
    };
    auto current = [c]() {
        Knobs k{{c->g[0], c->g[1], c->g[2]}, c->tune_shape.block, c->tune_shape.occ, c->tune_shape.queue,
                c->tune_walk_flat, c->tune_walk_steps, c->tune_shade_min, c->tune_refill_min, c->tune_dump_min};
        return k;
    };
    auto apply = [c](const Knobs &k) -> int {
        if (k.grid[0] != c->g[0] || k.grid[1] != c->g[1] || k.grid[2] != c->g[2])
            if (int rc = upload_grid(c, k.grid)) return rc;
        c->tune_shape.block = k.block;
        c->tune_shape.occ = k.occ;
        c->tune_shape.queue = k.queue;
        c->tune_walk_flat = k.walk_flat;
        c->tune_walk_steps = k.walk_steps;
        c->tune_shade_min = k.shade_min;
        c->tune_refill_min = k.refill_min;
        c->tune_dump_min = k.dump_min;
        return PACH_OK;
    };
    int trials = 0;
    auto rate = [&](double &r) -> int {
        if (int rc = pach_trace(c, &pp)) return rc;
        ++trials;
        const pach_stats &S = c->stats;
        r = (double)(S.bounces_broad + S.bounces_late + S.bounces_verylate) / (S.ms_trace * 1e-3);
        return PACH_OK;
    };
    const Knobs start = current();
    int rc = PACH_OK;
    double r0 = 0, best_r = 0;
    Knobs best = start;
    do {
        double warm;
        if ((rc = rate(warm))) break;  // first trial allocates the work buffers
        if (sample_rays <= 0 && sample < total) {
            // A trial must be long enough for the launch's drain (few rays left,
            // waves emptying) to be a small part of it, or settings that drain
            // well win on the sample and lose on the full trace: size the
            // default sample to ~2e9 bounces from the bounces per source ray
            // of the first trial.
            const pach_stats &S = c->stats;
            const double bpr = (double)(S.bounces_broad + S.bounces_late + S.bounces_verylate) /
                               (double)std::max<int64_t>(pp.ray_count, 1);
            const int64_t want = (int64_t)std::ceil(2e9 / std::max(bpr, 1.0));
            if (want > sample) {
                sample = std::min(total, want);
                set_sample(sample);
                if ((rc = rate(warm))) break;
            }
        }
        if ((rc = rate(r0))) break;
        best_r = r0;
        auto try_knob = [&](auto set, const std::vector<int> &vals) -> int {
            const Knobs base = best;
            for (int v : vals) {
                Knobs k = base;
                set(k, v);
                if (std::memcmp(&k, &base, sizeof(Knobs)) == 0) continue;
                if (int e = apply(k)) return e;
                double r;
                if (int e = rate(r)) return e;
                if (r > best_r * 1.01) { best_r = r; best = k; }
            }
            return apply(best);
        };
        // voxel resolution: linear scale of the starting grid (the dominant knob first)
        const std::vector<int> pct = {50, 71, 141, 200};
        if ((rc = try_knob([&](Knobs &k, int q) {
                 for (int a = 0; a < 3; ++a)
                     k.grid[a] = std::max(1, std::min(1024, (int)std::lround(start.grid[a] * q / 100.0)));
             }, pct)))
            break;
        if ((rc = try_knob([](Knobs &k, int v) { k.queue = v; }, {0, 1}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.occ = v; }, {0, 4, 5}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.block = v; }, {256, 512}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.walk_flat = v; }, {1, 2}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.walk_steps = v; }, {6, 8, 12, 16, 24, 32}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.shade_min = v; }, {8, 12, 16, 24, 32, 48}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.refill_min = v; }, {1, 4, 8, 16, 32}))) break;
        if ((rc = try_knob([](Knobs &k, int v) { k.dump_min = v; }, {64, 256, 1024}))) break;
    } while (false);
    c->rooms = saved;
    if (rc) {
        const std::string msg = c->err;
        apply(start);
        c->err = msg;
        return rc;
    }
    if ((rc = apply(best))) return rc;
    for (int a = 0; a < 3; ++a) out->grid[a] = best.grid[a];
    out->block = best.block;
    out->occupancy = best.occ;
    out->shade_queue = best.queue;
    out->walk_flat = best.walk_flat;
    out->walk_steps = best.walk_steps;
    out->shade_min = best.shade_min;
    out->refill_min = best.refill_min;
    out->dump_min = best.dump_min;
    out->trials = trials;
    out->rate_start = r0;
    out->rate_best = best_r;
    return PACH_OK;
}

extern "C" int64_t pach_convergence_state_size(int kind, int n, int fs)
{
    if ((kind != PACH_CONV_MINIMUM && kind != PACH_CONV_SPECMIN) || n <= 0 || fs <= 0) return -1;
    return conv_state_size(kind, n, fs);
}

extern "C" int pach_convergence_step(int kind, const double *energy, const double *spec, const double *diff,
                                     const double *rogue, int n, int fs, double direct_s, double *state,
                                     int *converged)
{
    if ((kind != PACH_CONV_MINIMUM && kind != PACH_CONV_SPECMIN) || !energy || !state || !converged || n <= 0 ||
        fs <= 0)
        return PACH_EINVAL;
    if (kind == PACH_CONV_SPECMIN && (!spec || !diff || !rogue)) return PACH_EINVAL;
    *converged = conv_step(kind, energy, spec, diff, rogue, n, fs, direct_s, state) ? 1 : 0;
    return PACH_OK;
}

extern "C" int pach_poll(pach_ctx *c, pach_progress *out)
{
    if (!c || !out) return PACH_EINVAL;
    out->rays_done = c->rays_done.load();
    out->rays_total = c->rays_total.load();
    out->octave_rays = c->oct_rays.load();
    out->running = c->running ? 1 : 0;
    out->status = c->status;
    return PACH_OK;
}

extern "C" int pach_wait(pach_ctx *c)
{
    if (!c) return PACH_EINVAL;
    if (c->worker.joinable()) c->worker.join();
    return c->status;
}

extern "C" int pach_cancel(pach_ctx *c)
{
    if (!c) return PACH_EINVAL;
    c->cancel = true;
    return PACH_OK;
}

extern "C" int pach_histogram(pach_ctx *c, int plane, double *out, int scaled)
{
    if (!c || !out) return PACH_EINVAL;
    if (plane < 0 || plane >= kPlanes) return fail(c, PACH_EINVAL, "bad plane");
    if (!c->has_recv) return fail(c, PACH_ESTATE, "no receivers");
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_hist + (size_t)plane * c->hist_n, c->hist_n * sizeof(double), hipMemcpyDeviceToHost));
    if (scaled) {
        // Receiver_Bank.Scale with each room's completed source rays
        const int64_t done = c->rays_done.load();
        const size_t per_rec = (size_t)8 * c->sct;
        for (const RoomTab &t : c->room_tab) {
            const double n = (double)std::min(std::max<int64_t>(done - t.ray_base, 0), t.ray_count);
            if (!(n > 0)) continue;
            const double f = plane == PACH_PLANE_PRESSURE ? 1.0 / std::sqrt(n) : 1.0 / n;
            for (size_t i = (size_t)t.rec_base * per_rec; i < (size_t)(t.rec_base + t.rec_count) * per_rec; ++i)
                out[i] *= f;
        }
    }
    return PACH_OK;
}

// ------------------------------------------------------------------ auralisation filters
namespace {

// GenerateBandLimitedSincPulse (Pach_SP.cs:1678-1700): sinc(bw t) cos(2 pi fc t)
// about the centre sample, scaled by 2 / (L1 norm x length).
void band_carrier(int fs, double fl, double fu, double *c)
{
    const double fc = (fl + fu) / 2.0, bw = fu - fl, t0 = (kCarrierLen - 1) / 2.0;
    double norm = 0;
    for (int n = 0; n < kCarrierLen; ++n) {
        const double t = (n - t0) / fs;
        const double sinc = bw == 0 ? 1.0 : std::sin(M_PI * bw * t) / (M_PI * bw * t);
        c[n] = sinc * std::cos(2 * M_PI * fc * t);
        norm += std::fabs(c[n]);
    }
    norm = 2 / (norm * kCarrierLen);
    if (norm > 0)
        for (int n = 0; n < kCarrierLen; ++n) c[n] *= norm;
}

// Medium_Properties.ISO9613_1_attencoef, scalar form (Classes_Medium_Props.cs:
// 193-202), with its C# integer exponents (-1/2, -1/3 and 1/2 are 0, -5/2 is -2).
double iso9613_scalar(double f, double tk, double pa, double hr)
{
    const double h = hr * std::pow(10, -6.8346 * std::pow(273.16 / tk, 1.261) + 4.6151) / (pa / 101.325);
    const double fro = (pa / 101.325) * (24 + 4.04 * std::pow(10, 4) * h * ((0.02 + h) / (0.391 + h)));
    const double frn = (pa / 101.325) * std::pow(tk / 293.15, 0) *
                       (9 + 280 * h * std::exp(-4.170 * (std::pow(tk / 293.15, 0) - 1)));
    const double z = 0.1068 * std::exp(-3352 / tk) * std::pow(frn + f * f / frn, -1);
    const double y = std::pow(tk / 293.15, -2) * (0.01275 * std::exp(-2239.1 / tk) * std::pow(fro + f * f / fro, -1) + z);
    return 8.686 * f * f * ((1.84 * std::pow(10, -11) * std::pow(pa / 101.325, -1) * std::pow(tk / 293.15, 0)) + y);
}

// Tables of Pach_SP.ETCToFilter (Pach_SP.cs:1584-1676) for T samples at fs:
// the 8 octave-band carriers and the 16 kHz one, the level factors
// 10^((120 - SWL) / 20) and the 16 kHz band's per-sample factors (the 8 kHz
// envelope with the extra air attenuation over 343 m/s x t, ISO 9613-1 at
// 20 C, 50 %, pressure passed in Pa).
struct FiltTables {
    std::vector<double> car, f16;
    double p_mod[8];
};

// The nine carriers at sample rate fs ([9][kCarrierLen]); they depend only on
// fs, so the last rate's set is kept.
const std::vector<double> &carriers(int fs)
{
    static std::mutex mu;
    static int cached_fs = 0;
    static std::vector<double> car;
    std::lock_guard<std::mutex> lock(mu);
    if (cached_fs != fs) {
        car.assign((size_t)kFiltBands * kCarrierLen, 0.0);
        for (int b = 0; b < 8; ++b) {
            const double fc = 62.5 * std::pow(2, b);
            band_carrier(fs, fc / std::sqrt(2.0), fc * std::sqrt(2.0), &car[(size_t)b * kCarrierLen]);
        }
        band_carrier(fs, 16000.0 / std::sqrt(2.0), 16000.0 * std::sqrt(2.0), &car[(size_t)8 * kCarrierLen]);
        cached_fs = fs;
    }
    return car;
}

FiltTables filt_tables(const pach_ctx *c, int T, const double swl[8], int fs)
{
    FiltTables f;
    // the device copy of the carriers is replaced when the rate changes
    if (!c->f_car || c->f_car_fs != fs) f.car = carriers(fs);
    for (int b = 0; b < 8; ++b) f.p_mod[b] = std::pow(10, (120 - swl[b]) / 20.0);
    const double tk = 20 + 273.15, pa = 101.325 * 1000, hr = 50.0;
    const double extra = iso9613_scalar(16000.0, tk, pa, hr) - iso9613_scalar(8000.0, tk, pa, hr);
    f.f16.resize(T);
    for (int t = 0; t < T; ++t) f.f16[t] = std::pow(10, -extra * (t * (1.0 / fs) * 343.0) / 20.0);
    return f;
}

// Convolution of the device envelopes d_env (launch_etc_conv's layout) with
// the carriers; out [n_ch][T + kCarrierLen - 1 - trim] on the host.
int run_filters(pach_ctx *c, const double *d_env, const FiltTables &f, int fs, int n_ch, int T, double *out, int trim)
{
    const int out_len = T + kCarrierLen - 1;
    const int n_chb = (n_ch + kFiltCh - 1) / kFiltCh;
    if (!f.car.empty()) {
        size_t car_n = 0;
        HIPCHK(c, dgrow(c->f_car, car_n, f.car.size()));
        HIPCHK(c, hipMemcpy(c->f_car, f.car.data(), f.car.size() * sizeof(double), hipMemcpyHostToDevice));
        c->f_car_fs = fs;
    }
    HIPCHK(c, dgrow(c->f_part, c->f_part_n, (size_t)kFiltBands * n_ch * out_len));
    HIPCHK(c, dgrow(c->f_out, c->f_out_n, (size_t)n_ch * out_len));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, launch_etc_conv(c->stream, d_env, c->f_car, T, n_chb, n_ch, c->f_part, c->f_out));
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->filter_stats.ms_kernel = ms;
    c->filter_stats.launches = 1;
    c->filter_stats.channels = n_ch;
    c->filter_stats.samples = T;
    const int keep = out_len - trim;
    HIPCHK(c, hipMemcpy2D(out, (size_t)keep * sizeof(double), c->f_out, (size_t)out_len * sizeof(double),
                          (size_t)keep * sizeof(double), n_ch, hipMemcpyDeviceToHost));
    return PACH_OK;
}

}  // namespace

extern "C" int pach_etc_to_filter(pach_ctx *c, const double *prms, int n_ch, int T, const double swl[8],
                                  int sample_frequency, double *out)
{
    if (!c || !prms || !swl || !out || n_ch <= 0 || T <= 0 || sample_frequency <= 0) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if ((int64_t)T > (int64_t)1 << 26) return fail(c, PACH_EINVAL, "envelope too long");
    const auto t_start = std::chrono::steady_clock::now();
    (void)hipSetDevice(c->device);
    const FiltTables f = filt_tables(c, T, swl, sample_frequency);
    // envelopes |prms| x p_mod in launch_etc_conv's layout
    const int n_chb = (n_ch + kFiltCh - 1) / kFiltCh;
    std::vector<double> env((size_t)kFiltBands * n_chb * T * kFiltCh, 0.0);
    for (int ch = 0; ch < n_ch; ++ch) {
        const double *pr = prms + (size_t)ch * 8 * T;
        const int cb = ch / kFiltCh, cl = ch % kFiltCh;
        for (int b = 0; b < 8; ++b)
            for (int t = 0; t < T; ++t)
                env[(((size_t)b * n_chb + cb) * T + t) * kFiltCh + cl] = std::fabs(pr[(size_t)b * T + t]) * f.p_mod[b];
        for (int t = 0; t < T; ++t)
            env[(((size_t)8 * n_chb + cb) * T + t) * kFiltCh + cl] =
                std::fabs(pr[(size_t)7 * T + t]) * f.p_mod[7] * f.f16[t];
    }
    int rc = PACH_OK;
    if (dgrow(c->f_env, c->f_env_n, env.size()) != hipSuccess ||
        hipMemcpy(c->f_env, env.data(), env.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, PACH_EHIP, "filter envelopes: device copy failed");
    if (!rc) rc = run_filters(c, c->f_env, f, sample_frequency, n_ch, T, out, 0);
    c->filter_stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return rc;
}

static_assert(kNoAllocScratch == PACH_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

extern "C" int pach_etc_to_filter_noalloc(pach_ctx *c, const double *d_prms, int n_ch, int T, const double swl[8],
                                          int sample_frequency, double *d_out)
{
    if (!c || !d_prms || !swl || !d_out || n_ch <= 0 || T <= 0 || sample_frequency <= 0) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    const auto t_start = std::chrono::steady_clock::now();
    // scratch layout in doubles, every part 256-byte aligned: carriers, the
    // p_mod / f16 tables, the envelopes, the nine band planes
    auto al = [](size_t n) { return (n + 31) & ~(size_t)31; };
    const size_t out_len = (size_t)T + kCarrierLen - 1;
    const size_t n_chb = ((size_t)n_ch + kFiltCh - 1) / kFiltCh;
    const size_t o_tab = al((size_t)kFiltBands * kCarrierLen);
    const size_t o_env = al(o_tab + 8 + (size_t)T);
    const size_t o_part = al(o_env + (size_t)kFiltBands * n_chb * kFiltCh * T);
    const size_t need = (o_part + (size_t)kFiltBands * n_ch * out_len) * sizeof(double);
    if (need > kNoAllocScratch)
        return fail(c, PACH_ERANGE,
                    "pach_etc_to_filter_noalloc: " + std::to_string(n_ch) + " channels of " + std::to_string(T) +
                        " samples need " + std::to_string(need) + " scratch bytes, above the fixed " +
                        std::to_string(kNoAllocScratch) + " (use pach_etc_to_filter)");
    (void)hipSetDevice(c->device);
    unsigned char *base = c->noalloc_scratch;
    if (!base) HIPCHK(c, noalloc_scratch(c->device, &base));
    double *sc = reinterpret_cast<double *>(base);
    const std::vector<double> &car = carriers(sample_frequency);
    std::vector<double> tab(8 + (size_t)T);
    {
        const FiltTables f = filt_tables(c, T, swl, sample_frequency);
        std::copy(f.p_mod, f.p_mod + 8, tab.begin());
        std::copy(f.f16.begin(), f.f16.end(), tab.begin() + 8);
    }
    hipStream_t st = c->stream;
    HIPCHK(c, hipMemcpyAsync(sc, car.data(), car.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(sc + o_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(c, launch_env_from_prms(st, d_prms, n_ch, T, sc + o_tab, sc + o_tab + 8, (int)n_chb, sc + o_env));
    HIPCHK(c, hipEventRecord(c->ev[2], st));
    HIPCHK(c, launch_etc_conv(st, sc + o_env, sc, T, (int)n_chb, n_ch, sc + o_part, d_out));
    HIPCHK(c, hipEventRecord(c->ev[3], st));
    HIPCHK(c, hipStreamSynchronize(st));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->filter_stats.ms_kernel = ms;
    c->filter_stats.launches = 1;
    c->filter_stats.channels = n_ch;
    c->filter_stats.samples = T;
    c->filter_stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return PACH_OK;
}

extern "C" int pach_receiver_filters(pach_ctx *c, const double swl[8], double *out)
{
    if (!c || !swl || !out) return PACH_EINVAL;
    if (!c->has_recv) return fail(c, PACH_ESTATE, "no receivers");
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    const int R = c->R, T = c->sct;
    if (T <= 1) return fail(c, PACH_ESTATE, "histograms too short for filters");
    const auto t_start = std::chrono::steady_clock::now();
    (void)hipSetDevice(c->device);
    const FiltTables f = filt_tables(c, T, swl, (int)c->fs);
    // Receiver_Bank.Scale: 1 / (completed source rays of each receiver's room)
    std::vector<double> scale(R, 0.0);
    const int64_t done = c->rays_done.load();
    for (const RoomTab &t : c->room_tab) {
        const double n = (double)std::min(std::max<int64_t>(done - t.ray_base, 0), t.ray_count);
        for (int r = t.rec_base; r < t.rec_base + t.rec_count && r < R; ++r) scale[r] = n > 0 ? 1.0 / n : 0.0;
    }
    const int n_ch = 7 * R, n_chb = (n_ch + kFiltCh - 1) / kFiltCh;
    auto body = [&]() -> int {
        const size_t env_n = (size_t)kFiltBands * n_chb * T * kFiltCh;
        HIPCHK(c, dgrow(c->f_env, c->f_env_n, env_n));
        HIPCHK(c, hipMemsetAsync(c->f_env, 0, env_n * sizeof(double), c->stream));
        std::vector<double> tab(scale);
        tab.insert(tab.end(), f.p_mod, f.p_mod + 8);
        tab.insert(tab.end(), f.f16.begin(), f.f16.end());
        HIPCHK(c, dgrow(c->f_tab, c->f_tab_n, tab.size()));
        HIPCHK(c, hipMemcpyAsync(c->f_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_receiver_env(c->stream, c->d_hist, c->hist_n, R, T, c->f_tab, c->f_tab + R,
                                      c->f_tab + R + 8, n_chb, c->f_env));
        return run_filters(c, c->f_env, f, (int)c->fs, n_ch, T, out, 4096);
    };
    const int rc = body();
    c->filter_stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return rc;
}

extern "C" int pach_filter_stats_get(pach_ctx *c, pach_filter_stats *out)
{
    if (!c || !out) return PACH_EINVAL;
    *out = c->filter_stats;
    return PACH_OK;
}

extern "C" int pach_shoot(pach_ctx *c, const double *origins, const double *dirs, int64_t n, int precision,
                          int *tri, double *t, int64_t *n_exact)
{
    if (!c || n < 0 || (n > 0 && (!origins || !dirs || !tri || !t))) return PACH_EINVAL;
    if (c->running) return fail(c, PACH_ESTATE, "trace running");
    if (!c->has_scene) return fail(c, PACH_ESTATE, "no scene");
    if (precision != PACH_PREC_FP32 && precision != PACH_PREC_FP64 && precision != PACH_PREC_MIXED)
        return fail(c, PACH_EINVAL, "bad precision");
    if (n > ((int64_t)1 << 31)) return fail(c, PACH_EINVAL, "at most 2^31 rays per call");
    if (n_exact) *n_exact = 0;
    if (n == 0) return PACH_OK;
    (void)hipSetDevice(c->device);
    hipStream_t st = c->stream;
    HIPCHK(c, dgrow(c->q_o, c->q_o_n, 3 * (size_t)n));
    HIPCHK(c, dgrow(c->q_d, c->q_d_n, 3 * (size_t)n));
    HIPCHK(c, dgrow(c->q_t, c->q_t_n, (size_t)n));
    HIPCHK(c, dgrow(c->q_tri, c->q_tri_n, (size_t)n));
    if (!c->q_exact) HIPCHK(c, dalloc(c->q_exact, 1));
    HIPCHK(c, hipMemcpyAsync(c->q_o, origins, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->q_d, dirs, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemsetAsync(c->q_exact, 0, sizeof(unsigned), st));
    TraceConst tc{};
    mix_bounds(c, tc);
    DevScene<float> sg = dev_scene<float>(c);
    if (precision == PACH_PREC_FP32) {
        HIPCHK(c, launch_shoot<float>(st, tc, sg, sg, c->q_o, c->q_d, n, c->q_tri, c->q_t, c->q_exact));
    } else {
        const DevScene<double> s = dev_scene<double>(c);
        if (precision == PACH_PREC_MIXED)
            HIPCHK(c, launch_shoot_mix(st, tc, s, sg, c->q_o, c->q_d, n, c->q_tri, c->q_t, c->q_exact));
        else
            HIPCHK(c, launch_shoot<double>(st, tc, s, sg, c->q_o, c->q_d, n, c->q_tri, c->q_t, c->q_exact));
    }
    unsigned ne = 0;
    HIPCHK(c, hipMemcpyAsync(tri, c->q_tri, n * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(t, c->q_t, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&ne, c->q_exact, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (n_exact) *n_exact = ne;
    return PACH_OK;
}

extern "C" int pach_stats_get(pach_ctx *c, pach_stats *out)
{
    if (!c || !out) return PACH_EINVAL;
    *out = c->stats;
    return PACH_OK;
}

extern "C" int pach_percent_lost(pach_ctx *c, double *out)
{
    if (!c || !out) return PACH_EINVAL;
    *out = c->stats.e_init > 0 ? 100 * c->stats.lost / c->stats.e_init : 0.0;
    return PACH_OK;
}

// ------------------------------------------------------------------ decay curves (Classes_Math.cs)
extern "C" int pach_schroeder_integral(const double *etc, int n, double *out)
{
    // Schroeder_Integral (:282-299): 1 - cumulative / total
    if (!etc || !out || n <= 0) return PACH_EINVAL;
    std::vector<double> b(n);
    double sum = 0;
    for (int i = 0; i < n; ++i) b[i] = sum += etc[i];
    for (int i = 0; i < n; ++i) out[i] = sum == 0 ? 0.0 : 1 - b[i] / sum;
    return PACH_OK;
}

static double decay_fit(const double *sch, int n, double first, double second, int fs)
{
    // Log10Data (:392-407; index 0 stays 0 dB), Schroeder_Limits (:417-440),
    // Least_Square_Fit (:446-478): value = -60 / slope
    std::vector<double> ls(n, 0.0);
    const double lim = std::pow(10, -60.0 / 10);
    for (int i = 1; i < n; ++i) ls[i] = sch[i] >= lim ? 10 * std::log10(sch[i]) : -60.0;
    int l0 = 0, l1 = 0;
    for (int i = 1; i < n; ++i)
        if (ls[i] < first) { l0 = i; break; }
    for (int i = l0; i < n; ++i)
        if (ls[i] < second) { l1 = i; break; }
    const int m = l1 - l0 + 1;
    if (m <= 0) return NAN;
    double ts = 0, ms = 0, tt = 0, tm = 0;
    for (int i = 0; i < m; ++i) {
        const double q = i / (double)fs;
        ts += q; ms += ls[l0 + i]; tt += q * q; tm += q * ls[l0 + i];
    }
    const double slope = ((m * tm) - (ts * ms)) / ((m * tt) - (ts * ts));
    return -60 / slope;
}

extern "C" int pach_t_x(const double *sch, int n, int depth, int fs, double *out)
{
    // T_X (:511-527): fit from -5 dB to -(5 + depth) dB
    if (!sch || !out || n <= 0 || fs <= 0) return PACH_EINVAL;
    *out = decay_fit(sch, n, -5, -5 - std::abs(depth), fs);
    return PACH_OK;
}

extern "C" int pach_early_decay_time(const double *sch, int n, int fs, double *out)
{
    // EarlyDecayTime (:534-551): fit from -1 dB to -10.0001 dB
    if (!sch || !out || n <= 0 || fs <= 0) return PACH_EINVAL;
    *out = decay_fit(sch, n, -1, -10.0001, fs);
    return PACH_OK;
}

extern "C" int pach_clarity(const double *etc, int n, double fs, double seconds, double start, int pressure,
                            double *out)
{
    // Clarity (:559-585): early/late energy with the boundary bin split
    if (!etc || !out || n <= 0 || !(fs > 0)) return PACH_EINVAL;
    if (pressure) seconds += 8192.0 / fs;
    const double bw = 1 / fs;
    const int si = (int)std::floor(start * fs), ei = (int)std::floor((seconds + start) * fs);
    if (si < 0 || ei >= n) return PACH_EINVAL;
    double early = 0, late = 0;
    for (int q = si; q < ei; ++q) early += etc[q];
    const double split = (seconds + start - (ei * bw)) / bw;
    early += split * etc[ei];
    for (int q = ei + 1; q < n; ++q) late += etc[q];
    late += (1 - split) * etc[ei];
    *out = 10 * std::log10(early / late);
    return PACH_OK;
}
