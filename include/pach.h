/*
 * pach.h -- C ABI of the MI355X stochastic split ray tracer (libpach.so).
 *
 * The library re-implements Pachyderm's Complete_Trace ray loop
 * (SplitRayTracer, "Classes_Complete_Trace - TaskAll.cs":33-536) on one
 * gfx950 GPU.  A host in any language (the C# GpuSplitRayTracer shim, Python
 * ctypes, C++) flattens Scene / Source / Receiver_Bank into the plain arrays
 * below and calls these entry points.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative PACH_E* code on
 *     failure; pach_last_error(ctx) then describes the failure.  No C++
 *     exception or longjmp crosses this boundary.
 *   - The caller owns every buffer.  Inputs are copied during the call;
 *     outputs are written into caller-allocated arrays.
 *   - A context is used from one host thread at a time.  Several contexts
 *     may coexist; each owns its own HIP stream.
 *   - Histogram layout is [R][8][sample_ct] doubles, band-major per
 *     receiver, the layout of Spherical_Receiver.Recs (Classes_Receivers.cs:1694-1760).
 */
#ifndef PACH_H
#define PACH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PACH_ABI_VERSION 1

/* error codes */
#define PACH_OK 0
#define PACH_EINVAL (-1)   /* bad argument */
#define PACH_EHIP (-2)     /* HIP runtime failure */
#define PACH_ESTATE (-3)   /* call out of order (no scene, trace running, ...) */
#define PACH_ENOMEM (-4)   /* device allocation failed */
#define PACH_ECANCEL (-5)  /* trace was cancelled (results hold completed rays) */
#define PACH_ERANGE (-6)   /* output buffer too small (the count needed is returned) */

/* trace modes */
#define PACH_MODE_STOCHASTIC 0  /* Lambert split scheme of Lambert_Scattering */
#define PACH_MODE_SPECULAR 1    /* deterministic: scattering/transmission off */

/* geometry precision of the traversal and reflection arithmetic */
#define PACH_PREC_FP32 0
#define PACH_PREC_FP64 1
/* FP64 rays, hits and shading (the FP64 results) with an FP32 screen for the
 * voxel walk: every hit is decided by the FP64 test; candidates the FP32 screen
 * cannot order with certainty are walked again in FP64 */
#define PACH_PREC_MIXED 2

/* source direction modes */
#define PACH_DIR_GEODESIC 0   /* GeodesicSource (Classes_Sources.cs:272-404) */
#define PACH_DIR_RANDOM 1     /* RandomSource (Classes_Sources.cs:238-266) */
#define PACH_DIR_SUPPLIED 2   /* caller's direction list, source ray i uses dirs[i % n] */

/* histogram planes (Directional_Histogram.Add, Classes_Receivers.cs:1813-1825,
 * and Reflection_Analysis_Data, :857-875) */
#define PACH_PLANE_ENERGY 0
#define PACH_PLANE_PRESSURE 1
#define PACH_PLANE_POS_X 2
#define PACH_PLANE_POS_Y 3
#define PACH_PLANE_POS_Z 4
#define PACH_PLANE_NEG_X 5
#define PACH_PLANE_NEG_Y 6
#define PACH_PLANE_NEG_Z 7
#define PACH_PLANE_SPEC 8        /* specular, never scattered */
#define PACH_PLANE_DIFFUSE 9     /* currently scattered */
#define PACH_PLANE_SPEC_SCAT 10  /* specular after an earlier scatter */
#define PACH_NPLANES 11

typedef struct pach_ctx pach_ctx;

typedef struct pach_params {
    double cutoff_m;      /* SplitRayTracer ctor `cutoffTime`: a path length in m; COTime = cutoff_m / c (TaskAll:83) */
    int oct_lo, oct_hi;   /* octaveRange[0], octaveRange[1] (TaskAll:81-82) */
    int is_order;         /* IS_Order: broadband deposits only for order > is_order (TaskAll:366, :378) */
    int mode;             /* PACH_MODE_* */
    int precision;        /* PACH_PREC_* */
    int async;            /* 0: return when done; 1: return at once, use pach_poll/pach_wait */
    int64_t ray_count;    /* Raycount: > 0 source rays; < 1 convergence mode (PACH_CONV_*) */
    uint64_t seed;        /* Philox key of every ray stream */
    int p1_max_bounces;   /* broadband-phase bounce cap (0 = 128) */
    int max_bounces;      /* per octave ray bounce guard, phases 2+3 (0 = 100000) */
    int64_t batch_rays;   /* source rays per phase-1 launch (0 = min(ray count, 131072)) */
} pach_params;

typedef struct pach_progress {
    int64_t rays_done;    /* _currentRay (TaskAll:40): completed source rays */
    int64_t rays_total;
    int64_t octave_rays;  /* _rayTotal analogue (sub-rays traced) */
    int running;          /* 1 = ThreadState Running, 0 = Stopped */
    int status;           /* last trace status (0 or PACH_E*) */
} pach_progress;

typedef struct pach_stats {
    int64_t bounces_broad;     /* closest-hit queries by broadband rays (incl. misses) */
    int64_t bounces_late;      /* ... by octave rays in the split phase (TaskAll:434-473) */
    int64_t bounces_verylate;  /* ... in End_Trace (TaskAll:498-534) */
    int64_t octave_rays;       /* octave rays created (split + scatter + transmission) */
    int64_t spawned;           /* of those, created by Scatter_Early/Late or transmission */
    int64_t misses;
    int64_t deposits;          /* receiver histogram additions */
    int64_t p1_cap_hits;       /* broadband rays stopped by p1_max_bounces */
    int64_t guard_hits;        /* octave rays stopped by max_bounces */
    int64_t spills;            /* rays moved to the overflow queue */
    int64_t resumes;           /* broadband rays resumed after an output-buffer overflow */
    double lost;               /* energy of rays leaving the model (TaskAll:372, :440, :530) */
    double e_init;             /* sum of initial ray energies (TaskAll:350) */
    int h_oct;                 /* band steering the broadband phase (TaskAll:88-109) */
    int64_t rays_done;
    double ms_phase1;          /* summed HIP-event durations of the broadband kernel */
    double ms_octave;          /* ... of the persistent octave-ray kernel */
    int64_t launches_phase1;
    int64_t launches_octave;
    double ms_trace;           /* wall time of the last trace (host clock) */
    int64_t loop_steps;        /* persistent-loop steps of the octave-ray kernel (per wave) */
    int64_t voxels;            /* voxels visited by octave-ray walks */
    int64_t tri_tests;         /* triangle tests of octave-ray walks */
    int64_t shade_passes;      /* shading passes (per wave) */
    int64_t lanes_shaded;      /* octave-ray bounces shaded (sum over the passes' lanes) */
    /* convergence mode (ray_count < 1) */
    int64_t revolutions;          /* revolutions of 80 source rays traced */
    int conv_receiver_visible;    /* farthest receiver the source sees (-1: none), checked */
    int conv_receiver_occluded;   /* farthest receiver it does not see (-1: none), checked */
    int converged;                /* 1: every check passed */
    /* mixed precision */
    int64_t exact_rewalks;        /* bounces the FP32 screen left undecided, redone by the FP64 walk */
    int64_t exact_walks;          /* grazing departures, walked in FP64 from the start */
    double ms_exact;              /* summed HIP-event durations of the FP64-walk kernel */
} pach_stats;

/* Convergence checks of a trace with ray_count < 1 (TaskAll:111-176). */
#define PACH_CONV_SPECMIN 0  /* ray_count == 0: SpecMin_Convergence_Check */
#define PACH_CONV_MINIMUM 1  /* ray_count < 0: Minimum_Convergence_Check */

/* Context lifetime.  device: HIP device ordinal. */
int pach_create(pach_ctx **out, int device);
void pach_destroy(pach_ctx *ctx);
const char *pach_last_error(const pach_ctx *ctx);
int pach_abi_version(void);

/* Uniform_Medium (Classes_Medium_Props.cs:283-296, :42-190; Classes_Math.cs:272):
 * octave attenuation in dB/m, speed of sound and density of air. */
int pach_uniform_medium(int air_choice, double temp_c, double hr, double pa_hpa,
                        int edge_correction, double atten_db_per_m[8], double *c,
                        double *rho);

/* Polygon_Scene.Construct + partition (Classes_Scenes.cs:447-630, :709-755).
 * verts [T][3][3]: triangles of the triangulated polygons; surf_of_tri [T]:
 * polygon index (Surf_ID) of each triangle; alpha/scat/trans [S][8] per
 * polygon: Coefficient_A_Broad, Scattering Coefficient, TransmissionValue.
 * grid_{x,y,z}: voxel counts of the uniform grid (<= 0: chosen automatically). */
int pach_scene_upload(pach_ctx *ctx, const double *verts, int T, const int *surf_of_tri,
                      int S, const double *alpha, const double *scat, const double *trans,
                      int grid_x, int grid_y, int grid_z);

/* The uniform voxel grid pach_scene_upload builds, computed on the host (no GPU
 * needed; for inspection and tests).  dims[3]: requested voxel counts (any
 * <= 0: automatic), overwritten with the counts used; origin[3], cell[3]: grid
 * corner and voxel size; ranges [ncell][2]: (first index, count) into tris;
 * dist [ncell]: L-inf distance in voxels to the nearest occupied voxel (0 =
 * occupied).  *n_list: capacity of tris on entry, list length on return; tris
 * is filled only when large enough, so call once with tris = NULL to size it. */
int pach_grid_build(const double *verts, int T, int dims[3], double origin[3], double cell[3],
                    uint32_t *ranges, uint32_t *tris, int64_t *n_list, uint8_t *dist);

/* The walk's per-face skip masks of the same grid (inspection and tests; no
 * counterpart in the reference, whose per-ray mailbox,
 * Classes_VoxelGrid_PolyRefractive.cs:162-164, they replace): skip
 * [ncell][6] (faces -x, +x, -y, +y, -z, +z), bit i set when entry i < 32 of
 * the voxel's list is also listed in the neighbour across that face. */
int pach_grid_face_skip(const double *verts, int T, int dims[3], uint32_t *skip);

/* The walk's empty cubes over an occupancy field (inspection and tests; no
 * counterpart in the reference): dist [ncell] (0 = occupied, as
 * pach_grid_build's), oct [ncell][8]: for each empty voxel and ray octant o
 * (bit 0: d_x < 0, bit 1: d_y < 0, bit 2: d_z < 0) the edge of the largest
 * empty cube cornered at the voxel and extending along o (voxels outside the
 * grid count as empty; at most 255), 0 for occupied voxels. */
int pach_grid_octant_cubes(const int dims[3], const uint8_t *dist, uint8_t *oct);

/* The walk's empty boxes over the same occupancy field (inspection and tests;
 * no counterpart in the reference): ext [ncell][3][8]: for each empty voxel
 * and ray octant o, the box a ray in octant o crosses in one move, cornered at
 * the voxel and extending in o's directions (voxels outside the grid count as
 * empty; lengths at most 255).  With K the octant cube's edge and E_a the
 * length of the longest empty box K long on the two axes other than a: first
 * axis A = the a with the largest E_a (the lowest on a tie), second axis B =
 * the longer of the other two (the one after A on a tie), E2 = the longest
 * empty length on B with A at E_A and the third axis at K.  ext[v][0][o] =
 * E_A, ext[v][1][o] = E2, ext[v][2][o] = A | B << 2; 0 for occupied voxels. */
int pach_grid_octant_boxes(const int dims[3], const uint8_t *dist, uint8_t *ext);

/* Source origin, per-band power in W (Source.SoundPower, Classes_Sources.cs:96-104)
 * and direction mode.  dirs [n_dirs][3] only for PACH_DIR_SUPPLIED. */
int pach_source_set(pach_ctx *ctx, const double origin[3], const double power_w[8],
                    int dir_mode, const double *dirs, int64_t n_dirs);

/* Receiver_Bank of spherical receivers (Classes_Receivers.cs:139-164, :748-767).
 * radius: the reference hard-codes 1 m (:751); SizeMod = 1/(pi r^3) reduces to
 * the reference's 1/pi at r = 1.  atten: medium attenuation in dB/m; c, rho_c
 * as Uniform_Medium; sample_ct = floor(COTime_ms * fs / 1000) (:144). */
int pach_receivers_set(pach_ctx *ctx, const double *centers, int R, double radius,
                       const double atten_db_per_m[8], double c, double rho_c,
                       int sample_rate, int sample_ct);

/* Batched rooms (no counterpart in the reference: several independent rooms
 * traced in one launch).  The uploaded scene holds the rooms side by side
 * (disjoint in space), their polygons room by room (surf_count[i] polygons for
 * room i) and the receiver list room by room (rec_count[i] receivers).  Room i
 * has its own source (src[i][3], power[i][8] in W) and traces rays[i] source
 * rays; pach_trace's ray_count must then be 0 or their sum.  Histograms keep
 * the receiver order; scaled copies divide each room by its own ray count.
 * n = 0 returns to one room (pach_source_set's source, all receivers). */
int pach_rooms_set(pach_ctx *ctx, int n, const double *src, const double *power, const int64_t *rays,
                   const int *surf_count, const int *rec_count);

/* Mapping receivers (PachMapReceiver, Pach_Map_Receiver.cs:34-1690): M
 * spheres of diameter `increment` at centers [M][3] that replace the
 * spherical receivers in later traces (one room; RecMain is the map), with
 * their own voxel grid (Partition :99-303) and voxel walks per ray segment
 * (CheckRay :1493-1631, CheckBroadbandRay :1359-1491, as written there).
 * Each receiver records at 1000 Hz, sample_ct samples, times relative to its
 * direct arrival, deposits I 10^(-0.1 a t) / (pi r^2) at the chord midpoint;
 * directional: Directional_Histogram (pressure sqrt(E rho c), direction
 * planes) else Histogram (pressure sqrt(E)); both drop deposits outside
 * [0, sample_ct).  time1pt = 1 (the ctor's Time1Pt, :65-93, which sets
 * SampleCT to 1; sample_ct must be 1): Histogram_1Pt (Classes_Receivers.cs:
 * 1369-1402), every deposit's energy added to sample 0 of plane 0 whatever
 * its time, the other planes left zero.  pach_receivers_set still supplies
 * the medium. */
int pach_map_receivers_set(pach_ctx *ctx, const double *centers, int M, double increment, int directional,
                           int sample_ct, int time1pt);
int pach_map_receivers_clear(pach_ctx *ctx);
/* plane 0-7 (energy, pressure, +x, +y, +z, -x, -y, -z): out [M][8][sample_ct],
 * scaled by the source rays traced (pressure by their square root). */
int pach_map_histogram(pach_ctx *ctx, int plane, double *out, int scaled);

/* Launch tuning of later traces (no counterpart in the reference; results do
 * not depend on it).  Keys: "dist_lds" (0/1): stage the voxel distance field
 * in LDS when it fits; "receiver_masks" (0/1): test only receivers whose
 * voxels the walk crossed (up to 64 receivers); "walk_steps" (1-64): voxels a
 * walk advances per step of the persistent loop (with "walk_flat" = 1: voxel
 * steps plus triangle tests, 1-256); "shade_min" (1-64): lanes that must wait
 * before a wave shades; "walk_flat" (0/1/2): nested walk loop, flattened walk
 * loop, or flattened with one kind of unit (tests or voxel moves) per wave
 * iteration; "refill_min"
 * (1-64): idle lanes a wave gathers before refilling them (the refill
 * threshold); "dump_min" (>= 1): loop steps a drained wave runs before handing
 * its leftover rays to the next launch; "block" (256/512): workgroup size of
 * the octave kernel; "occupancy" (-1/0/3/4/5): 4 and 5 cap registers at 128 and
 * 102 VGPRs so four or five waves share each SIMD, 0 leaves the compiler's
 * allocation, -1 (default) takes 4 for FP32 and 3 for mixed and FP64, 3 caps
 * at 168 VGPRs; "shade_queue" (0/1): a lane whose walk ended queues
 * its hit and walks on, and the wave shades 64 queued hits at a time;
 * "exact_queue" (>= 1, mixed precision): rays the queue of bounces left to the
 * FP64 walk holds per launch beyond a reserve of 128 per wave (a wave drains
 * once its push lands past that; default 2^25). */
int pach_tune(pach_ctx *ctx, const char *key, int64_t value);

/* Exact (order-independent) histogram accumulation for later traces (on = 1)
 * or FP64 atomics (0, the default).  Replaces the reference's unsynchronised
 * FP64 adds into the shared histograms (Directional_Histogram.Add,
 * Classes_Receivers.cs:1159-1166, from every task of TaskAll:217-282), whose
 * sums depend on the thread schedule.  With on = 1 each deposit is scaled by a
 * power of two chosen from the source power and receiver radius (not the ray
 * count, so a deposit is the same integer in every trace of a scene), rounded
 * to an integer of up to 123 bits and added as 32-bit limbs into int64 sums,
 * which are converted to FP64 when the trace ends (and after each convergence
 * revolution): the histograms are then bit-for-bit the same whatever the launch
 * shape, batch size or schedule, at 4x the histogram memory (4 x 11 x R x 8 x
 * SampleCT int64) and up to 4 integer atomics per plane per deposit.  The unit
 * is 2^-76 of the largest possible deposit u = 2 P / (pi r^2), and the range
 * holds 16 deposits of u per bin from each of 2^40 source rays; the spherical
 * receivers only (a trace with mapping receivers returns PACH_EINVAL).  A
 * deposit beyond the range makes the trace return PACH_ERANGE. */
int pach_exact_histograms(pach_ctx *ctx, int on);

/* Polygon_Scene.shoot (Classes_Scenes.cs:757-819; Hare Voxel_Grid.Shoot, whose
 * contract is restated at Classes_VoxelGrid_PolyRefractive.cs:70-316) for n
 * rays on the uploaded scene: origins, dirs [n][3] (unit directions); out
 * tri [n] (triangle index, -1: the ray leaves the model; surf_of_tri gives
 * the polygon) and t [n] (hit distance, 0 on a miss).  PACH_PREC_FP64 and
 * PACH_PREC_MIXED return the FP64 closest hit (Moller-Trumbore, t > 1e-10,
 * the lower triangle index on equal t); PACH_PREC_FP32 the FP32 walk of the
 * FP32 tracer (no triangle excluded).  n_exact (may be NULL): mixed walks
 * repeated with every candidate decided in FP64. */
int pach_shoot(pach_ctx *ctx, const double *origins, const double *dirs, int64_t n, int precision, int *tri,
               double *t, int64_t *n_exact);

/* Deterministic direct sound of a point source (Direct_Sound,
 * Classes_Direct.cs:654-840), for the energy-time curve that
 * IR_Construction.ETCurve (Classes_Math.cs:1417-1474) assembles.  Host
 * arithmetic, no context needed.  verts [T][3][3], surf_of_tri [T], trans
 * [S][8] (may be NULL); src, power [8] (W), centers [R][3], atten [8] (dB/m),
 * c (m/s).  Out per receiver: energy [R][8] = power 10^(-0.1 a d) transmod /
 * (4 pi d^2) (the reference computes it whether or not the path is blocked),
 * time_s [R] = d / c, audible [R] (Check_Validity: no surface nearer than the
 * receiver, continuing through transmissive ones, whose coefficients make
 * transmod). */
int pach_direct_sound(const double *verts, int T, const int *surf_of_tri, const double *trans, int S,
                      const double src[3], const double power[8], const double *centers, int R,
                      const double atten[8], double c, double *energy, double *time_s, int *audible);

/* Image sources: one valid specular path of a reflection sequence to one
 * receiver (Specular_Path, Classes_ImageSource.cs:1891-1982). */
#define PACH_IS_MAX_ORDER 8
typedef struct pach_is_path {
    int32_t receiver;
    int32_t order;                         /* reflections */
    int32_t seq[PACH_IS_MAX_ORDER];        /* object reflected from, source side first */
    int32_t poly[PACH_IS_MAX_ORDER];       /* polygon hit at each reflection */
    double pts[PACH_IS_MAX_ORDER + 2][3];  /* source, reflection points, receiver */
    double length;                         /* m */
    double time_s;                         /* length / c */
    double energy[8];                      /* PathEnergy per octave band */
} pach_is_path;

/* ImageSourceData for planar objects without edge diffraction
 * (Classes_ImageSource.cs: TraverseOrders :266-348, ProcessImages :350-476,
 * OcclusionIntersect :1350-1392, FinalOcclusion :1245-1281, Power_Recursion
 * :535-803, B_Function_NoInterp / H_Function :859-1030, Specular_Path
 * :1914-1982), on the uploaded scene, source and receiver centres.  Every
 * sequence of 1..max_order objects (repeats allowed) is validated for every
 * receiver on the GPU in FP64.  object_of_surface [S] groups polygons into the
 * reference's objects (Polygon_Scene.Construct's Model[i]); an object reflects
 * in the plane of its lowest polygon.  NULL: each polygon is its own object
 * (n_objects = S).  The valid paths are sorted by receiver, then in the
 * reference's depth-first sequence order; min(total, cap) are written to out
 * and *n_paths = total.  Returns PACH_ERANGE when total > cap. */
int pach_image_sources(pach_ctx *ctx, const int *object_of_surface, int n_objects, int max_order,
                       pach_is_path *out, int64_t cap, int64_t *n_paths);

/* Work and timing of the last pach_image_sources call (no counterpart in the
 * reference). */
typedef struct pach_is_stats {
    int64_t checks;     /* (sequence, receiver) pairs validated */
    int64_t paths;      /* valid paths found */
    int64_t launches;   /* kernel launches */
    double ms_kernel;   /* summed HIP-event durations of the validation kernel */
    double ms_total;    /* wall time of the call (host clock) */
} pach_is_stats;
int pach_image_sources_stats(pach_ctx *ctx, pach_is_stats *out);

/* Auralisation filters.  Pach_SP.ETCToFilter (Pach_SP.cs:1584-1676) of n_ch
 * channels on the GPU: prms [n_ch][8][T] octave-band pressure envelopes
 * (62.5 Hz .. 8 kHz) at sample_frequency, swl [8] source power levels (dB).
 * Each band's |envelope| x 10^((120 - SWL) / 20) is convolved with the band's
 * 16,384-tap band-limited carrier (GenerateBandLimitedSincPulse at
 * sample_frequency) and the bands are summed, plus a 16 kHz band made from the
 * 8 kHz envelope with ISO 9613-1's extra air attenuation over 343 m/s x t.
 * out [n_ch][T + 16383]. */
int pach_etc_to_filter(pach_ctx *ctx, const double *prms, int n_ch, int T, const double swl[8],
                       int sample_frequency, double *out);

/* Allocation-free pach_etc_to_filter.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  d_prms [n_ch][8][T]
 * and d_out [n_ch][T + 16383] are DEVICE pointers the caller owns; the work
 * buffers (carriers, envelopes, the nine band planes) live in one fixed
 * device array of PACH_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, so the call allocates nothing.  Every call on a GPU shares that
 * array: calls that overlap on one GPU race.  Returns PACH_ERANGE when
 * 8 (9 x 16384 + 8 + T + 72 ceil(n_ch / 8) T + 9 n_ch (T + 16383)) bytes
 * (each of the four parts rounded up to 256 bytes) exceed the array;
 * otherwise the same result as pach_etc_to_filter. */
#define PACH_NOALLOC_SCRATCH_BYTES (64ull << 20)
int pach_etc_to_filter_noalloc(pach_ctx *ctx, const double *d_prms, int n_ch, int T, const double swl[8],
                               int sample_frequency, double *d_out);

/* Directional_Histogram.Create_Filter(SWL, Rho_C) (Classes_Receivers.cs:
 * 1911-1936) for every receiver of the last trace, from its scaled histograms:
 * 7 filters per receiver -- the pressure envelope, then for x, y, z the
 * envelopes P neg / E and P pos / E (0 where E = 0) -- each ETCToFilter at the
 * receivers' sample rate shortened by 4096 samples.
 * out [R][7][SampleCT + 16383 - 4096]. */
int pach_receiver_filters(pach_ctx *ctx, const double swl[8], double *out);

/* Work and timing of the last filter call (no counterpart in the reference). */
typedef struct pach_filter_stats {
    int64_t channels;   /* channels convolved */
    int64_t samples;    /* envelope samples per band and channel */
    int64_t launches;
    double ms_kernel;   /* HIP-event duration of the convolution kernel */
    double ms_total;    /* wall time of the call (host clock) */
} pach_filter_stats;
int pach_filter_stats_get(pach_ctx *ctx, pach_filter_stats *out);

/* One revolution of a convergence check on host arrays, the function the
 * convergence mode applies to each checked receiver (TaskAll:870-956 for
 * PACH_CONV_SPECMIN, :1063-1112 for PACH_CONV_MINIMUM).  energy, spec, diff,
 * rogue: the receiver's running (unscaled) energy histogram in the steering
 * band and its specular / diffuse / specular-scattered analysis arrays (only
 * energy for PACH_CONV_MINIMUM), n bins at fs; direct_s: source-receiver
 * distance / c.  state: pach_convergence_state_size() doubles, zero before the
 * first revolution. */
int64_t pach_convergence_state_size(int kind, int n, int fs);
int pach_convergence_step(int kind, const double *energy, const double *spec, const double *diff,
                          const double *rogue, int n, int fs, double direct_s, double *state, int *converged);

/* Autotuner result (no counterpart in the reference). */
typedef struct pach_tuning {
    int grid[3];           /* voxel grid dimensions chosen (now uploaded) */
    int block, occupancy;  /* octave-kernel workgroup size and register cap */
    int shade_queue;       /* 1: finished walks are shaded 64 at a time from a per-wave queue */
    int walk_flat;         /* walk loop form (pach_tune "walk_flat") */
    int walk_steps, shade_min, refill_min, dump_min;
    int trials;            /* sample traces run */
    double rate_start;     /* ray-bounces/s of the sample under the settings on entry */
    double rate_best;      /* ... under the chosen settings */
} pach_tuning;

/* Picks voxel resolution, shade queue, workgroup size, register cap, walk budget, shading
 * and refill thresholds for this scene, receivers and trace parameters by
 * timing sample traces (coordinate descent over each knob's candidates,
 * starting from the current settings), then applies the best: the grid is
 * rebuilt at the chosen resolution and the knobs are set as by pach_tune.
 * sample_rays: source rays per trial (<= 0: enough for ~2e9 bounces per
 * trial, from the bounces per source ray of a first 8192-ray trial, at most
 * the ray count); with rooms set, every room's ray count is scaled down to the
 * sample in proportion.
 * Overwrites the histograms and statistics (tune before the real trace).
 * No counterpart in the reference (its Voxel_Grid resolution is a user input,
 * Classes_Scenes.cs:709-755). */
int pach_autotune(pach_ctx *ctx, const pach_params *p, int64_t sample_rays, pach_tuning *out);

/* SplitRayTracer(...).Begin() (TaskAll:68-240).  Clears the histograms first.
 * With ray_count < 1 (and no rooms set) the trace runs in convergence mode
 * (TaskAll:111-176, :288-329): revolutions of 80 source rays, each followed by
 * the convergence checks of the farthest receiver the source sees and the
 * farthest it does not (ray_count < 0: Minimum_Convergence_Check, 0:
 * SpecMin_Convergence_Check, both on the steering band h_oct), until every
 * check passes (at most 2^20 revolutions); rays_done then counts the
 * revolutions' rays. */
int pach_trace(pach_ctx *ctx, const pach_params *params);
/* ThreadState / ProgressMsg analogue (TaskAll:246-252, :742-763). */
int pach_poll(pach_ctx *ctx, pach_progress *out);
/* Block until an async trace ends; returns its status. */
int pach_wait(pach_ctx *ctx);
/* Cooperative cancellation (I_Conv_Progress.CancellationToken, TaskAll:276-280):
 * the trace stops after the batch in flight; results cover completed rays. */
int pach_cancel(pach_ctx *ctx);

/* Copy one histogram plane [R][8][sample_ct].  scaled != 0 applies
 * Combine_ThreadLocal_Results' Receiver_Bank.Scale(rays_done)
 * (Classes_Receivers.cs:1088-1102, :1774-1786): pressure / sqrt(N), other
 * planes / N.  scaled == 0 returns raw sums. */
int pach_histogram(pach_ctx *ctx, int plane, double *out, int scaled);
int pach_stats_get(pach_ctx *ctx, pach_stats *out);
/* PercentLost (TaskAll:773-782) */
int pach_percent_lost(pach_ctx *ctx, double *out);

/* Decay-curve integration, FP64 (Classes_Math.cs:282-585). */
int pach_schroeder_integral(const double *etc, int n, double *out);
int pach_t_x(const double *schroeder, int n, int decay_depth, int sample_rate, double *out);
int pach_early_decay_time(const double *schroeder, int n, int sample_rate, double *out);
int pach_clarity(const double *etc, int n, double sample_rate, double seconds,
                 double start_time, int pressure, double *out);

#ifdef __cplusplus
}
#endif
#endif
