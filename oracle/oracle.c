/*
 * oracle.c -- CPU FP64 restatement of Pachyderm's stochastic split ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP tracer, loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Never shipped.
 *
 * Every function cites the reference lines it restates (paths relative to
 * Pachyderm_Acoustic_Universal/; TaskAll = "Classes_Complete_Trace - TaskAll.cs").
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   philox          Random123 published known-answer vectors.
 *   attenuation     ISO 9613-1 table value at 1 kHz / 20 C / 50 % and the
 *                   closed form with the C# integer-division exponents.
 *   shoot           grid DDA == brute force on random rays; analytic hits on a box.
 *   check_ray       hand-computed known answers (sphere chord, binning edges).
 *   tracer          specular shoebox vs. the image-source closed form (arrival
 *                   bins and energies), energy budget, sampler KS tests,
 *                   geodesic stratification, Sabine/Eyring diffuse-field trend.
 *   decay math      exponential decays with closed-form T30/EDT/C80.
 *   image sources   shoebox image lattice (path count 4n^2+2 and lengths per
 *                   order), mirror law at every reflection point, spherical
 *                   spreading at zero losses, per-wall material products,
 *                   segment-box occlusion and transmissive crossings
 *                   (tests/test_image_source_oracle.py).
 * Parity unpinned (no executable reference exists anywhere offline):
 *   Hare GetRandomPoint's exact (u,v) mapping, Hare Ray.Reverse (taken as
 *   d -> -d), Hare tie-breaking at shared edges, Hare's DistToPlane /
 *   Normalize rounding and its X_Event on a miss (taken as t = 0), and .NET System.Random
 *   streams (replaced by per-ray Philox streams, matched only in distribution).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                        */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* splitmix64 finaliser (Steele, Lea, Flood 2014) used to derive child-ray ids */
uint64_t or_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t or_child_uid(uint64_t parent_uid, uint64_t tag)
{
    return or_mix64(parent_uid + 0x9E3779B97F4A7C15ull * (tag + 1ull));
}

/* Tags of spawned rays: each is unique per parent, so every ray of a
 * source-ray tree owns a distinct Philox stream whatever the processing order. */
#define TAG_EARLY(k, o) ((1ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define TAG_EARLY_TR(k, o) ((2ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define TAG_EARLY_TD(k, o) ((3ull << 40) | ((uint64_t)(k) << 8) | (uint64_t)(o))
#define TAG_SPLIT(o) ((4ull << 40) | (uint64_t)(o))
#define TAG_LATE(k) ((5ull << 40) | (uint64_t)(k))

typedef struct {
    uint32_t key[2];
    uint32_t src;
} rng_ctx;

static void rng_block(const rng_ctx *r, uint64_t uid, uint32_t block, uint32_t w[4])
{
    uint32_t ctr[4] = {block, (uint32_t)uid, (uint32_t)(uid >> 32), r->src};
    or_philox4x32_10(ctr, r->key, w);
}

static double u01(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }

/* ------------------------------------------------------------------------ */
/* Medium: Classes_Medium_Props.cs:42-190, :283-296; Classes_Math.cs:272     */
/* ------------------------------------------------------------------------ */
void or_attenuation(int air_choice, double temp_c, double hr, double pa_hpa,
                    int edge_correction, double out[8])
{
    static const double F[8] = {62.5, 125, 250, 500, 1000, 2000, 4000, 8000};
    double Tk = temp_c + 273.15;
    double Pa = pa_hpa / 10.0; /* Uniform_Medium passes Pa/10 (:289) */
    for (int i = 0; i < 8; ++i) out[i] = 0.0;
    if (air_choice == 0) {
        /* C# integer division: -1/2 == 0, -1/3 == 0, -5/2 == -2, 1/2 == 0 (:84-87) */
        double h = hr * pow(10.0, -6.8346 * pow(273.16 / Tk, 1.261) + 4.6151) / (Pa / 101.325);
        for (int o = 0; o < 8; ++o) {
            double frO = (Pa / 101.325) * (24 + 4.04 * pow(10.0, 4) * h * ((0.02 + h) / (0.391 + h)));
            double frN = (Pa / 101.325) * pow(Tk / 293.15, 0.0) *
                         (9 + 280 * h * exp(-4.170 * (pow(Tk / 293.15, 0.0) - 1)));
            double z = 0.1068 * exp(-3352 / Tk) * pow(frN + F[o] * F[o] / frN, -1);
            double y = pow(Tk / 293.15, -2.0) *
                       (0.01275 * exp(-2239.1 / Tk) * pow(frO + F[o] * F[o] / frO, -1) + z);
            out[o] = 8.686 * F[o] * F[o] *
                     ((1.84 * pow(10.0, -11) * pow(Pa / 101.325, -1) * pow(Tk / 293.15, 0.0)) + y);
        }
    } else if (air_choice == 1) {
        for (int q = 0; q < 8; ++q) {
            out[q] = ((85 / hr) * (pow(F[q] / 1000, 2) * 0.0001)) / 8.7;
            out[q] -= 0.04 * (Tk - 273.15 - 20) * out[q];
            out[q] *= 0.50248756218905477;
        }
    } else if (air_choice == 2) {
        static const double A[6][8] = {
            {0.0001, 0.0001, 0.0003, 0.0006, 0.001, 0.0019, 0.0058, 0.0203},
            {0.0001, 0.0001, 0.0003, 0.0006, 0.001, 0.0017, 0.0041, 0.0135},
            {0.0001, 0.0001, 0.0003, 0.0006, 0.0011, 0.0017, 0.0035, 0.0106},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0011, 0.0027, 0.0094, 0.029},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0008, 0.0018, 0.0059, 0.0211},
            {0.0001, 0.0001, 0.0002, 0.0005, 0.0007, 0.0014, 0.0044, 0.0158}};
        int row = -1;
        int base = (Tk - 273.15 > 15) ? 0 : 3;
        if (hr < 50) row = base;
        else if (hr < 70) row = base + 1;
        else if (hr > 70) row = base + 2; /* hr == 70 leaves zeros, as the reference does */
        if (row >= 0)
            for (int q = 0; q < 8; ++q) out[q] = A[row][q];
        if (edge_correction)
            for (int q = 0; q < 8; ++q) out[q] /= sqrt(2.0);
    }
}

/* Calculate_Density(Tk, Pa*100, hr) (:42-49, called with hPa*100 at :288) */
double or_density(double temp_c, double hr, double pa_hpa)
{
    double Tk = temp_c + 273.15, Pa = pa_hpa * 100.0, Rd = 287.05, T = temp_c;
    double Pv = hr * 6.1078 * pow(10.0, (7.5 * T) / (237.3 + T));
    return (Pa / (Rd * Tk)) * (1 - 0.378 * Pv / Pa);
}

double or_sound_speed(double temp_c) { return 331 + 0.6 * temp_c; }

/* ------------------------------------------------------------------------ */
/* Scene: triangles (Hare Topology polygons, triangulated), per-polygon      */
/* materials (Classes_Scenes.cs:484-487), uniform voxel grid.                */
/* ------------------------------------------------------------------------ */
struct or_scene {
    int T, S;
    double *v0, *e1, *e2, *n; /* [T][3] */
    int *surf;
    double *alpha, *scat, *trans, *lo, *hi; /* [S][8] */
    double *area;                           /* [S] */
    int g[3];
    double gmin[3], cs[3];
    int *cell_start; /* [ncell+1] */
    int *cell_tris;
};

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void normalize3(double *a)
{
    double l = sqrt(dot3(a, a));
    if (l > 0) { a[0] /= l; a[1] /= l; a[2] /= l; }
}

or_scene *or_scene_create(const double *verts, int T, const int *surf, int S,
                          const double *alpha, const double *scat, const double *trans,
                          int gx, int gy, int gz)
{
    if (T <= 0 || S <= 0) return NULL;
    or_scene *s = (or_scene *)calloc(1, sizeof(or_scene));
    s->T = T; s->S = S;
    s->v0 = malloc(sizeof(double) * 3 * T);
    s->e1 = malloc(sizeof(double) * 3 * T);
    s->e2 = malloc(sizeof(double) * 3 * T);
    s->n = malloc(sizeof(double) * 3 * T);
    s->surf = malloc(sizeof(int) * T);
    s->alpha = malloc(sizeof(double) * 8 * S);
    s->scat = malloc(sizeof(double) * 8 * S);
    s->trans = malloc(sizeof(double) * 8 * S);
    s->lo = malloc(sizeof(double) * 8 * S);
    s->hi = malloc(sizeof(double) * 8 * S);
    s->area = calloc(S, sizeof(double));
    double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
    for (int t = 0; t < T; ++t) {
        const double *p = verts + 9 * t;
        for (int k = 0; k < 3; ++k) {
            s->v0[3 * t + k] = p[k];
            s->e1[3 * t + k] = p[3 + k] - p[k];
            s->e2[3 * t + k] = p[6 + k] - p[k];
            for (int j = 0; j < 3; ++j) {
                if (p[3 * j + k] < bmin[k]) bmin[k] = p[3 * j + k];
                if (p[3 * j + k] > bmax[k]) bmax[k] = p[3 * j + k];
            }
        }
        double c[3];
        cross3(s->e1 + 3 * t, s->e2 + 3 * t, c);
        double l = sqrt(dot3(c, c));
        s->surf[t] = surf ? surf[t] : t;
        if (s->surf[t] >= 0 && s->surf[t] < S) s->area[s->surf[t]] += 0.5 * l;
        normalize3(c);
        memcpy(s->n + 3 * t, c, sizeof(c));
    }
    /* Lambert_Scattering bounds (Classes_Materials.cs:4466-4484) */
    for (int i = 0; i < 8 * S; ++i) {
        s->alpha[i] = alpha[i];
        s->scat[i] = scat[i];
        s->trans[i] = trans ? trans[i] : 0.0;
        if (scat[i] < 0.25) { s->lo[i] = 0.1; s->hi[i] = 0.4; }
        else {
            double mod = fabs(1 - scat[i]) / 5;
            s->lo[i] = scat[i] - mod;
            s->hi[i] = scat[i] + mod;
        }
    }
    /* uniform grid over the padded bounding box; triangle AABB binning */
    s->g[0] = gx > 0 ? gx : 1; s->g[1] = gy > 0 ? gy : 1; s->g[2] = gz > 0 ? gz : 1;
    for (int k = 0; k < 3; ++k) {
        double ext = bmax[k] - bmin[k];
        double pad = 1e-6 * (ext > 1 ? ext : 1) + 1e-9;
        s->gmin[k] = bmin[k] - pad;
        s->cs[k] = (ext + 2 * pad) / s->g[k];
    }
    int nc = s->g[0] * s->g[1] * s->g[2];
    int *cnt = calloc(nc + 1, sizeof(int));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            s->cell_start = malloc(sizeof(int) * (nc + 1));
            s->cell_start[0] = 0;
            for (int i = 0; i < nc; ++i) s->cell_start[i + 1] = s->cell_start[i] + cnt[i];
            s->cell_tris = malloc(sizeof(int) * (s->cell_start[nc] > 0 ? s->cell_start[nc] : 1));
            memset(cnt, 0, sizeof(int) * (nc + 1));
        }
        for (int t = 0; t < T; ++t) {
            const double *p = verts + 9 * t;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                double mn = p[k], mx = p[k];
                for (int j = 1; j < 3; ++j) {
                    if (p[3 * j + k] < mn) mn = p[3 * j + k];
                    if (p[3 * j + k] > mx) mx = p[3 * j + k];
                }
                double e = 1e-9 * (s->cs[k] > 1 ? s->cs[k] : 1);
                lo[k] = (int)floor((mn - e - s->gmin[k]) / s->cs[k]);
                hi[k] = (int)floor((mx + e - s->gmin[k]) / s->cs[k]);
                if (lo[k] < 0) lo[k] = 0;
                if (hi[k] > s->g[k] - 1) hi[k] = s->g[k] - 1;
            }
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        int c = (z * s->g[1] + y) * s->g[0] + x;
                        if (pass == 0) cnt[c]++;
                        else s->cell_tris[s->cell_start[c] + cnt[c]++] = t;
                    }
        }
    }
    free(cnt);
    return s;
}

void or_scene_destroy(or_scene *s)
{
    if (!s) return;
    free(s->v0); free(s->e1); free(s->e2); free(s->n); free(s->surf);
    free(s->alpha); free(s->scat); free(s->trans); free(s->lo); free(s->hi);
    free(s->area); free(s->cell_start); free(s->cell_tris);
    free(s);
}

double or_surface_area(const or_scene *s, int surf) { return s->area[surf]; }

void or_tri_normal(const or_scene *s, int tri, double n[3]) { memcpy(n, s->n + 3 * tri, sizeof(double) * 3); }

/* Moller-Trumbore, inclusive edges; accepts t > 1e-10 as Hare's Voxel_Grid does
 * (Classes_VoxelGrid_PolyRefractive.cs:166). */
static inline int tri_hit(const or_scene *s, int t, const double *o, const double *d, double *tout)
{
    const double *e1 = s->e1 + 3 * t, *e2 = s->e2 + 3 * t, *v0 = s->v0 + 3 * t;
    double p[3], q[3], sv[3];
    cross3(d, e2, p);
    double det = dot3(e1, p);
    if (det == 0.0) return 0;
    double inv = 1.0 / det;
    sv[0] = o[0] - v0[0]; sv[1] = o[1] - v0[1]; sv[2] = o[2] - v0[2];
    double u = dot3(sv, p) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    cross3(sv, e1, q);
    double v = dot3(d, q) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    double tt = dot3(e2, q) * inv;
    if (!(tt > 1e-10)) return 0;
    *tout = tt;
    return 1;
}

int or_shoot_brute(const or_scene *s, const double o[3], const double d[3], double *t_out)
{
    int best = -1;
    double bt = INFINITY, t;
    for (int i = 0; i < s->T; ++i)
        if (tri_hit(s, i, o, d, &t) && (t < bt || (t == bt && i < best))) { bt = t; best = i; }
    if (best >= 0) *t_out = bt;
    return best;
}

/* Amanatides-Woo DDA (Classes_VoxelGrid_PolyRefractive.cs:84-130, :199-313):
 * a hit ends the walk once its t lies within the current voxel's exit. */
void or_shoot_brute_n(const or_scene *s, const double *o, const double *d, int64_t n, int *tri, double *t)
{
    for (int64_t i = 0; i < n; ++i) {
        double tt = 0;
        tri[i] = or_shoot_brute(s, o + 3 * i, d + 3 * i, &tt);
        t[i] = tri[i] >= 0 ? tt : 0.0;
    }
}

int or_shoot_grid(const or_scene *s, const double o[3], const double d[3], double *t_out)
{
    int c[3], step[3];
    double tmax[3], tdel[3];
    for (int k = 0; k < 3; ++k) {
        int ci = (int)floor((o[k] - s->gmin[k]) / s->cs[k]);
        if (ci < 0) ci = 0;
        if (ci > s->g[k] - 1) ci = s->g[k] - 1;
        c[k] = ci;
        if (d[k] > 0) {
            step[k] = 1;
            tmax[k] = (s->gmin[k] + (ci + 1) * s->cs[k] - o[k]) / d[k];
            tdel[k] = s->cs[k] / d[k];
        } else if (d[k] < 0) {
            step[k] = -1;
            tmax[k] = (s->gmin[k] + ci * s->cs[k] - o[k]) / d[k];
            tdel[k] = -s->cs[k] / d[k];
        } else {
            step[k] = 0;
            tmax[k] = INFINITY;
            tdel[k] = INFINITY;
        }
    }
    int best = -1;
    double bt = INFINITY, t;
    for (;;) {
        int cell = (c[2] * s->g[1] + c[1]) * s->g[0] + c[0];
        for (int j = s->cell_start[cell]; j < s->cell_start[cell + 1]; ++j) {
            int i = s->cell_tris[j];
            if (tri_hit(s, i, o, d, &t) && (t < bt || (t == bt && i < best))) { bt = t; best = i; }
        }
        int ax = (tmax[0] < tmax[1]) ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
        double texit = tmax[ax];
        if (best >= 0 && bt <= texit) break;
        c[ax] += step[ax];
        if (step[ax] == 0 || c[ax] < 0 || c[ax] >= s->g[ax]) break;
        tmax[ax] += tdel[ax];
    }
    if (best >= 0) *t_out = bt;
    return best;
}

static void sincos_turn(double u, double *s, double *c);

/* ------------------------------------------------------------------------ */
/* Sources (Classes_Sources.cs)                                              */
/* ------------------------------------------------------------------------ */
static void icos_tri(const double *a, const double *b, const double *c, int ord, int max,
                     double out[80][3][3], int *fnum)
{
    if (ord < max) {
        double p3[3], p4[3], p5[3];
        for (int k = 0; k < 3; ++k) {
            p3[k] = (a[k] + b[k]) / 2;
            p4[k] = (b[k] + c[k]) / 2;
            p5[k] = (c[k] + a[k]) / 2;
        }
        normalize3(p3); normalize3(p4); normalize3(p5);
        icos_tri(a, p3, p5, ord + 1, max, out, fnum);
        icos_tri(b, p4, p3, ord + 1, max, out, fnum);
        icos_tri(c, p5, p4, ord + 1, max, out, fnum);
        icos_tri(p3, p4, p5, ord + 1, max, out, fnum);
    } else {
        memcpy(out[*fnum][0], a, 3 * sizeof(double));
        memcpy(out[*fnum][1], b, 3 * sizeof(double));
        memcpy(out[*fnum][2], c, 3 * sizeof(double));
        (*fnum)++;
    }
}

/* GeodesicSource.GeoSphere(1) (Classes_Sources.cs:290-350) */
void or_geodesic_faces(double out[80][3][3])
{
    double sqr5 = sqrt(5.0), phi = (1.0 + sqr5) * 0.5;
    double ratio = sqrt(10.0 + 2.0 * sqr5) / (4.0 * phi);
    double a = (.25 / ratio) * 0.5, b = (.25 / ratio) / (2.0 * phi);
    double P[12][3] = {{0, b, -a}, {b, a, 0}, {-b, a, 0}, {0, b, a}, {0, -b, a}, {-a, 0, b},
                       {0, -b, -a}, {a, 0, -b}, {a, 0, b}, {-a, 0, -b}, {b, -a, 0}, {-b, -a, 0}};
    for (int i = 0; i < 12; ++i) normalize3(P[i]);
    static const int F[20][3] = {{0, 1, 2}, {3, 2, 1}, {3, 4, 5}, {3, 8, 4}, {0, 6, 7},
                                 {0, 9, 6}, {4, 10, 11}, {6, 11, 10}, {2, 5, 9}, {11, 9, 5},
                                 {1, 7, 8}, {10, 8, 7}, {3, 5, 2}, {3, 1, 8}, {0, 2, 9},
                                 {0, 7, 1}, {6, 9, 11}, {6, 10, 7}, {4, 11, 5}, {4, 8, 10}};
    int fnum = 0;
    for (int f = 0; f < 20; ++f) icos_tri(P[F[f][0]], P[F[f][1]], P[F[f][2]], 0, 1, out, &fnum);
}

/* Directions(): GeodesicSource (:390-403: face = ray % 80, random point on the
 * face, normalised); RandomSource (:256-265: theta, phi ~ U(0, 2 pi));
 * SUPPLIED: caller's list (the deterministic mode's seeded directions). */
void or_source_direction(int dir_mode, uint64_t seed, int64_t src_id,
                         const double *dirs, int64_t n_dirs, double d[3])
{
    rng_ctx r = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, (uint32_t)src_id};
    uint32_t w[4];
    rng_block(&r, 0xFFFFFFFFFFFFFFFFull, 0, w);
    if (dir_mode == OR_DIR_SUPPLIED && dirs && n_dirs > 0) {
        const double *p = dirs + 3 * (src_id % n_dirs);
        d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
        normalize3(d);
    } else if (dir_mode == OR_DIR_RANDOM) {
        /* theta, phi ~ U(0, 2 pi) */
        double st, ct, sp, cp;
        sincos_turn(u01(w[0]), &st, &ct);
        sincos_turn(u01(w[1]), &sp, &cp);
        d[0] = st * cp; d[1] = st * sp; d[2] = ct;
    } else {
        static double faces[80][3][3];
        static int init = 0;
        if (!__atomic_load_n(&init, __ATOMIC_ACQUIRE)) {
            double tmp[80][3][3];
            or_geodesic_faces(tmp);
            memcpy(faces, tmp, sizeof(tmp));
            __atomic_store_n(&init, 1, __ATOMIC_RELEASE);
        }
        const double(*f)[3] = faces[src_id % 80];
        double su = sqrt(u01(w[0])), v = u01(w[1]);
        for (int k = 0; k < 3; ++k) d[k] = (1 - su) * f[0][k] + su * (1 - v) * f[1][k] + su * v * f[2][k];
        normalize3(d);
    }
}

/* ------------------------------------------------------------------------ */
/* Receivers (Classes_Receivers.cs:803-877) and Directional_Histogram.Add    */
/* (:1813-1825); analysis arrays (:857-875).                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const or_setup *su;
    double **planes;
    double inv_c, sizemod;
    or_counters *cnt;
} rx_ctx;

static void hist_add(const rx_ctx *x, int r, int band, double time, double e, const double *d, int ana)
{
    const or_setup *su = x->su;
    int sample = (int)(time * su->sample_rate);
    if (sample >= su->sample_ct || sample < 0) return;
    size_t idx = ((size_t)r * 8 + band) * su->sample_ct + sample;
    double **P = x->planes;
    if (P[0]) P[0][idx] += e;
    if (P[1]) P[1][idx] += sqrt(e * su->rho_c);
    for (int k = 0; k < 3; ++k) {
        double dk = -d[k];
        double v = (double)(float)(dk * e);
        if (dk > 0) { if (P[2 + k]) P[2 + k][idx] += v; }
        else { if (P[5 + k]) P[5 + k][idx] += v; }
    }
    if (P[8 + ana]) P[8 + ana][idx] += e;
    x->cnt->deposits++;
}

/* CheckRay / CheckBroadbandRay.  end: EndPt (hit point, or o + 1e6 d on a miss).
 * nb == 1: octave ray (E[0] in `band`); nb == 8: broadband over [lo, hi]. */
static void check_rays(const rx_ctx *x, const double *o, const double *d, const double *end,
                       const double *E, int band, int lo, int hi, double t_sum, int ana)
{
    const or_setup *su = x->su;
    double r2 = su->radius * su->radius;
    double dx = end[0] - o[0], dy = end[1] - o[1], dz = end[2] - o[2];
    double sqd = dx * dx + dy * dy + dz * dz;
    for (int r = 0; r < su->R; ++r) {
        const double *C = su->centers + 3 * r;
        double m[3] = {o[0] - C[0], o[1] - C[1], o[2] - C[2]};
        double b = dot3(m, d);
        double c = dot3(m, m) - r2;
        if (c > 0 && b > 0) continue;
        double disc = b * b - c;
        if (disc < 0) continue;
        double t1 = -b - sqrt(disc), t2 = -b + sqrt(disc);
        double t = (t1 + t2) * .5;
        if (t > 0 && t * t < sqd) {
            double ts = (t2 - t1) * 0.5;
            double rt = t * x->inv_c + t_sum;
            double dist = rt * su->c;
            if (band >= 0) {
                double e = E[0] * pow(10, -.1 * su->atten[band] * dist) * x->sizemod * ts;
                hist_add(x, r, band, rt, e, d, ana);
            } else {
                for (int oc = lo; oc <= hi; ++oc) {
                    double e = E[oc] * pow(10, -.1 * su->atten[oc] * dist) * x->sizemod * ts;
                    hist_add(x, r, oc, rt, e, d, 0);
                }
            }
        }
    }
}

void or_receiver_check(const or_setup *su, double **planes, const double o[3], const double d[3],
                       const double end[3], double I, int band, double t_sum, int ana,
                       int64_t *deposits)
{
    or_counters cnt;
    memset(&cnt, 0, sizeof(cnt));
    rx_ctx x = {su, planes, 1.0 / su->c, 1.0 / (M_PI * su->radius * su->radius * su->radius), &cnt};
    check_rays(&x, o, d, end, &I, band, 0, 0, t_sum, ana);
    if (deposits) *deposits = cnt.deposits;
}

/* ------------------------------------------------------------------------ */
/* Mapping receivers (PachMapReceiver, Pach_Map_Receiver.cs).                */
/*   Partition :99-303 (bounding box, voxel counts, receiver lists)          */
/*   CheckBroadbandRay :1359-1491, CheckRay :1493-1631 (voxel walks)         */
/*   CheckReceivers_* :349-403 (each receiver once per segment)              */
/*   Map_Receiver :1721-1826 (sphere test, deposit)                          */
/*   Histogram.Add / Directional_Histogram.Add (Classes_Receivers.cs:1159-  */
/*   1166, :1813-1825)                                                       */
/* Hare's AABB.IsPointInBox / Intersect are taken as the inclusive box and   */
/* the slab entry point (t >= 0); parity of those two is unpinned.           */
/* ------------------------------------------------------------------------ */
struct or_map {
    int M;
    double *ctr;      /* [M][3] */
    double *direct;   /* [M] Direct_Time */
    double r, r2, sizemod, c, rho_c, atten[8];
    int directional, sample_ct, time1pt;
    double bmin[3], bmax[3], vd[3];
    int ct[3];
    int *start;       /* [ncell + 1], cell = (x * ctY + y) * ctZ + z */
    int *list;
};

or_map *or_map_create(const double *centers, int M, double increment, const double src[3], double c,
                      double rho_c, const double atten[8], int directional, int sample_ct)
{
    if (M <= 0 || !(increment > 0) || sample_ct <= 0) return NULL;
    or_map *m = calloc(1, sizeof(or_map));
    m->M = M;
    m->ctr = malloc(sizeof(double) * 3 * M);
    m->direct = malloc(sizeof(double) * M);
    memcpy(m->ctr, centers, sizeof(double) * 3 * M);
    m->r = increment * 0.5;
    m->r2 = m->r * m->r;
    m->sizemod = 1 / (M_PI * m->r2);
    m->c = c;
    m->rho_c = rho_c;
    memcpy(m->atten, atten, sizeof(m->atten));
    m->directional = directional;
    m->sample_ct = sample_ct;
    for (int i = 0; i < M; ++i) {
        const double *P = centers + 3 * i;
        double lx = src[0] - P[0], ly = src[1] - P[1], lz = src[2] - P[2];
        m->direct[i] = sqrt(lx * lx + ly * ly + lz * lz) / c;
    }
    /* bounds of all spheres (:183-192) */
    for (int k = 0; k < 3; ++k) { m->bmin[k] = INFINITY; m->bmax[k] = -INFINITY; }
    for (int i = 0; i < M; ++i)
        for (int k = 0; k < 3; ++k) {
            double v = centers[3 * i + k];
            if (v + increment * .5 > m->bmax[k]) m->bmax[k] = v + increment * .5;
            if (v - increment * .5 < m->bmin[k]) m->bmin[k] = v - increment * .5;
        }
    m->ct[0] = (int)ceil((m->bmax[0] - m->bmin[0]) / (increment * 4));
    m->ct[1] = (int)ceil((m->bmax[1] - m->bmin[1]) / (increment * 4));
    m->ct[2] = (int)ceil((m->bmax[2] - m->bmin[2]) / (increment * 2.5));
    double inv[3];
    for (int k = 0; k < 3; ++k) {
        m->vd[k] = (m->bmax[k] - m->bmin[k]) / m->ct[k];
        inv[k] = 1 / m->vd[k];
    }
    int nc = m->ct[0] * m->ct[1] * m->ct[2];
    int *cnt = calloc(nc + 1, sizeof(int));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            m->start = malloc(sizeof(int) * (nc + 1));
            m->start[0] = 0;
            for (int i = 0; i < nc; ++i) m->start[i + 1] = m->start[i] + cnt[i];
            m->list = malloc(sizeof(int) * (m->start[nc] > 0 ? m->start[nc] : 1));
            memset(cnt, 0, sizeof(int) * (nc + 1));
        }
        for (int i = 0; i < M; ++i) {
            const double *o = centers + 3 * i;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = (int)floor((o[k] - m->r - m->bmin[k]) * inv[k]);
                hi[k] = (int)floor((o[k] + m->r - m->bmin[k]) * inv[k]);
                if (lo[k] < 0) lo[k] = 0;
                if (hi[k] > m->ct[k] - 1) hi[k] = m->ct[k] - 1;
            }
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        /* voxel box as Partition builds it, then the sphere-AABB test */
                        double mn[3] = {x * m->vd[0] + m->bmin[0], y * m->vd[1] + m->bmin[1], z * m->vd[2] + m->bmin[2]};
                        double mx[3] = {(x + 1) * m->vd[0] + m->bmin[0], (y + 1) * m->vd[1] + m->bmin[1],
                                        (z + 1) * m->vd[2] + m->bmin[2]};
                        double dd[3];
                        for (int k = 0; k < 3; ++k) {
                            double a = mn[k] - o[k], b = o[k] - mx[k];
                            dd[k] = a > (b > 0 ? b : 0) ? a : (b > 0 ? b : 0);
                        }
                        if (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < m->r2) {
                            int cell = (x * m->ct[1] + y) * m->ct[2] + z;
                            if (pass == 0) cnt[cell]++;
                            else m->list[m->start[cell] + cnt[cell]++] = i;
                        }
                    }
        }
    }
    free(cnt);
    return m;
}

void or_map_destroy(or_map *m)
{
    if (!m) return;
    free(m->ctr); free(m->direct); free(m->start); free(m->list);
    free(m);
}

int or_map_grid(const or_map *m, int ct[3], double box_min[3], double voxel[3])
{
    for (int k = 0; k < 3; ++k) { ct[k] = m->ct[k]; box_min[k] = m->bmin[k]; voxel[k] = m->vd[k]; }
    return m->ct[0] * m->ct[1] * m->ct[2];
}

int or_map_voxel(const or_map *m, int x, int y, int z, int *out, int cap)
{
    int cell = (x * m->ct[1] + y) * m->ct[2] + z;
    int n = m->start[cell + 1] - m->start[cell];
    for (int i = 0; i < n && i < cap; ++i) out[i] = m->list[m->start[cell] + i];
    return n;
}

/* Histogram.Add / Directional_Histogram.Add at 1000 Hz; samples before 0
 * are skipped (the directional form has no such test and would fail). */
static void map_add(const or_map *m, double **P, int i, int band, double time, double e, const double *d,
                    int64_t *dep)
{
    if (m->time1pt) {
        if (P[0]) P[0][(size_t)i * 8 + band] += e;
        if (dep) (*dep)++;
        return;
    }
    int sample = (int)(time * 1000);
    if (sample >= m->sample_ct || sample < 0) return;
    size_t idx = ((size_t)i * 8 + band) * m->sample_ct + sample;
    if (P[0]) P[0][idx] += e;
    if (m->directional) {
        if (P[1]) P[1][idx] += sqrt(e * m->rho_c);
        for (int k = 0; k < 3; ++k) {
            double dk = -d[k];
            double v = (double)(float)(dk * e);
            if (dk > 0) { if (P[2 + k]) P[2 + k][idx] += v; }
            else { if (P[5 + k]) P[5 + k][idx] += v; }
        }
    } else if (P[1]) {
        P[1][idx] += sqrt(e);
    }
    if (dep) (*dep)++;
}

/* Map_Receiver.CheckRay / CheckBroadbandRay (:1780-1826): chord midpoint
 * before EndPt, deposit at t / c + t_sum - Direct_Time (the air attenuation
 * is applied over that time, as written). */
static void map_receiver(const or_map *m, double **P, int i, const double *o, const double *d,
                         const double *end, const double *E, int band, double t_sum, int64_t *dep)
{
    const double *C = m->ctr + 3 * i;
    double mx = o[0] - C[0], my = o[1] - C[1], mz = o[2] - C[2];
    double b = mx * d[0] + my * d[1] + mz * d[2];
    double c = (mx * mx + my * my + mz * mz) - m->r2;
    if (c > 0 && b > 0) return;
    double discr = b * b - c;
    if (discr < 0) return;
    double t1 = -b - sqrt(discr), t2 = -b + sqrt(discr);
    double t = (t1 + t2) / 2;
    double ex = end[0] - o[0], ey = end[1] - o[1], ez = end[2] - o[2];
    if (!(t > 0 && t * t < ex * ex + ey * ey + ez * ez)) return;
    double raydist = t / m->c + t_sum - m->direct[i];
    if (band >= 0) {
        map_add(m, P, i, band, raydist, E[0] * pow(10, -.1 * m->atten[band] * raydist) * m->sizemod, d, dep);
    } else {
        for (int oc = 0; oc < 8; ++oc)
            map_add(m, P, i, oc, raydist, E[oc] * pow(10, -.1 * m->atten[oc] * raydist) * m->sizemod, d, dep);
    }
}

/* receivers of voxel (x, y, z) not yet checked on this segment */
static void map_voxel_check(const or_map *m, double **P, unsigned char *seen, int x, int y, int z, const double *o,
                            const double *d, const double *end, const double *E, int band, double t_sum,
                            int64_t *dep)
{
    int cell = (x * m->ct[1] + y) * m->ct[2] + z;
    for (int j = m->start[cell]; j < m->start[cell + 1]; ++j) {
        int i = m->list[j];
        if (seen[i]) continue;
        seen[i] = 1;
        map_receiver(m, P, i, o, d, end, E, band, t_sum, dep);
    }
}

static int map_in_box(const or_map *m, const double *p)
{
    for (int k = 0; k < 3; ++k)
        if (p[k] < m->bmin[k] || p[k] > m->bmax[k]) return 0;
    return 1;
}

/* slab entry of the ray into the box (t >= 0) */
static int map_box_entry(const or_map *m, const double *o, const double *d, double *entry)
{
    double t0 = 0, t1 = INFINITY;
    for (int k = 0; k < 3; ++k) {
        if (d[k] == 0) {
            if (o[k] < m->bmin[k] || o[k] > m->bmax[k]) return 0;
            continue;
        }
        double a = (m->bmin[k] - o[k]) / d[k], b = (m->bmax[k] - o[k]) / d[k];
        if (a > b) { double t = a; a = b; b = t; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
    }
    if (t0 > t1) return 0;
    for (int k = 0; k < 3; ++k) entry[k] = o[k] + d[k] * t0;
    return 1;
}

static void map_walk(const or_map *m, double **P, const double *o, const double *d, const double *end,
                     const double *E, int band, double t_sum, int broadband, int64_t *dep)
{
    double start[3] = {o[0], o[1], o[2]};
    if (!map_in_box(m, o))
        if (!map_box_entry(m, o, d, start)) return;
    /* the broadband walk starts from the ray's own origin and steers by the
     * signs of its position; the octave walk uses the entry point and the
     * direction (the reference's two versions differ so) */
    const double *from = broadband ? o : start;
    const double *sgn = broadband ? o : d;
    int X[3], Xe[3], step[3];
    double tmax[3], tdel[3];
    for (int k = 0; k < 3; ++k) {
        X[k] = (int)floor((from[k] - m->bmin[k]) / m->vd[k]);
        Xe[k] = (int)floor((end[k] - m->bmin[k]) / m->vd[k]);
        if (X[k] < 0) X[k] = 0;
        if (X[k] >= m->ct[k]) X[k] = m->ct[k] - 1;
    }
    for (int k = 0; k < 3; ++k)
        if ((sgn[k] < 0 && X[k] < Xe[k]) || (sgn[k] >= 0 && X[k] > Xe[k])) return;
    for (int k = 0; k < 3; ++k) {
        double lo = X[k] * m->vd[k] + m->bmin[k], hi = (X[k] + 1) * m->vd[k] + m->bmin[k];
        if (sgn[k] < 0) {
            step[k] = -1;
            tmax[k] = (lo - from[k]) / sgn[k];
        } else {
            step[k] = 1;
            tmax[k] = (hi - from[k]) / sgn[k];
        }
        tdel[k] = m->vd[k] / sgn[k] * step[k];
    }
    unsigned char *seen = calloc(m->M, 1);
    for (;;) {
        map_voxel_check(m, P, seen, X[0], X[1], X[2], o, d, end, E, band, t_sum, dep);
        int ax;
        if (tmax[0] < tmax[1]) ax = tmax[0] < tmax[2] ? 0 : 2;
        else ax = tmax[1] < tmax[2] ? 1 : 2;
        X[ax] += step[ax];
        if (X[ax] < 0 || X[ax] >= m->ct[ax]) break;
        tmax[ax] = tmax[ax] + tdel[ax];
        int more = broadband ? (X[0] != Xe[0] && X[1] != Xe[1] && X[2] != Xe[2])
                             : (X[0] != Xe[0] || X[1] != Xe[1] || X[2] != Xe[2]);
        if (!more) break;
    }
    free(seen);
}

void or_map_set_time1pt(or_map *m, int on)
{
    m->time1pt = on && m->sample_ct == 1;
}

void or_map_check_octave(const or_map *m, double **planes, const double o[3], const double d[3],
                         const double end[3], double I, int band, double t_sum, int64_t *deposits)
{
    map_walk(m, planes, o, d, end, &I, band, t_sum, 0, deposits);
}

void or_map_check_broadband(const or_map *m, double **planes, const double o[3], const double d[3],
                            const double end[3], const double E[8], double t_sum, int64_t *deposits)
{
    map_walk(m, planes, o, d, end, E, -1, t_sum, 1, deposits);
}

/* ------------------------------------------------------------------------ */
/* Tracer (TaskAll:335-536)                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    double o[3], d[3];
    double I, t;
    int band, tri, mask, phase;
    uint64_t uid;
    uint32_t k; /* bounce (= Philox block) counter of this ray */
} oray;

typedef struct {
    oray *a;
    size_t n, cap;
} ostack;

static void push(ostack *s, const oray *r)
{
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->a = realloc(s->a, s->cap * sizeof(oray));
    }
    s->a[s->n++] = *r;
}

/* sin(2 pi u), cos(2 pi u) for u in [0, 1): quarter-turn reduction, then the
 * Taylor series of sin and cos on [-pi/4, pi/4] to degree 17/18 (truncation
 * below 1e-19), evaluated in a fixed Horner order.  Only +, -, * on doubles, so
 * the CPU oracle and the FP64 GPU kernel produce the same bits. */
static void sincos_turn(double u, double *s, double *c)
{
    int k = (int)floor(4.0 * u + 0.5);
    double r = u - 0.25 * k;
    double a = r * 6.283185307179586476925286766559;
    double z = a * a;
    double sp = 1.0 + z * (-1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 +
                z * (-1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
                z * (1.0 / 355687428096000.0))))))));
    double cp = 1.0 + z * (-1.0 / 2.0 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 +
                z * (-1.0 / 3628800.0 + z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 +
                z * (1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0)))))))));
    double sa = a * sp, ca = cp;
    switch (k & 3) {
    case 0: *s = sa; *c = ca; break;
    case 1: *s = ca; *c = -sa; break;
    case 2: *s = -sa; *c = -ca; break;
    default: *s = -ca; *c = sa; break;
    }
}

/* diffuse direction of Lambert_Scattering (Classes_Materials.cs:4519-4549):
 * frame about the (flipped) normal, azimuth 2 pi U, cos(zenith) = U.  The
 * reference's cos(acos(U)) is U and sin(acos(U)) is sqrt((1-U)(1+U)); both are
 * written that way so no libm transcendental decides a path. */
static void diffuse_dir(const double *N, double u_az, double u_zen, int very_late, double out[3])
{
    (void)very_late; /* Scatter_VeryLate's z = Scat_Mod equals cos(acos(Scat_Mod)) */
    double dz[3] = {N[0], N[1], N[2]}, dx[3] = {0, 0, 1}, dy[3];
    double proj = fabs(dot3(dz, dx));
    if (0.99 < proj && 1.01 > proj) { dx[0] = 1; dx[1] = 0; dx[2] = 0; }
    cross3(dz, dx, dy);
    cross3(dy, dz, dx);
    normalize3(dx); normalize3(dy); normalize3(dz);
    double s1, c1;
    sincos_turn(u_az, &s1, &c1);
    double st = sqrt((1.0 - u_zen) * (1.0 + u_zen));
    double x = c1 * st, y = s1 * st, z = u_zen;
    for (int k = 0; k < 3; ++k) out[k] = dx[k] * x + dy[k] * y + dz[k] * z;
    normalize3(out);
}

void or_sample_diffuse(const double N[3], double u_az, double u_zen, double out[3])
{
    diffuse_dir(N, u_az, u_zen, 0, out);
}

void or_sincos_turn(double u, double *s, double *c) { sincos_turn(u, s, c); }

static void specular(double *d, const double *N, double cos_t)
{
    d[0] -= N[0] * cos_t * 2;
    d[1] -= N[1] * cos_t * 2;
    d[2] -= N[2] * cos_t * 2;
}

int or_h_oct(const or_scene *s, const double power[8], int oct_lo, int oct_hi)
{
    double total[8] = {0};
    for (int q = 0; q < s->S; ++q)
        for (int o = oct_lo; o <= oct_hi; ++o) total[o] += s->area[q] * s->alpha[8 * q + o];
    double mn = INFINITY, op = 0;
    int h = 0;
    for (int o = oct_lo; o <= oct_hi; ++o) if (power[o] > op) op = power[o];
    for (int o = oct_lo; o <= oct_hi; ++o) {
        if (power[o] < op * 1e-6) continue;
        if (mn > total[o]) { mn = total[o]; h = o; }
    }
    return h;
}

/* one closest-hit query; returns triangle or -1; fills hit point */
static int shoot(const or_scene *s, const or_params *p, const double *o, const double *d,
                 double *hit, double *leg)
{
    double t;
    int tri = p->use_grid ? or_shoot_grid(s, o, d, &t) : or_shoot_brute(s, o, d, &t);
    if (tri < 0) {
        for (int k = 0; k < 3; ++k) hit[k] = o[k] + d[k] * 1000000;
        return -1;
    }
    for (int k = 0; k < 3; ++k) hit[k] = o[k] + d[k] * t;
    *leg = t;
    return tri;
}


typedef struct {
    double T2[8], T3[8];
    double COTime;
} thresholds;

/* Scatter_Late (Classes_Materials.cs:4629-4702) */
static void scatter_late(const or_scene *s, oray *r, const double *N, double cos_t, int sf,
                         const uint32_t w[4], int transmit, uint32_t k_event, ostack *st,
                         or_counters *cnt)
{
    int b = r->band;
    double sel = u01(w[1]);
    if (sel > s->hi[8 * sf + b]) {
        specular(r->d, N, cos_t);
        if (transmit) { r->d[0] = -r->d[0]; r->d[1] = -r->d[1]; r->d[2] = -r->d[2]; }
        r->mask &= ~1;
        return;
    } else if (sel > s->lo[8 * sf + b]) {
        /* OctaveRay.SplitRay(1 - s) (Classes_Rays.cs:123-129): child keeps I(1-s) */
        double emc = 1 - s->scat[8 * sf + b];
        oray tr = *r;
        tr.I = r->I * emc;
        r->I *= (1 - emc);
        specular(tr.d, N, cos_t);
        if (transmit) { tr.d[0] = -tr.d[0]; tr.d[1] = -tr.d[1]; tr.d[2] = -tr.d[2]; }
        tr.mask &= ~1;
        tr.uid = or_child_uid(r->uid, TAG_LATE(k_event));
        tr.k = 0;
        tr.phase = 0;
        push(st, &tr);
        cnt->spawned++;
        cnt->octave_rays++;
    }
    double Nf[3] = {N[0], N[1], N[2]};
    if (cos_t > 0) { Nf[0] = -N[0]; Nf[1] = -N[1]; Nf[2] = -N[2]; }
    diffuse_dir(Nf, u01(w[2]), u01(w[3]), 0, r->d);
    r->mask |= 3;
    if (transmit) { r->d[0] = -r->d[0]; r->d[1] = -r->d[1]; r->d[2] = -r->d[2]; }
}

/* Scatter_VeryLate (Classes_Materials.cs:4572-4627) */
static void scatter_verylate(const or_scene *s, oray *r, const double *N, double cos_t, int sf,
                             const uint32_t w[4], int transmit)
{
    if (u01(w[1]) < s->scat[8 * sf + r->band]) {
        double Nf[3] = {N[0], N[1], N[2]};
        if (cos_t > 0) { Nf[0] = -N[0]; Nf[1] = -N[1]; Nf[2] = -N[2]; }
        diffuse_dir(Nf, u01(w[2]), u01(w[3]), 1, r->d);
        r->mask |= 3;
    } else {
        specular(r->d, N, cos_t);
        r->mask &= ~1;
    }
    if (transmit) { r->d[0] = -r->d[0]; r->d[1] = -r->d[1]; r->d[2] = -r->d[2]; }
}

/* Lambert_Scattering.Scatter_Early (Classes_Materials.cs:4498-4570) of a
 * broadband ray at hit point o on triangle ht of polygon sf, energies E after
 * Absorb: per band with E > 0 a diffuse child BroadRay.SplitRay(oct, s)
 * (E s; the ray keeps E (1 - s), Classes_Rays.cs:345-351); with transmission
 * tau > 0 the ray splits again, tr = SplitRay(oct, tau) along the incoming
 * direction, and the child splits td = R.SplitRay(tau), reversed
 * (:4554-4563).  Children are pushed tr, td, R (the reference's enqueue
 * order); d becomes the specular reflection about the flipped normal. */
static void scatter_early(const or_scene *s, int sf, int ht, const double o_[3], double d[3], const double *N,
                          double cos_t, double tb, double E[8], uint64_t root, uint32_t k, const uint32_t w[16],
                          ostack *st, or_counters *cnt)
{
    double Nf[3] = {N[0], N[1], N[2]}, cf = cos_t;
    if (cf > 0) { Nf[0] = -N[0]; Nf[1] = -N[1]; Nf[2] = -N[2]; cf = -cf; }
    for (int oc = 0; oc < 8; ++oc) {
        if (E[oc] == 0) continue;
        double sc = s->scat[8 * sf + oc];
        oray R;
        memcpy(R.o, o_, sizeof(R.o));
        R.I = E[oc] * sc;
        E[oc] *= (1 - sc);
        R.t = tb; R.band = oc; R.tri = ht; R.phase = 0; R.k = 0;
        R.uid = or_child_uid(root, TAG_EARLY(k, oc));
        diffuse_dir(Nf, u01(w[2 * oc]), u01(w[2 * oc + 1]), 0, R.d);
        R.mask = 3;
        double tau = s->trans[8 * sf + oc];
        if (tau > 0) {
            oray tr = R;
            memcpy(tr.d, d, sizeof(tr.d)); /* incoming direction */
            tr.I = E[oc] * tau;
            E[oc] *= (1 - tau);
            tr.mask = 0;
            tr.uid = or_child_uid(root, TAG_EARLY_TR(k, oc));
            push(st, &tr);
            oray td = R;
            td.I = R.I * tau;
            R.I *= (1 - tau);
            td.d[0] = -R.d[0]; td.d[1] = -R.d[1]; td.d[2] = -R.d[2];
            td.mask = 3;
            td.uid = or_child_uid(root, TAG_EARLY_TD(k, oc));
            push(st, &td);
            cnt->spawned += 2;
            cnt->octave_rays += 2;
        }
        push(st, &R);
        cnt->spawned++;
        cnt->octave_rays++;
    }
    specular(d, Nf, cf);
}

/* ---- single scattering events, for the pins of the Lambert scheme ---- */
int or_lambert_bounds(const or_scene *s, int surf, int band, double *lo, double *hi)
{
    if (!s || surf < 0 || surf >= s->S || band < 0 || band > 7) return -1;
    *lo = s->lo[8 * surf + band];
    *hi = s->hi[8 * surf + band];
    return 0;
}

static void ray_out(const oray *r, or_ray_out *o)
{
    memcpy(o->d, r->d, sizeof(o->d));
    o->I = r->I;
    o->mask = r->mask;
    o->band = r->band;
}

int or_scatter_late_once(const or_scene *s, int sf, int band, const double d_in[3], const double N[3], double I,
                         int mask_in, int transmit, const uint32_t w[4], or_ray_out *parent, or_ray_out *child)
{
    oray r;
    memset(&r, 0, sizeof(r));
    memcpy(r.d, d_in, sizeof(r.d));
    r.I = I; r.band = band; r.mask = mask_in;
    ostack st = {NULL, 0, 0};
    or_counters cnt;
    memset(&cnt, 0, sizeof(cnt));
    const double cos_t = dot3(N, r.d);
    scatter_late(s, &r, N, cos_t, sf, w, transmit, 0, &st, &cnt);
    ray_out(&r, parent);
    const int n = (int)st.n;
    if (n > 0 && child) ray_out(&st.a[0], child);
    free(st.a);
    return n;
}

void or_scatter_verylate_once(const or_scene *s, int sf, int band, const double d_in[3], const double N[3], double I,
                              int mask_in, int transmit, const uint32_t w[4], or_ray_out *out)
{
    oray r;
    memset(&r, 0, sizeof(r));
    memcpy(r.d, d_in, sizeof(r.d));
    r.I = I; r.band = band; r.mask = mask_in;
    scatter_verylate(s, &r, N, dot3(N, r.d), sf, w, transmit);
    ray_out(&r, out);
}

int or_scatter_early_once(const or_scene *s, int sf, const double d_in[3], const double N[3], const double E_in[8],
                          const uint32_t w[16], double d_out[3], double E_out[8], or_ray_out *children, int cap)
{
    double o_[3] = {0, 0, 0}, d[3] = {d_in[0], d_in[1], d_in[2]}, E[8];
    memcpy(E, E_in, sizeof(E));
    ostack st = {NULL, 0, 0};
    or_counters cnt;
    memset(&cnt, 0, sizeof(cnt));
    scatter_early(s, sf, 0, o_, d, N, dot3(d, N), 0.0, E, 0, 0, w, &st, &cnt);
    memcpy(d_out, d, sizeof(d));
    memcpy(E_out, E, sizeof(E));
    const int n = (int)st.n;
    for (int i = 0; i < n && i < cap; ++i) ray_out(&st.a[i], &children[i]);
    free(st.a);
    return n;
}

/* Phase 2 (TaskAll:427-480) then End_Trace (TaskAll:489-536) for one octave ray. */
static void trace_octave(const or_scene *s, const or_setup *su, const or_params *p,
                         const rx_ctx *x, const rng_ctx *rng, const thresholds *th,
                         oray r, ostack *st, or_counters *cnt)
{
    const int stoch = (p->mode == OR_MODE_STOCHASTIC);
    double hit[3], leg = 0;
    for (;;) {
        if ((int64_t)r.k >= p->max_bounces) { cnt->guard_hits++; cnt->remaining += r.I; return; }
        int ht = shoot(s, p, r.o, r.d, hit, &leg);
        if (r.phase == 0) cnt->bounces_late++;
        else cnt->bounces_verylate++;
        int ana = (r.mask & 1) ? 1 : ((r.mask & 2) ? 2 : 0);
        if (ht < 0) {
            if (su->map) or_map_check_octave(su->map, su->map_planes, r.o, r.d, hit, r.I, r.band, r.t, &cnt->deposits);
            else check_rays(x, r.o, r.d, hit, &r.I, r.band, 0, 0, r.t, ana);
            cnt->lost += r.I;
            cnt->misses++;
            return;
        }
        if (su->map) or_map_check_octave(su->map, su->map_planes, r.o, r.d, hit, r.I, r.band, r.t, &cnt->deposits);
        else check_rays(x, r.o, r.d, hit, &r.I, r.band, 0, 0, r.t, ana);
        r.t += leg / su->c;
        memcpy(r.o, hit, sizeof(hit));
        r.tri = ht;
        const double *N = s->n + 3 * ht;
        int sf = s->surf[ht];
        /* Basic_Material.Absorb (Classes_Materials.cs:160-164) */
        double cos_t = dot3(N, r.d);
        double I0 = r.I;
        r.I *= (1 - s->alpha[8 * sf + r.band]);
        cnt->absorbed += I0 - r.I;
        uint32_t kev = r.k++;
        if (stoch) {
            uint32_t w[4];
            rng_block(rng, r.uid, kev, w);
            int transmit = u01(w[0]) < s->trans[8 * sf + r.band];
            if (r.phase == 0) scatter_late(s, &r, N, cos_t, sf, w, transmit, kev, st, cnt);
            else scatter_verylate(s, &r, N, cos_t, sf, w, transmit);
        } else {
            specular(r.d, N, cos_t);
        }
        if (r.phase == 0) {
            if (!(r.t < th->COTime && r.I > th->T2[r.band])) r.phase = 1; /* hand-off to End_Trace */
        } else {
            if (!(r.t < th->COTime && r.I > th->T3[r.band])) { cnt->remaining += r.I; return; }
        }
    }
}

/* cast_ray_distributed (TaskAll:335-486) for source ray `src`. */
static void trace_source_ray(const or_scene *s, const or_setup *su, const or_params *p,
                             const rx_ctx *x, int64_t src, int h, ostack *st, or_counters *cnt)
{
    const int stoch = (p->mode == OR_MODE_STOCHASTIC);
    rng_ctx rng = {{(uint32_t)p->seed, (uint32_t)(p->seed >> 32)}, (uint32_t)src};
    thresholds th;
    double E[8];
    th.COTime = p->cutoff_m / su->c;
    for (int o = 0; o < 8; ++o) {
        E[o] = su->power[o];
        cnt->e_init += E[o];
        th.T2[o] = th.T3[o] = 0.0;
    }
    double T1 = E[h] * 1e-2;
    for (int o = p->oct_lo; o <= p->oct_hi; ++o) {
        th.T2[o] = E[o] * pow(10, -4.5);
        th.T3[o] = E[o] * 1e-9;
    }
    double o_[3], d[3], hit[3], leg = 0, tb = 0;
    memcpy(o_, su->src, sizeof(o_));
    or_source_direction(p->dir_mode, p->seed, src, p->dirs, p->n_dirs, d);
    const uint64_t root = 0;
    int order = 0, tri = -1, split = 1;
    uint32_t k = 0;
    st->n = 0;
    /* ---- phase 1: broadband ray (TaskAll:360-417) ---- */
    for (;;) {
        int ht = shoot(s, p, o_, d, hit, &leg);
        cnt->bounces_broad++;
        if (ht < 0) {
            if (order > p->is_order) {
                if (su->map) or_map_check_broadband(su->map, su->map_planes, o_, d, hit, E, tb, &cnt->deposits);
                else check_rays(x, o_, d, hit, E, -1, p->oct_lo, p->oct_hi, tb, 0);
            }
            for (int j = 0; j < 8; ++j) cnt->lost += E[j];
            cnt->misses++;
            split = 0;
            break;
        }
        if (order > p->is_order) {
            if (su->map) or_map_check_broadband(su->map, su->map_planes, o_, d, hit, E, tb, &cnt->deposits);
            else check_rays(x, o_, d, hit, E, -1, p->oct_lo, p->oct_hi, tb, 0);
        }
        tb += leg / su->c;
        memcpy(o_, hit, sizeof(hit));
        order++;
        tri = ht;
        const double *N = s->n + 3 * ht;
        int sf = s->surf[ht];
        /* Basic_Material.Absorb for BroadRay (Classes_Materials.cs:166-178) */
        double cos_t = dot3(d, N);
        for (int o = 0; o < 8; ++o) {
            double e0 = E[o];
            E[o] *= (1 - s->alpha[8 * sf + o]);
            cnt->absorbed += e0 - E[o];
        }
        if (stoch) {
            uint32_t w[16];
            for (int j = 0; j < 4; ++j) rng_block(&rng, root, 4 * k + j, w + 4 * j);
            scatter_early(s, sf, ht, o_, d, N, cos_t, tb, E, root, k, w, st, cnt);
        } else {
            specular(d, N, cos_t);
        }
        k++;
        if (!(E[h] > T1)) break;
        if ((int)order >= p->p1_max_bounces) { cnt->p1_cap_hits++; break; }
    }
    if (split) {
        for (int oc = 0; oc < 8; ++oc)
            if (oc < p->oct_lo || oc > p->oct_hi) cnt->remaining += E[oc];
        /* BroadRay.SplitRay(o) per octave (TaskAll:419-423; Classes_Rays.cs:331-337) */
        for (int oc = p->oct_lo; oc <= p->oct_hi; ++oc) {
            oray R;
            memcpy(R.o, o_, sizeof(o_));
            memcpy(R.d, d, sizeof(d));
            R.I = E[oc];
            E[oc] = 0;
            R.t = tb; R.band = oc; R.tri = tri; R.mask = 0; R.phase = 0; R.k = 0;
            R.uid = or_child_uid(root, TAG_SPLIT(oc));
            push(st, &R);
            cnt->octave_rays++;
        }
    }
    /* ---- phases 2 and 3 over the queued octave rays ---- */
    while (st->n > 0) {
        oray r = st->a[--st->n];
        trace_octave(s, su, p, x, &rng, &th, r, st, cnt);
    }
}

int or_trace(const or_scene *s, const or_setup *su, const or_params *p,
             double **planes, or_counters *cnt)
{
    if (!s || !su || !p || !planes || !cnt) return -1;
    if (p->oct_lo < 0 || p->oct_hi > 7 || p->oct_lo > p->oct_hi) return -2;
    rx_ctx x = {su, planes, 1.0 / su->c, 1.0 / (M_PI * su->radius * su->radius * su->radius), cnt};
    int h = or_h_oct(s, su->power, p->oct_lo, p->oct_hi);
    cnt->h_oct = h;
    ostack st = {0};
    for (int64_t src = p->ray_begin; src < p->ray_end; ++src)
        trace_source_ray(s, su, p, &x, src, h, &st, cnt);
    free(st.a);
    return 0;
}

typedef struct {
    const or_scene *s;
    const or_setup *su;
    or_setup su_own;            /* the caller's setup with this thread's map planes */
    double *map_planes[8];
    or_params p;
    double *planes[OR_NPLANES];
    or_counters cnt;
} mt_job;

static void *mt_worker(void *arg)
{
    mt_job *j = (mt_job *)arg;
    or_trace(j->s, j->su, &j->p, j->planes, &j->cnt);
    return NULL;
}

int or_trace_mt(const or_scene *s, const or_setup *su, const or_params *p,
                double **planes, or_counters *cnt, int n_threads)
{
    if (n_threads <= 1) return or_trace(s, su, p, planes, cnt);
    size_t plane_n = (size_t)su->R * 8 * su->sample_ct;
    size_t map_n = su->map ? (size_t)su->map->M * 8 * su->map->sample_ct : 0;
    mt_job *jobs = calloc(n_threads, sizeof(mt_job));
    pthread_t *th = calloc(n_threads, sizeof(pthread_t));
    int64_t n = p->ray_end - p->ray_begin;
    for (int i = 0; i < n_threads; ++i) {
        jobs[i].s = s; jobs[i].su = su; jobs[i].p = *p;
        jobs[i].p.ray_begin = p->ray_begin + n * i / n_threads;
        jobs[i].p.ray_end = p->ray_begin + n * (i + 1) / n_threads;
        for (int k = 0; k < OR_NPLANES; ++k)
            jobs[i].planes[k] = planes[k] ? calloc(plane_n, sizeof(double)) : NULL;
        if (su->map) {
            jobs[i].su_own = *su;
            for (int k = 0; k < 8; ++k)
                jobs[i].map_planes[k] = su->map_planes[k] ? calloc(map_n, sizeof(double)) : NULL;
            jobs[i].su_own.map_planes = jobs[i].map_planes;
            jobs[i].su = &jobs[i].su_own;
        }
        pthread_create(&th[i], NULL, mt_worker, &jobs[i]);
    }
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
    for (int i = 0; i < n_threads; ++i) {
        for (int k = 0; k < OR_NPLANES; ++k) {
            if (!planes[k]) continue;
            for (size_t q = 0; q < plane_n; ++q) planes[k][q] += jobs[i].planes[k][q];
            free(jobs[i].planes[k]);
        }
        for (int k = 0; k < 8 && su->map; ++k) {
            if (!su->map_planes[k]) continue;
            for (size_t q = 0; q < map_n; ++q) su->map_planes[k][q] += jobs[i].map_planes[k][q];
            free(jobs[i].map_planes[k]);
        }
        or_counters *c = &jobs[i].cnt;
        cnt->bounces_broad += c->bounces_broad;
        cnt->bounces_late += c->bounces_late;
        cnt->bounces_verylate += c->bounces_verylate;
        cnt->spawned += c->spawned;
        cnt->misses += c->misses;
        cnt->deposits += c->deposits;
        cnt->octave_rays += c->octave_rays;
        cnt->p1_cap_hits += c->p1_cap_hits;
        cnt->guard_hits += c->guard_hits;
        cnt->lost += c->lost;
        cnt->e_init += c->e_init;
        cnt->absorbed += c->absorbed;
        cnt->remaining += c->remaining;
        cnt->h_oct = c->h_oct;
    }
    free(jobs);
    free(th);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Decay-curve integration (Classes_Math.cs:282-585) and diffuse-field       */
/* estimates (:216-270).                                                     */
/* ------------------------------------------------------------------------ */
void or_schroeder(const double *etc, int n, double *out)
{
    double sum = 0;
    double *b = malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) b[i] = sum += etc[i];
    for (int i = 0; i < n; ++i) out[i] = (sum == 0) ? 0 : 1 - (b[i] / sum);
    free(b);
}

void or_log10data(const double *x, int n, double limit_db, double *out)
{
    double lim = pow(10, limit_db / 10);
    if (n > 0) out[0] = 0;
    for (int i = 1; i <= n - 1; ++i) out[i] = (x[i] >= lim) ? 10 * log10(x[i]) : limit_db;
}

void or_schroeder_limits(const double *ls, int n, double first, double second, int lim[2])
{
    lim[0] = lim[1] = 0;
    for (int i = 1; i <= n - 1; ++i)
        if (ls[i] < first) { lim[0] = i; break; }
    for (int i = lim[0]; i <= n - 1; ++i)
        if (ls[i] < second) { lim[1] = i; break; }
}

void or_lsf(const double *mag, int n, double fs, double out[2])
{
    double ts = 0, ms = 0, tt = 0, tm = 0;
    for (int i = 0; i <= n - 1; ++i) {
        double q = i / fs;
        ts += q; ms += mag[i]; tt += q * q; tm += q * mag[i];
    }
    double slope = ((n * tm) - (ts * ms)) / ((n * tt) - (ts * ts));
    out[0] = slope;
    out[1] = (ms - slope * ts) / n;
}

static double fit_window(const double *sch, int n, double first, double second, int fs)
{
    double *ls = malloc(sizeof(double) * (n > 0 ? n : 1));
    or_log10data(sch, n, -60, ls);
    int lim[2];
    or_schroeder_limits(ls, n, first, second, lim);
    int m = lim[1] - lim[0] + 1;
    double c[2] = {0, 0};
    if (m > 0) or_lsf(ls + lim[0], m, fs, c);
    free(ls);
    return -60 / c[0];
}

double or_t_x(const double *sch, int n, int depth, int fs)
{
    return fit_window(sch, n, -5, -5 - abs(depth), fs);
}

double or_edt(const double *sch, int n, int fs) { return fit_window(sch, n, -1, -10.0001, fs); }

double or_clarity(const double *etc, int n, double fs, double seconds, double start, int pressure)
{
    if (pressure) seconds += 8192.0 / fs;
    double bw = 1 / fs;
    int si = (int)floor(start * fs), ei = (int)floor((seconds + start) * fs);
    double early = 0, late = 0;
    for (int q = si; q < ei; ++q) early += etc[q];
    double split = (seconds + start - (ei * bw)) / bw;
    early += split * etc[ei];
    for (int q = ei + 1; q < n; ++q) late += etc[q];
    late += (1 - split) * etc[ei];
    return 10 * log10(early / late);
}

void or_sabine(const or_scene *s, double volume, const double atten[8], double out[8])
{
    for (int t = 0; t < 8; ++t) {
        double ta = 0;
        for (int q = 0; q < s->S; ++q) ta += s->area[q] * s->alpha[8 * q + t];
        out[t] = 0.161 * (volume / (ta + (4 * atten[t] * volume)));
    }
}

void or_eyring(const or_scene *s, double volume, const double atten[8], double out[8])
{
    for (int t = 0; t < 8; ++t) {
        double ta = 0;
        for (int q = 0; q < s->S; ++q) ta += s->area[q] * (-log(1 - s->alpha[8 * q + t]));
        out[t] = 0.161 * volume / (ta + 4 * atten[t] * volume);
    }
}

/* ------------------------------------------------------------------------ */
/* Image sources: ImageSourceData for planar objects without edge            */
/* diffraction (Classes_ImageSource.cs).                                     */
/*   TraverseOrders :266-348   every sequence of 1..MaxOrder objects,         */
/*                              repeats allowed, depth-first                  */
/*   ProcessImages  :350-476   backward validation, final checks, energy     */
/*   OcclusionIntersect :1350-1392, FinalOcclusion :1245-1281                 */
/*   Power_Recursion :535-803 (planar branch), project_wavefront :805-857     */
/*   B_Function_NoInterp :859-890, H_Function :892-1030 (planar branch)       */
/*   Specular_Path :1914-1982                                                */
/*   Polygon_Scene.Image :701-707                                            */
/* Hare's X_Event of a miss is taken as t = 0 (the library is not part of the */
/* reference): OcclusionIntersect then fails and FinalOcclusion reports open. */
/* ------------------------------------------------------------------------ */
/* transmissive surfaces a check passes before it gives up (as a miss) */
#define IS_PASS_LIMIT (1 << 16)

typedef struct {
    const or_scene *s;
    const int *obj;    /* [S] */
    const double *pl;  /* [n_obj][6]: unit normal, point */
} is_ctx;

static int is_transmissive(const or_scene *s, int p)
{
    for (int b = 0; b < 8; ++b)
        if (s->trans[8 * p + b] > 0) return 1;
    return 0;
}

static void is_normalize(double *v)
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
}

static double is_len(const double *a, const double *b)
{
    double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return sqrt(x * x + y * y + z * z);
}

static int is_sign(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

/* closest hit; X = o + d t */
static int is_shoot(const is_ctx *k, const double o[3], const double d[3], double *t, double X[3])
{
    int tri = or_shoot_grid(k->s, o, d, t);
    if (tri >= 0)
        for (int j = 0; j < 3; ++j) X[j] = o[j] + d[j] * *t;
    return tri;
}

/* OcclusionIntersect (:1350-1392): from O towards the image E, passing
 * transmissive polygons of other objects, the first surface must belong to
 * object polyX. */
static int is_occl_intersect(const is_ctx *k, const double O[3], const double E[3], int polyX, double tm[8],
                             double Xout[3], int *poly_out)
{
    const or_scene *s = k->s;
    double D[3] = {E[0] - O[0], E[1] - O[1], E[2] - O[2]};
    double L = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    D[0] /= L; D[1] /= L; D[2] /= L;
    double o[3] = {O[0], O[1], O[2]}, X[3], t = 0;
    int tri = -1;
    for (int guard = 0; guard < IS_PASS_LIMIT; ++guard) {
        tri = is_shoot(k, o, D, &t, X);
        if (tri >= 0) {
            int p = s->surf[tri];
            if (is_transmissive(s, p) && polyX != k->obj[p]) {
                for (int b = 0; b < 8; ++b) tm[b] *= (1 - s->alpha[8 * p + b]) * s->trans[8 * p + b];
                memcpy(o, X, sizeof(o));
                continue;
            }
        }
        break;
    }
    if (tri < 0) return 0;
    if (t < 0.000001) return 0;
    if (polyX != k->obj[s->surf[tri]]) return 0;
    *poly_out = s->surf[tri];
    memcpy(Xout, X, sizeof(X));
    return 1;
}

/* FinalOcclusion (:1245-1281): from the first reflection point back to the
 * source; 1 = occluded. */
static int is_final_occlusion(const is_ctx *k, const double S[3], const double P1[3], int polyX, double tm[8])
{
    const or_scene *s = k->s;
    double D[3] = {S[0] - P1[0], S[1] - P1[1], S[2] - P1[2]};
    double L = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    double L2 = 0;
    is_normalize(D);
    double o[3] = {P1[0], P1[1], P1[2]}, X[3], t = 0;
    int tri = -1;
    for (int guard = 0; guard < IS_PASS_LIMIT; ++guard) {
        tri = is_shoot(k, o, D, &t, X);
        if (tri >= 0) {
            int p = s->surf[tri];
            if (is_transmissive(s, p) && polyX != k->obj[p] && t < L) {
                for (int b = 0; b < 8; ++b) tm[b] *= (1 - s->alpha[8 * p + b]) * s->trans[8 * p + b];
                memcpy(o, X, sizeof(o));
                L2 += t;
                continue;
            }
        }
        break;
    }
    if (tri < 0) return 0;
    L2 += t;
    return L2 < L;
}

/* ProcessImages (:350-476) for one sequence and one receiver, planar objects:
 * one image per reflection, one path.  Returns 1 and fills *out when valid. */
static int is_process(const is_ctx *k, const int *seq, int n, const double (*img)[3], const double src[3],
                      const double power[8], const double E[3], const double atten[8], double c, double rho_c,
                      or_is_path *out)
{
    const or_scene *s = k->s;
    double P[OR_IS_MAX_ORDER + 2][3];
    int polys[OR_IS_MAX_ORDER];
    double tm[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    memcpy(P[0], src, sizeof(P[0]));
    memcpy(P[n + 1], E, sizeof(P[0]));
    for (int q = n - 1; q >= 0; --q)
        if (!is_occl_intersect(k, P[q + 2], img[q], seq[q], tm, P[q + 1], &polys[q])) return 0;
    /* final checks (:423-441): both legs on one side of each reflecting plane,
     * then the first leg unobstructed */
    for (int i = 0; i < n; ++i) {
        double d1[3], d2[3];
        for (int j = 0; j < 3; ++j) {
            d1[j] = P[i][j] - P[i + 1][j];
            d2[j] = P[i + 2][j] - P[i + 1][j];
        }
        const double *nrm = k->pl + 6 * seq[i];
        is_normalize(d1);
        is_normalize(d2);
        if (is_sign(nrm[0] * d1[0] + nrm[1] * d1[1] + nrm[2] * d1[2]) !=
            is_sign(nrm[0] * d2[0] + nrm[1] * d2[1] + nrm[2] * d2[2]))
            return 0;
        if (i == n - 1 && is_final_occlusion(k, P[0], P[1], seq[0], tm)) return 0;
    }
    /* source power along the first leg (Source.DirPower: no directivity) */
    for (int b = 0; b < 8; ++b) tm[b] *= power[b];
    /* Power_Recursion, planar branch: wavefront curvature 1 / length of the
     * first leg, projected over every later leg */
    double kv = 1.0 / is_len(P[1], P[0]);
    double l0 = is_len(P[0], P[1]);
    double t0 = l0 / c, t1 = l0 / c;
    for (int o = 0; o < n; ++o) {
        double l = is_len(P[o + 1], P[o + 2]);
        kv = (kv == INFINITY) ? 0 : 1 / kv;
        kv += l;
        kv = (kv == INFINITY) ? 0 : 1 / kv;
    }
    double h = sqrt(fabs((1 / (4 * M_PI)) * (kv * kv - 0.0 * 0.0)));
    double p = fabs(h);
    /* Specular_Path */
    double length = 0;
    for (int q = 1; q < n + 2; ++q) {
        double x = P[q][0] - P[q - 1][0], y = P[q][1] - P[q - 1][1], z = P[q][2] - P[q - 1][2];
        length += sqrt(x * x + y * y + z * z);
    }
    for (int b = 0; b < 8; ++b) {
        /* B_Function_NoInterp: amplitude with the air attenuation of the mean
         * of the time limits (here the first leg's travel time) */
        double mod = sqrt(tm[b] * pow(10, -.05 * atten[b] * c * 0.5 * (t0 + t1)));
        double B = p * mod;
        /* H_Function: pressure of the source's DirPressure (Source.Rho_C 411.6) */
        double H = B * sqrt(power[b] * 411.6);
        double Hrc = H * H / rho_c;
        double e = pow(10, -.1 * atten[b] * length) * Hrc;
        e *= tm[b];
        out->energy[b] = e;
    }
    for (int q = 0; q < n; ++q) {
        int pq = polys[q];
        for (int b = 0; b < 8; ++b) {
            out->energy[b] *= (1 - s->alpha[8 * pq + b]) * (1 - s->scat[8 * pq + b]) * (1 - s->trans[8 * pq + b]);
            if (out->energy[b] == 0) out->energy[b] = 1E-22;
        }
    }
    out->order = n;
    for (int q = 0; q < n; ++q) {
        out->seq[q] = seq[q];
        out->poly[q] = polys[q];
    }
    memcpy(out->pts, P, sizeof(double) * 3 * (n + 2));
    out->length = length;
    out->time = length / c;
    return 1;
}

typedef struct {
    const is_ctx *k;
    int n_obj, max_order, R;
    const double *src, *power, *recs, *atten;
    double c, rho_c;
    or_is_path *out;
    int64_t cap, n;
} is_run;

static void is_traverse(is_run *r, int *seq, int order, double (*img)[3])
{
    const double *prev = order > 0 ? img[order - 1] : r->src;
    for (int q = 0; q < r->n_obj; ++q) {
        if (order == 0 && q > 0) break; /* caller sets seq[0] */
        if (order > 0) seq[order] = q;
        /* Polygon_Scene.Image: mirror over the plane of the object's first polygon */
        const double *pl = r->k->pl + 6 * seq[order];
        double dist = pl[0] * (prev[0] - pl[3]) + pl[1] * (prev[1] - pl[4]) + pl[2] * (prev[2] - pl[5]);
        for (int j = 0; j < 3; ++j) img[order][j] = prev[j] - 2 * pl[j] * dist;
        for (int rec = 0; rec < r->R; ++rec) {
            or_is_path tmp;
            memset(&tmp, 0, sizeof(tmp));
            if (is_process(r->k, seq, order + 1, (const double (*)[3])img, r->src, r->power, r->recs + 3 * rec,
                           r->atten, r->c, r->rho_c, &tmp)) {
                tmp.rec = rec;
                if (r->n < r->cap) r->out[r->n] = tmp;
                r->n++;
            }
        }
        if (order + 1 < r->max_order) is_traverse(r, seq, order + 1, img);
    }
}

static int64_t is_range(const or_scene *s, const int *obj_of_surf, int n_obj, int max_order, const double src[3],
                        const double power[8], const double *recs, int R, const double atten[8], double c,
                        double rho_c, or_is_path *out, int64_t cap, int q0_begin, int q0_end);

int64_t or_image_sources(const or_scene *s, const int *obj_of_surf, int n_obj, int max_order,
                         const double src[3], const double power[8], const double *recs, int R,
                         const double atten[8], double c, double rho_c, or_is_path *out, int64_t cap)
{
    return is_range(s, obj_of_surf, n_obj, max_order, src, power, recs, R, atten, c, rho_c, out, cap, 0, n_obj);
}

typedef struct {
    const or_scene *s;
    const int *obj;
    int n_obj, max_order, R, q0_begin, q0_end;
    const double *src, *power, *recs, *atten;
    double c, rho_c;
    int64_t n;
} is_mt_arg;

static void *is_mt_worker(void *p)
{
    is_mt_arg *a = (is_mt_arg *)p;
    a->n = is_range(a->s, a->obj, a->n_obj, a->max_order, a->src, a->power, a->recs, a->R, a->atten, a->c,
                    a->rho_c, NULL, 0, a->q0_begin, a->q0_end);
    return NULL;
}

/* Counts the valid paths of the sequences whose first object lies in
 * [q0_begin, q0_end), split over n_threads as Begin splits the objects over
 * its threads (:117-150).  The CPU timing sample. */
int64_t or_image_sources_count_mt(const or_scene *s, const int *obj_of_surf, int n_obj, int max_order,
                                  const double src[3], const double power[8], const double *recs, int R,
                                  const double atten[8], double c, double rho_c, int q0_begin, int q0_end,
                                  int n_threads)
{
    if (n_threads < 1) n_threads = 1;
    if (q0_begin < 0 || q0_end > n_obj || q0_begin >= q0_end) return -1;
    int span = q0_end - q0_begin;
    if (n_threads > span) n_threads = span;
    pthread_t th[256];
    is_mt_arg args[256];
    if (n_threads > 256) n_threads = 256;
    for (int i = 0; i < n_threads; ++i) {
        is_mt_arg a = {s, obj_of_surf, n_obj, max_order, R, q0_begin + (int)((int64_t)span * i / n_threads),
                       q0_begin + (int)((int64_t)span * (i + 1) / n_threads), src, power, recs, atten, c, rho_c, 0};
        args[i] = a;
        pthread_create(&th[i], NULL, is_mt_worker, &args[i]);
    }
    int64_t n = 0;
    for (int i = 0; i < n_threads; ++i) {
        pthread_join(th[i], NULL);
        if (args[i].n < 0) return -1;
        n += args[i].n;
    }
    return n;
}

static int64_t is_range(const or_scene *s, const int *obj_of_surf, int n_obj, int max_order, const double src[3],
                        const double power[8], const double *recs, int R, const double atten[8], double c,
                        double rho_c, or_is_path *out, int64_t cap, int q0_begin, int q0_end)
{
    if (max_order < 1 || max_order > OR_IS_MAX_ORDER || n_obj < 1) return -1;
    int *obj = malloc(sizeof(int) * s->S);
    double *pl = calloc((size_t)n_obj * 6, sizeof(double));
    int *first_tri = malloc(sizeof(int) * n_obj);
    for (int o = 0; o < n_obj; ++o) first_tri[o] = -1;
    for (int p = 0; p < s->S; ++p) obj[p] = obj_of_surf ? obj_of_surf[p] : p;
    /* ObjectMembers[o][0]: the object's lowest polygon; its first triangle */
    int *first_poly = malloc(sizeof(int) * n_obj);
    for (int o = 0; o < n_obj; ++o) first_poly[o] = -1;
    for (int p = 0; p < s->S; ++p)
        if (obj[p] >= 0 && obj[p] < n_obj && first_poly[obj[p]] < 0) first_poly[obj[p]] = p;
    for (int t = 0; t < s->T; ++t) {
        int p = s->surf[t], o = obj[p];
        if (o >= 0 && o < n_obj && first_poly[o] == p && first_tri[o] < 0) first_tri[o] = t;
    }
    for (int o = 0; o < n_obj; ++o) {
        int t = first_tri[o];
        if (t < 0) { pl[6 * o] = 0; continue; } /* object without triangles: never hit */
        memcpy(pl + 6 * o, s->n + 3 * t, sizeof(double) * 3);
        memcpy(pl + 6 * o + 3, s->v0 + 3 * t, sizeof(double) * 3);
    }
    is_ctx k = {s, obj, pl};
    is_run r = {&k, n_obj, max_order, R, src, power, recs, atten, c, rho_c, out, cap, 0};
    int seq[OR_IS_MAX_ORDER];
    double img[OR_IS_MAX_ORDER][3];
    /* Begin / Calculate (:117-170): one depth-first traversal per first object */
    for (int q0 = q0_begin; q0 < q0_end; ++q0) {
        seq[0] = q0;
        is_traverse(&r, seq, 0, img);
    }
    free(obj); free(pl); free(first_tri); free(first_poly);
    return r.n;
}
