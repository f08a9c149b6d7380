// image_source.h -- launcher of the image-source kernel (image_source.hip).
#pragma once

#include "pach.h"
#include "trace_types.h"

namespace pach {

// Plane of an object: unit normal and a point (its lowest polygon's first
// triangle), the plane Polygon_Scene.Image mirrors in (Classes_Scenes.cs:701-707).
struct IsPlane {
    double n[3];
    double p[3];
};

// One order of the image-source enumeration: item i takes sequence i / R
// (base-n_obj digits, source side most significant) and receiver i % R.  A
// launch covers items [lo, lo + count) with count < 2^31; base_rec and
// base_seq are the receiver and digits of item lo, so threads decode their
// item with 32-bit arithmetic.
struct IsJob {
    const int *obj;        // [S] object of each polygon
    const IsPlane *plane;  // [n_obj]
    const double4 *rec;    // [R] receiver centres
    const int *obj_tri0;   // [n_obj + 1] first entry of each object in obj_tris
    const int *obj_tris;   // triangles of each object
    int n_obj, order, R;
    int any_trans;         // a polygon transmits: checks pass through it
    double src[3];
    double power[8];       // W per band
    double atten[8];       // dB/m
    double c, rho_c;
    double inv_R, inv_N;   // 1 / R, 1 / n_obj
    int base_rec;
    int base_seq[PACH_IS_MAX_ORDER];
    pach_is_path *out;     // [cap]
    unsigned long long *count;
    unsigned long long cap;
};

// queue: null, or room for count item indices, with *queued zero: the items
// are screened first and only the survivors validated.
hipError_t launch_image_sources(hipStream_t st, int grid, const DevScene<double> &s, const IsJob &j, uint32_t count,
                                uint32_t *queue, unsigned *queued);

}  // namespace pach
