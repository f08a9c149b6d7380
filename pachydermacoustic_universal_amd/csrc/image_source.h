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

// One order of the image-source enumeration: thread i takes sequence i / R
// (base-n_obj digits, source side most significant) and receiver i % R.
struct IsJob {
    const int *obj;        // [S] object of each polygon
    const IsPlane *plane;  // [n_obj]
    const double4 *rec;    // [R] receiver centres
    int n_obj, order, R;
    double src[3];
    double power[8];       // W per band
    double atten[8];       // dB/m
    double c, rho_c;
    pach_is_path *out;     // [cap]
    unsigned long long *count;
    unsigned long long cap;
};

hipError_t launch_image_sources(hipStream_t st, int grid, const DevScene<double> &s, const IsJob &j, int64_t lo,
                                int64_t hi);

}  // namespace pach
