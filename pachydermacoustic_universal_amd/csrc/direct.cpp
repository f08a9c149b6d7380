// direct.cpp -- deterministic direct sound of a point source (Direct_Sound,
// Classes_Direct.cs:654-840), the first part of the energy-time curve that
// IR_Construction.ETCurve (Classes_Math.cs:1417-1474) assembles from the direct
// sound, image sources and the ray tracer's histograms.  Host code: O(R x T).
#include "../../include/pach.h"

#include <cmath>
#include <vector>

namespace {

// Closest hit over every triangle (FP64 Moller-Trumbore, t > 1e-10, ties to
// the lower index): Room.shoot as Check_Validity uses it.
int shoot_all(const double *v, int T, const double o[3], const double d[3], double &tbest)
{
    int best = -1;
    tbest = 1e300;
    for (int t = 0; t < T; ++t) {
        const double *a = v + 9 * t;
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
            best = t;
        }
    }
    return best;
}

}  // namespace

extern "C" int pach_direct_sound(const double *verts, int T, const int *surf_of_tri, const double *trans, int S,
                                 const double src[3], const double power[8], const double *centers, int R,
                                 const double atten[8], double c, double *energy, double *time_s, int *audible)
{
    if (!verts || T <= 0 || !surf_of_tri || S <= 0 || !src || !power || !centers || R <= 0 || !atten || !(c > 0) ||
        !energy || !time_s || !audible)
        return PACH_EINVAL;
    for (int t = 0; t < T; ++t)
        if (surf_of_tri[t] < 0 || surf_of_tri[t] >= S) return PACH_EINVAL;
    for (int i = 0; i < R; ++i) {
        const double *P = centers + 3 * i;
        double d[3] = {P[0] - src[0], P[1] - src[1], P[2] - src[2]};
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(dist > 0)) return PACH_EINVAL;
        for (double &x : d) x /= dist;
        // Check_Validity (Classes_Direct.cs:654-686): audible when the first
        // surface lies at or past the receiver; through a transmissive
        // surface the search continues from the hit point, scaling by the
        // transmission coefficients and still comparing with the whole
        // source-receiver distance, as the reference does.
        double tm[8] = {1, 1, 1, 1, 1, 1, 1, 1};
        double o[3] = {src[0], src[1], src[2]};
        bool valid = true;
        for (int guard = 0; guard < T + 1; ++guard) {
            double th;
            const int hit = shoot_all(verts, T, o, d, th);
            if (hit < 0) { valid = true; break; }
            valid = th >= dist;
            bool transmissive = false;
            if (trans)
                for (int b = 0; b < 8; ++b) transmissive |= trans[8 * surf_of_tri[hit] + b] > 0;
            if (!valid && transmissive) {
                for (int b = 0; b < 8; ++b) tm[b] *= trans[8 * surf_of_tri[hit] + b];
                for (int k = 0; k < 3; ++k) o[k] += d[k] * th;
                continue;
            }
            break;
        }
        audible[i] = valid ? 1 : 0;
        // Time_Pt = d / c (+ Delay_ms = 0); a point source's arrival is sample
        // 0 of the direct-sound window (Classes_Direct.cs:749, :812-823)
        time_s[i] = dist / c;
        for (int b = 0; b < 8; ++b)
            energy[8 * i + b] = power[b] * std::pow(10, -.1 * atten[b] * dist) * tm[b] / (4 * M_PI * dist * dist);
    }
    return PACH_OK;
}
