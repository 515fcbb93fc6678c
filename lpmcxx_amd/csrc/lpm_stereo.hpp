// Stereographic far field of the sphere Biot-Savart kernel.
//
// For unit vectors t, y and their stereographic coordinates z_t, z_y (complex),
//   1 - t.y = 2 |z_t - z_y|^2 / ((1 + |z_t|^2) (1 + |z_y|^2)),
// so the streamfunction psi(t) = -(1/4 pi) sum_j g_j log(1 - t.y_j) is the
// plane log kernel plus terms that separate, and the velocity
// u = -t x grad psi follows from W(z_t) = sum_j g_j / (z_t - z_j) and the
// total strength.  About a node centre z_c the Cauchy sum is the multipole
// series  sum_j g_j / (z - z_j) = sum_k a_k / (z - z_c)^(k+1),
// a_k = sum_j g_j (z_j - z_c)^k:  P + 1 complex terms for the truncation
// error (rho / |z - z_c|)^(P+1) that the Cartesian series reaches with
// (P + 1)(P + 2)(P + 3) / 6 terms.
//
// Two projections keep every target batch within 90 degrees of its
// projection centre: 0, from the south pole, z = (x + i y) / (1 + z3)
// (batches with z3 >= 0); 1, from the north pole, z = (x - i y) / (1 - z3),
// which is projection 0 after the rotation (x, y, z) -> (x, -y, -z).
//
// Unscaled velocity sum of a target (the sum of g (t x y) / (1 - t.y) that
// the far and near fields add up): with W~ = S - G conj(z) / (1 + |z|^2),
// S = sum over the target's stereographic far nodes of their Cauchy series,
// G = the sum of their total strengths, and z = xi + i eta the target's
// coordinate in the batch's projection,
//   U = -(Im W~ (1 + eta^2 - xi^2, -2 xi eta, -2 xi)
//         + Re W~ (-2 xi eta, 1 + xi^2 - eta^2, -2 eta)),
// rotated back by (x, y, z) -> (x, -y, -z) for projection 1.
#pragma once

#include <hip/hip_runtime.h>

namespace lpm {

constexpr int kStereoTerms = 16;  // a_0 .. a_15 per node and projection

// Stereographic coordinate of the unit vector (x, y, z) in projection
// `proj`; false (and 0) within 1e-12 of the projection's pole.
__device__ inline bool stereo_point(int proj, double x, double y, double z, double& re, double& im) {
    const double den = proj == 0 ? 1.0 + z : 1.0 - z;
    if (!(den > 1e-12)) {
        re = im = 0.0;
        return false;
    }
    const double inv = 1.0 / den;
    re = x * inv;
    im = (proj == 0 ? y : -y) * inv;
    return true;
}

}  // namespace lpm
