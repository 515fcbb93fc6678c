// Stereographic node moments for the learned FP32C evaluation
// (lpm_stereo.hpp): for each node and projection the centre z_c (the
// projection of the node's box centroid direction), the radius rho of its
// sources about z_c, and a_k = sum_j g_j (z_j - z_c)^k, k < kStereoTerms,
// in FP64 (unscaled) and as FP32 a_k / rho^k.  Leaves sum their sources;
// internal nodes shift their children's moments (a binomial transform) level
// by level.  A node whose centre or one of whose sources lies within 1e-12 of
// the projection's pole gets rho = +inf: it is never evaluated in that
// projection.  No counterpart in the reference.
#include "lpm_stereo.hpp"
#include "lpm_tree.hpp"

#include <algorithm>
#include <cmath>

namespace lpm {
namespace {

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

__device__ inline bool node_centre(const double4& c, int proj, double& re, double& im) {
    const double r = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    if (!(r > 0.0)) {
        re = im = 0.0;
        return false;
    }
    return stereo_point(proj, c.x / r, c.y / r, c.z / r, re, im);
}

__device__ inline void store_node(int64_t i, bool ok, double cr, double ci, double rho, const double2* a,
                                  double4* __restrict__ sgeo, double2* __restrict__ smom,
                                  float2* __restrict__ smom32) {
    sgeo[i] = make_double4(cr, ci, ok ? rho : INFINITY, 0.0);
    const double s = (ok && rho > 0.0) ? 1.0 / rho : 1.0;
    double f = 1.0;
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) {
        const double2 v = ok ? a[k] : make_double2(0.0, 0.0);
        smom[i * kStereoTerms + k] = v;
        smom32[i * kStereoTerms + k] = make_float2(float(v.x * f), float(v.y * f));
        f *= s;
    }
}

// Leaves; thread = (projection, node); terms k < T (the others stay 0).
__global__ __launch_bounds__(256) void stereo_leaf(int nn, int T, const int* __restrict__ nkids, const int* __restrict__ nbeg,
                                                   const int* __restrict__ nend, const double4* __restrict__ ncent,
                                                   const double4* __restrict__ spos, double4* __restrict__ sgeo,
                                                   double2* __restrict__ smom, float2* __restrict__ smom32) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2 * int64_t(nn)) return;
    const int proj = int(i / nn), n = int(i % nn);
    if (nkids[n] != 0) return;
    double cr, ci;
    bool ok = node_centre(ncent[n], proj, cr, ci);
    double2 a[kStereoTerms];
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) a[k] = make_double2(0.0, 0.0);
    double rho2 = 0.0;
    const int e = nend[n];
    for (int j = nbeg[n]; ok && j < e; ++j) {
        const double4 s = spos[j];
        double zr, zi;
        ok = stereo_point(proj, s.x, s.y, s.z, zr, zi);
        const double dr = zr - cr, di = zi - ci;
        rho2 = fmax(rho2, dr * dr + di * di);
        double pr = s.w, pim = 0.0;  // g (z - z_c)^k
#pragma unroll
        for (int k = 0; k < kStereoTerms; ++k) {
            if (k < T) {
                a[k].x += pr;
                a[k].y += pim;
                const double t = pr * dr - pim * di;
                pim = pr * di + pim * dr;
                pr = t;
            }
        }
    }
    store_node(i, ok, cr, ci, sqrt(rho2), a, sgeo, smom, smom32);
}

// Internal nodes of one level (first node f, cnt nodes); thread = (projection, node).
// With d = z_c' - z_c, a_k = sum_l C(k, l) a'_l d^(k-l): the in-place
// transform b_k += d b_(k-1) for k = K-1 .. j, j = 1 .. K-1.
__global__ __launch_bounds__(256) void stereo_m2m(int f, int cnt, int nn, int T, const int* __restrict__ nfirst,
                                                  const int* __restrict__ nkids, const double4* __restrict__ ncent,
                                                  double4* __restrict__ sgeo, double2* __restrict__ smom,
                                                  float2* __restrict__ smom32) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2 * int64_t(cnt)) return;
    const int proj = int(i / cnt), n = f + int(i % cnt);
    const int nk = nkids[n];
    if (nk == 0) return;
    double cr, ci;
    bool ok = node_centre(ncent[n], proj, cr, ci);
    double2 a[kStereoTerms];
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) a[k] = make_double2(0.0, 0.0);
    double rho = 0.0;
    const int fc = nfirst[n];
    for (int q = 0; ok && q < nk; ++q) {
        const int64_t ci_ = int64_t(proj) * nn + fc + q;
        const double4 g = sgeo[ci_];
        ok = g.z < INFINITY;
        const double dr = g.x - cr, di = g.y - ci;
        rho = fmax(rho, sqrt(dr * dr + di * di) + g.z);
        double2 b[kStereoTerms];
#pragma unroll
        for (int k = 0; k < kStereoTerms; ++k) b[k] = smom[ci_ * kStereoTerms + k];
#pragma unroll
        for (int j = 1; j < kStereoTerms; ++j)
#pragma unroll
            for (int k = kStereoTerms - 1; k >= j; --k) {
                if (k < T) {
                    b[k].x += dr * b[k - 1].x - di * b[k - 1].y;
                    b[k].y += dr * b[k - 1].y + di * b[k - 1].x;
                }
            }
#pragma unroll
        for (int k = 0; k < kStereoTerms; ++k) {
            a[k].x += b[k].x;
            a[k].y += b[k].y;
        }
    }
    store_node(int64_t(proj) * nn + n, ok, cr, ci, rho, a, sgeo, smom, smom32);
}

}  // namespace

int tree_stereo_moments(lpm_ctx* c, int p, hipStream_t st) {
    Tree* t = c->tree;
    // an accepted pair of Cartesian order q <= p becomes a series of order
    // q + extra: terms up to p + extra are needed
    const int T = std::min(kStereoTerms, p + c->stereo_extra + 1);
    const int64_t nn = t->nn;
    LPM_CHECK(t->sgeo.reserve(size_t(2 * nn) * sizeof(double4)));
    LPM_CHECK(t->smom.reserve(size_t(2 * nn) * kStereoTerms * sizeof(double2)));
    LPM_CHECK(t->smom32.reserve(size_t(2 * nn) * kStereoTerms * sizeof(float2)));
    double4* g = t->sgeo.as<double4>();
    double2* m = t->smom.as<double2>();
    float2* m32 = t->smom32.as<float2>();
    hipLaunchKernelGGL(stereo_leaf, dim3(blocks(2 * nn)), dim3(256), 0, st, int(nn), T, t->nkids.as<int>(),
                       t->nbeg.as<int>(), t->nend.as<int>(), t->ncent.as<double4>(), t->spos.as<double4>(), g, m, m32);
    for (int l = int(t->level_start.size()) - 3; l >= 0; --l) {
        const int64_t f = t->level_start[l], cnt = t->level_start[l + 1] - f;
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(stereo_m2m, dim3(blocks(2 * cnt)), dim3(256), 0, st, int(f), int(cnt), int(nn), T,
                           t->nfirst.as<int>(), t->nkids.as<int>(), t->ncent.as<double4>(), g, m, m32);
    }
    LPM_HIP(c, hipGetLastError());
    t->stereo_terms = T;
    return LPM_OK;
}

}  // namespace lpm
