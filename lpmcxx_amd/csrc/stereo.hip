// Stereographic node moments for the learned FP32C evaluation
// (lpm_stereo.hpp): for each node and projection the centre z_c (the
// projection of the node's box centroid direction), the radius rho of its
// sources about z_c, and a_k = sum_j g_j (z_j - z_c)^k, k < kStereoTerms,
// in FP64 (unscaled) and as FP32 a_k / rho^k.  Leaves sum their sources
// (16 lanes per leaf); internal nodes shift their children's moments (a
// binomial transform, one lane per child) level by level.  A node whose centre or one of whose sources lies within 1e-12 of
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

// Sum / max / and over the GS-lane group of this lane (xor butterfly).
template <int GS>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int o = 1; o < GS; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <int GS>
__device__ inline double group_max(double v) {
#pragma unroll
    for (int o = 1; o < GS; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
template <int GS>
__device__ inline bool group_all(bool b) {
    const uint64_t m = __ballot(!b);
    const int g0 = (threadIdx.x & 63) & ~(GS - 1);
    return ((m >> g0) & ((GS == 64) ? ~0ull : ((1ull << GS) - 1ull))) == 0ull;
}

// Leaves: kLeafLanes lanes per (leaf, projection), each summing every
// kLeafLanes-th source of the leaf, the lane sums then added by a butterfly;
// terms k < T (the others stay 0).
constexpr int kLeafLanes = 16;
__global__ __launch_bounds__(256) void stereo_leaf(int nn, int T, const int* __restrict__ nkids, const int* __restrict__ nbeg,
                                                   const int* __restrict__ nend, const double4* __restrict__ ncent,
                                                   const double4* __restrict__ spos, double4* __restrict__ sgeo,
                                                   double2* __restrict__ smom, float2* __restrict__ smom32) {
    const int64_t g = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kLeafLanes;
    const int l = threadIdx.x & (kLeafLanes - 1);
    if (g >= 2 * int64_t(nn)) return;  // whole lane groups
    const int proj = int(g / nn), n = int(g % nn);
    if (nkids[n] != 0) return;
    double cr, ci;
    bool ok = node_centre(ncent[n], proj, cr, ci);
    double2 a[kStereoTerms];
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) a[k] = make_double2(0.0, 0.0);
    double rho2 = 0.0;
    const int e = nend[n];
    // every source is summed (a pole point only spoils the result, which is
    // then discarded), so the loads of successive sources do not wait for
    // each other's pole test
    for (int j = nbeg[n] + l; j < e; j += kLeafLanes) {
        const double4 s = spos[j];
        double zr, zi;
        ok = stereo_point(proj, s.x, s.y, s.z, zr, zi) && ok;
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
    ok = group_all<kLeafLanes>(ok);
    rho2 = group_max<kLeafLanes>(rho2);
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) {
        if (k < T) {
            a[k].x = group_sum<kLeafLanes>(a[k].x);
            a[k].y = group_sum<kLeafLanes>(a[k].y);
        }
    }
    if (l == 0) store_node(g, ok, cr, ci, sqrt(rho2), a, sgeo, smom, smom32);
}

// Internal nodes of one level (first node f, cnt nodes): 8 lanes per
// (projection, node), lane q shifting child q.  With d = z_c' - z_c,
// a_k = sum_l C(k, l) a'_l d^(k-l): the in-place transform
// b_k += d b_(k-1) for k = K-1 .. j, j = 1 .. K-1; the children's shifted
// moments are then added by a butterfly over the 8 lanes.
__global__ __launch_bounds__(256) void stereo_m2m(int f, int cnt, int nn, int T, const int* __restrict__ nfirst,
                                                  const int* __restrict__ nkids, const double4* __restrict__ ncent,
                                                  double4* __restrict__ sgeo, double2* __restrict__ smom,
                                                  float2* __restrict__ smom32) {
    const int64_t g = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3;
    const int q = threadIdx.x & 7;
    if (g >= 2 * int64_t(cnt)) return;  // whole lane groups
    const int proj = int(g / cnt), n = f + int(g % cnt);
    const int nk = nkids[n];
    if (nk == 0) return;
    double cr, ci;
    bool ok = node_centre(ncent[n], proj, cr, ci);
    double2 b[kStereoTerms];
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) b[k] = make_double2(0.0, 0.0);
    double rho = 0.0;
    if (q < nk) {
        const int64_t ci_ = int64_t(proj) * nn + nfirst[n] + q;
        const double4 gz = sgeo[ci_];
        ok = ok && gz.z < INFINITY;
        const double dr = gz.x - cr, di = gz.y - ci;
        rho = sqrt(dr * dr + di * di) + gz.z;
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
    }
    ok = group_all<8>(ok);
    rho = group_max<8>(rho);
#pragma unroll
    for (int k = 0; k < kStereoTerms; ++k) {
        if (k < T) {
            b[k].x = group_sum<8>(b[k].x);
            b[k].y = group_sum<8>(b[k].y);
        }
    }
    if (q == 0) store_node(int64_t(proj) * nn + n, ok, cr, ci, rho, b, sgeo, smom, smom32);
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
    hipLaunchKernelGGL(stereo_leaf, dim3(blocks(2 * nn * kLeafLanes)), dim3(256), 0, st, int(nn), T, t->nkids.as<int>(),
                       t->nbeg.as<int>(), t->nend.as<int>(), t->ncent.as<double4>(), t->spos.as<double4>(), g, m, m32);
    for (int l = int(t->level_start.size()) - 3; l >= 0; --l) {
        const int64_t f = t->level_start[l], cnt = t->level_start[l + 1] - f;
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(stereo_m2m, dim3(blocks(2 * cnt * 8)), dim3(256), 0, st, int(f), int(cnt), int(nn), T,
                           t->nfirst.as<int>(), t->nkids.as<int>(), t->ncent.as<double4>(), g, m, m32);
    }
    LPM_HIP(c, hipGetLastError());
    t->stereo_terms = T;
    t->stereo_order = p;
    return LPM_OK;
}

}  // namespace lpm
