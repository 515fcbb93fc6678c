// Batch-MAC treecode and the learned order / acceptance (the TreeSum
// evaluation of src/LpmTreeSum.cpp:232-311, reorganised for wave64).
//
// A target batch is 64 consecutive Morton-sorted targets, one wave.  The wave
// walks the node table depth-first with a wave-uniform stack, and every
// decision is taken for the whole batch (c_b, r_b: centroid and radius of the
// batch's targets; c_n, r_n: centroid and maxRadius of the node):
//   plain    accept when r_n + r_b < theta |c_b - c_n|; every lane then
//            evaluates the order-p Cartesian far field.  Otherwise descend,
//            and at a leaf every lane sums the leaf's particles, which all
//            lanes read at one address (scalar loads, one broadcast per
//            source).  Because no lane idles, the far and near loops run at
//            full SIMD width, unlike the per-target MAC of the reference
//            (which tree_eval keeps for parity).
//   learned  pairs with r_n + r_b < theta_max |c_b - c_n| go through the order
//            model on the MFMA units (mlp_classify_regs, all children of a
//            node in one pass): class 0 descends, class q <= max_order
//            accepts with order min(q + margin, max_order).
//   labels   for a sample of batches the wave visits every node; for a sample
//            of the pairs that pass theta_max it records the ten features and
//            the error of the order-1..12 far field against the exact sum over
//            the node's particles (max over the batch, in units of u_ref).
// Precision: LPM_F64 evaluates everything in FP64 (reference pair formula in
// the near field).  LPM_F32C evaluates the far field in FP32 from FP32
// moments (graded order) with d = t - c_n rounded once, and the near field in
// FP32 from offsets to the leaf centroid with |d|^2 / 2 for 1 - x.y; the
// per-node results are added to FP64 per-target sums.
#include "lpm_mlp.hpp"
#include "lpm_series.hpp"
#include "lpm_stereo.hpp"
#include "lpm_tree.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <utility>

namespace lpm {
namespace {

constexpr int kStackB = 160;  // >= 7 * kMaxDepth + 1
// Planning-stack entries per wave (the bound for a given number of rows per
// model call is plan_stack_bound): with the model, the rows' inputs and the
// row map, a block of four waves stays within a fifth of the CU's LDS.
constexpr int kStackPlan = 640;
constexpr int kLabelOrder = 12;
constexpr int kMaxLearnedOrder = 12;

__device__ inline double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Global index of this wave (4 waves per block), marked wave-uniform so that
// loads addressed by it become scalar loads.
__device__ inline int64_t wave_index() {
    return int64_t(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
}

// true in every lane when cond holds in lane 0 (cond is wave-uniform by construction)
__device__ inline bool uniform(bool cond) { return (__ballot(cond) & 1ull) != 0; }

struct Batch {
    double cx, cy, cz, r;
    int nvalid;
};

// Centroid and radius of the valid targets of the wave.  The butterfly
// reductions leave identical values in every lane.
__device__ inline Batch batch_of(bool valid, double t0, double t1, double t2) {
    Batch b;
    b.nvalid = __popcll(__ballot(valid));
    const double inv = 1.0 / double(b.nvalid);
    b.cx = wsum(valid ? t0 : 0.0) * inv;
    b.cy = wsum(valid ? t1 : 0.0) * inv;
    b.cz = wsum(valid ? t2 : 0.0) * inv;
    const double dx = t0 - b.cx, dy = t1 - b.cy, dz = t2 - b.cz;
    b.r = sqrt(wmax(valid ? dx * dx + dy * dy + dz * dz : 0.0));
    return b;
}

// Ten features of a (batch, node) pair:
//   log2(r_n/s), log2(r_b/s), (r_n + r_b)/s, log2(M0/(u_ref s)),
//   log2(M1/(M0 r_n)), log2(M2/(M0 r_n^2)), log2(npart), nvalid/64, log2 s,
//   leaf flag, with s = |c_b - c_n| and M0, M1, M2 the moment shell norms.
// FP32 with the hardware log2 and reciprocal: the label pass and the
// evaluation call this same function, so training and inference see
// identical inputs.
__device__ inline float flog2(float v) {
    return v > 0.f ? fmaxf(__builtin_amdgcn_logf(v), -100.f) : -100.f;
}
__device__ inline void pair_features(float* f, double rn, double rb, double s, float4 st, bool leaf, int nvalid,
                                     double u_ref) {
    const float sf = float(s);
    const float is = __builtin_amdgcn_rcpf(sf);
    const float rnf = float(fmax(rn, 1e-12 * s)), rbf = float(fmax(rb, 1e-12 * s));
    const float m0 = st.x, m1 = st.y, m2 = st.z;
    f[0] = flog2(rnf * is);
    f[1] = flog2(rbf * is);
    f[2] = (rnf + rbf) * is;
    f[3] = flog2(m0 * is * __builtin_amdgcn_rcpf(float(u_ref)));
    const float mr = m0 * rnf;
    f[4] = mr > 0.f ? flog2(m1 * __builtin_amdgcn_rcpf(mr)) : -100.f;
    f[5] = mr > 0.f ? flog2(m2 * __builtin_amdgcn_rcpf(mr * rnf)) : -100.f;
    f[6] = flog2(st.w);
    f[7] = float(nvalid) * (1.f / 64.f);
    f[8] = flog2(sf);
    f[9] = leaf ? 1.f : 0.f;
#pragma unroll
    for (int k = kFeatures; k < kIn; ++k) f[k] = 0.f;
}

template <typename T, int... Ps>
__device__ __forceinline__ void far_switch(std::integer_sequence<int, Ps...>, int p, T d0, T d1, T d2, T dd,
                                           const T* __restrict__ mm, T& S0, T& S1, T& S2) {
    ((p == Ps + 1 ? (far_cartesian<Ps + 1>(d0, d1, d2, dd, mm, S0, S1, S2), 0) : 0), ...);
}

template <int... Ps>
__device__ __forceinline__ void far_switch_scaled(std::integer_sequence<int, Ps...>, int p, float d0, float d1,
                                                  float d2, float dd, float rho, const float* __restrict__ mm,
                                                  float& S0, float& S1, float& S2) {
    ((p == Ps + 1 ? (far_cartesian_scaled<Ps + 1>(d0, d1, d2, dd, rho, mm, S0, S1, S2), 0) : 0), ...);
}

// s += t x v
__device__ inline void cross_add(double t0, double t1, double t2, double v0, double v1, double v2, double& s0,
                                 double& s1, double& s2) {
    s0 += t1 * v2 - t2 * v1;
    s1 += t2 * v0 - t0 * v2;
    s2 += t0 * v1 - t1 * v0;
}

// Far field of node n at every lane's target.  FP64: t x S added to (s0, s1,
// s2).  FP32: S added to (v0, v1, v2); t x v is formed once per target at the
// end (t x sum S = sum t x S).
template <int P, bool F32, bool RUNTIME_P>
__device__ __forceinline__ void far_node(int p, int n, int nterm, double t0, double t1, double t2, const double4& c,
                                         double d2, const double* __restrict__ momg,
                                         const float* __restrict__ mom32, double& s0, double& s1, double& s2,
                                         double& v0, double& v1, double& v2) {
    double S0, S1, S2;
    if constexpr (F32) {
        float A0 = 0.f, A1 = 0.f, A2 = 0.f;
        const float e0 = float(t0 - c.x), e1 = float(t1 - c.y), e2 = float(t2 - c.z), dd = float(2.0 * d2);
        const float rho = float(c.w);  // box radius: the moments' scale
        const float* mm = mom32 + int64_t(n) * nterm * 3;
        if constexpr (RUNTIME_P)
            far_switch_scaled(std::make_integer_sequence<int, P>{}, p, e0, e1, e2, dd, rho, mm, A0, A1, A2);
        else far_cartesian_scaled<P>(e0, e1, e2, dd, rho, mm, A0, A1, A2);
        v0 += A0;
        v1 += A1;
        v2 += A2;
        return;
    } else {
        S0 = S1 = S2 = 0.0;
        const double* mm = momg + int64_t(n) * nterm * 3;
        if constexpr (RUNTIME_P)
            far_switch<double>(std::make_integer_sequence<int, P>{}, p, t0 - c.x, t1 - c.y, t2 - c.z, 2.0 * d2, mm, S0,
                               S1, S2);
        else far_cartesian<P>(t0 - c.x, t1 - c.y, t2 - c.z, 2.0 * d2, mm, S0, S1, S2);
    }
    s0 += t1 * S2 - t2 * S1;
    s1 += t2 * S0 - t0 * S2;
    s2 += t0 * S1 - t1 * S0;
}

// Leaf [b, e) summed directly at every lane's target; the sources are
// wave-uniform loads.  CHECK = false (FP32 only): no target of the batch
// comes within 1e-6 of the leaf's sources, so no pair meets the zero rule
// and its test is left out.
template <bool F32, bool CHECK = true>
__device__ __forceinline__ void near_leaf(int b, int e, double t0, double t1, double t2, const double4& c, double d2,
                                          float eps2v, const double4* __restrict__ spos,
                                          const float4* __restrict__ soff, double& s0, double& s1, double& s2,
                                          double& v0, double& v1, double& v2) {
    if constexpr (F32) {
        const float q0 = float(t0 - c.x), q1 = float(t1 - c.y), q2 = float(t2 - c.z);
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        auto pair = [&](const float4& o) {
            const float dx = q0 - o.x, dy = q1 - o.y, dz = q2 - o.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2v)));
            // the reference's |den| <= 1e-13 rule: |d|^2 + 2 delta^2 <= 2e-13 contributes 0
            const float wgt = (!CHECK || r2 > 2e-13f) ? o.w * __builtin_amdgcn_rcpf(r2) : 0.f;
            a0 = fmaf(wgt, dx, a0);
            a1 = fmaf(wgt, dy, a1);
            a2 = fmaf(wgt, dz, a2);
        };
        // four sources a step, the next four requested before the current
        // four are summed
        const float4* src = soff + b;
        const int n = e - b;
        int j = 0;
        if (n >= 8) {
            float4 c0 = src[0], c1 = src[1], c2 = src[2], c3 = src[3];
            for (j = 4; j + 4 <= n; j += 4) {
                const float4 n0 = src[j], n1 = src[j + 1], n2 = src[j + 2], n3 = src[j + 3];
                pair(c0);
                pair(c1);
                pair(c2);
                pair(c3);
                c0 = n0;
                c1 = n1;
                c2 = n2;
                c3 = n3;
            }
            pair(c0);
            pair(c1);
            pair(c2);
            pair(c3);
        }
        for (; j < n; ++j) pair(src[j]);
        // (t x y) gamma / den = -2 t x d gamma / (|d|^2 + 2 delta^2): -2 d
        // gamma / den joins v, crossed with t once per target
        v0 -= 2.0 * double(a0);
        v1 -= 2.0 * double(a1);
        v2 -= 2.0 * double(a2);
    } else {
        for (int j = b; j < e; ++j) {
            const double4 s = spos[j];
            const double cx = t1 * s.z - t2 * s.y;
            const double cy = t2 * s.x - t0 * s.z;
            const double cz = t0 * s.y - t1 * s.x;
            const double den = 1.0 - (t0 * s.x + t1 * s.y + t2 * s.z) + d2;
            if (fabs(den) > 1.0e-13) {
                const double inv = 1.0 / den;
                s0 += (cx * inv) * s.w;
                s1 += (cy * inv) * s.w;
                s2 += (cz * inv) * s.w;
            }
        }
    }
}

struct TreeView {
    const double4* ncent;
    const double* nrad;
    const int *nfirst, *nkids, *nbeg, *nend;
    const int4* nodew;  // the four words above, packed
    const float4* nstat;
    const double4* spos;
    const float4* soff;
    const double* momg;
    const float* mom32;
    int nterm;
};

struct Targets {
    const double *x, *y, *z;
    const int* perm;    // sorted position -> target index
    int64_t n;
    const int* bstart;  // first sorted position of every batch, then n
    int64_t nb;         // bound on the batch count (launch size)
    const int* nbd;     // batch count (device)
};

// Whether wave index b is a batch (wave-uniform).
__device__ inline bool is_batch(const Targets& tg, int64_t b) { return b < tg.nb && b < *tg.nbd; }

// Target of this lane in batch b; pos = its sorted position.
__device__ inline int64_t load_target(const Targets& tg, int64_t b, bool& valid, int& orig, double& t0, double& t1,
                                      double& t2) {
    const int64_t b0 = tg.bstart[b], b1 = tg.bstart[b + 1];
    const int64_t it = b0 + (threadIdx.x & 63);
    valid = it < b1;
    orig = valid ? tg.perm[it] : 0;
    t0 = valid ? tg.x[orig] : 0.0;
    t1 = valid ? tg.y[orig] : 0.0;
    t2 = valid ? tg.z[orig] : 0.0;
    return it;
}

__device__ inline void flush_stats(unsigned long long* stats, unsigned long long far, unsigned long long near,
                                   unsigned long long terms, unsigned long long calls,
                                   unsigned long long far_hi = 0, unsigned long long terms_hi = 0,
                                   unsigned long long far_st = 0, unsigned long long terms_st = 0) {
    if ((threadIdx.x & 63) == 0) {
        stats += (blockIdx.x % kStatSlots) * kStatStride;
        atomicAdd(stats + 0, far);
        atomicAdd(stats + 1, near);
        atomicAdd(stats + 2, terms);
        atomicAdd(stats + 3, calls);
        if (far_hi) atomicAdd(stats + 4, far_hi);
        if (terms_hi) atomicAdd(stats + 5, terms_hi);
        if (far_st) atomicAdd(stats + 6, far_st);
        if (terms_st) atomicAdd(stats + 7, terms_st);
    }
}

// ---------------------------------------------------------------- stereographic far field
// Node data of both projections and the switch of the learned FP32C
// evaluation (lpm_stereo.hpp).
struct StereoView {
    const double4* sgeo;   // [2][nn] centre (re, im), rho
    const float2* smom32;  // [2][nn][kStereoTerms] a_k / rho^k
    int64_t nn;
    int on;                // evaluate accepted pairs as plane series where they qualify
    int extra;             // order P = p + extra
    int maxp;              // highest P the moments allow
    double ratio;          // plane ratio must be <= ratio x the sphere ratio
    int cart_max;          // highest order of a Cartesian entry (the Cartesian moments')
};

// The batch's projection and plane centre / radius; ok is false when a target
// lies at the projection's pole (the batch then keeps the Cartesian series).
struct StereoBatch {
    int proj;
    bool ok;
    double zr, zi;    // this lane's target
    double cr, ci, r; // batch centre and radius in the plane
};

// a wave-uniform double held in scalar registers
__device__ inline double uniform_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffffll)), hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// The batch in projection proj (wave-uniform).
__device__ inline StereoBatch stereo_batch_in(int proj, const Batch& bt, bool valid, double t0, double t1, double t2) {
    StereoBatch s;
    s.proj = proj;
    const bool good = stereo_point(s.proj, t0, t1, t2, s.zr, s.zi) || !valid;
    s.ok = __ballot(!good) == 0ull;
    const double inv = 1.0 / double(bt.nvalid);
    s.cr = uniform_d(wsum(valid ? s.zr : 0.0) * inv);
    s.ci = uniform_d(wsum(valid ? s.zi : 0.0) * inv);
    const double er = s.zr - s.cr, ei = s.zi - s.ci;
    s.r = uniform_d(sqrt(wmax(valid ? er * er + ei * ei : 0.0)));
    return s;
}

// The batch in its own projection: from the pole opposite its centre's
// hemisphere, which keeps its targets within 90 degrees of the projection
// centre.  A pair that fails the geometry test there may pass it in the other
// projection (a node near the far pole is distorted in the first, near the
// centre of the second).
__device__ inline StereoBatch stereo_batch(const Batch& bt, bool valid, double t0, double t1, double t2) {
    return stereo_batch_in(__builtin_amdgcn_readfirstlane(bt.cz >= 0.0 ? 0 : 1), bt, valid, t0, t1, t2);
}

// Stereographic geometry test of the pair (node n, batch sb): the plane ratio
// no worse (by sv.ratio) than the sphere ratio (r_n + r_b) / s, and the plane
// series convergent (radius below 0.95 of the distance); squared, so no FP64
// division or square root (both sides are >= 0; rho = +inf fails both).
__device__ inline bool stereo_fits(const StereoView& sv, const StereoBatch& sb, int n, double sd, double rn,
                                   double rb) {
    const double4 gz = sv.sgeo[int64_t(sb.proj) * sv.nn + n];
    const double er = sb.cr - gz.x, ei = sb.ci - gz.y;
    const double dist2 = er * er + ei * ei;
    const double rz = gz.z + sb.r;
    const double lhs = rz * sd, rhs = sv.ratio * (rn + rb);
    return lhs * lhs < rhs * rhs * dist2 && rz * rz < 0.9025 * dist2;
}

// Node data of one stereographic entry: plane centre and radius (FP64) and
// the first kStPre terms a_k / rho^k.  gi is wave-uniform, so these are
// scalar loads; the list kernel requests the next entry's data before it
// sums the current one.
// Terms of a stereographic entry fetched ahead of its sum (orders P < kStPre;
// higher orders load the rest when they are reached).
constexpr int kStPre = 10;
static_assert(kStPre == 10 && kStPre <= kStereoTerms, "stereo_sum unrolls terms 9 .. 0");

struct StNode {
    double cr, ci, rho;
    float2 a[kStPre];
};

__device__ __forceinline__ StNode stereo_fetch(const StereoView& sv, int64_t gi) {
    StNode s;
    const double4 g = sv.sgeo[gi];
    s.cr = g.x;
    s.ci = g.y;
    s.rho = g.z;
    const float2* a = sv.smom32 + gi * kStereoTerms;
#pragma unroll
    for (int k = 0; k < kStPre; ++k) s.a[k] = a[k];
    return s;
}

// Cauchy series of node entry gi at the lane's target from its fetched data:
// S += w sum_k (a_k / rho^k) (rho w)^k, w = 1 / (z - z_c), by Horner in FP32
// entering the unrolled sum at term P; G += the node's total strength.
// Orders P >= kStPre read their remaining terms here.
__device__ __forceinline__ void stereo_sum(const StNode& nd, int P, int64_t gi, const StereoView& sv, double zr,
                                           double zi, float& Sr, float& Si, float& G) {
    const float dr = float(zr - nd.cr), di = float(zi - nd.ci);
    const float ir2 = __builtin_amdgcn_rcpf(fmaf(dr, dr, di * di));
    const float wr = dr * ir2, wi = -di * ir2;
    const float rho = float(nd.rho);
    const float xr = rho * wr, xi = rho * wi;
    float sr = 0.f, si = 0.f;
    if (P >= kStPre) {
        const float2* a = sv.smom32 + gi * kStereoTerms;
        for (int k = P; k >= kStPre; --k) {
            const float2 ak = a[k];
            const float tr = fmaf(sr, xr, fmaf(-si, xi, ak.x));
            si = fmaf(sr, xi, fmaf(si, xr, ak.y));
            sr = tr;
        }
    }
#define LPM_HORNER(K)                                                \
    case K: {                                                        \
        const float tr = fmaf(sr, xr, fmaf(-si, xi, nd.a[K].x));     \
        si = fmaf(sr, xi, fmaf(si, xr, nd.a[K].y));                  \
        sr = tr;                                                     \
    }                                                                \
        [[fallthrough]];
    switch (P < kStPre ? P : kStPre - 1) {
        LPM_HORNER(9) LPM_HORNER(8) LPM_HORNER(7) LPM_HORNER(6) LPM_HORNER(5) LPM_HORNER(4)
        LPM_HORNER(3) LPM_HORNER(2) LPM_HORNER(1) LPM_HORNER(0)
        default: break;
    }
#undef LPM_HORNER
    Sr += fmaf(sr, wr, -si * wi);
    Si += fmaf(sr, wi, si * wr);
    G += nd.a[0].x;
}

// The unscaled velocity sum sum g (t x y) / (1 - t.y) of the plane series
// (S, G) at the target z = zr + i zi, added to (s0, s1, s2).
__device__ inline void stereo_add(int proj, double zr, double zi, double Sr, double Si, double G, double& s0,
                                  double& s1, double& s2) {
    const double q = 1.0 + zr * zr + zi * zi;
    const double wr = Sr - G * zr / q, wi = Si + G * zi / q;  // S - G conj(z) / (1 + |z|^2)
    const double u0 = -(wi * (1.0 + zi * zi - zr * zr) - 2.0 * wr * zr * zi);
    const double u1 = -(wr * (1.0 + zr * zr - zi * zi) - 2.0 * wi * zr * zi);
    const double u2 = 2.0 * (wi * zr + wr * zi);
    s0 += u0;
    s1 += proj == 0 ? u1 : -u1;
    s2 += proj == 0 ? u2 : -u2;
}

// ---------------------------------------------------------------- plain batch treecode
template <int P, bool F32>
__global__ __launch_bounds__(256) void tree_eval_batch(Targets tg, TreeView tv, double theta, double d2, float eps2v,
                                                       double mult, double* __restrict__ ux, double* __restrict__ uy,
                                                       double* __restrict__ uz,
                                                       unsigned long long* __restrict__ stats) {
    __shared__ int s_node[4][kStackB];
    const int w = threadIdx.x >> 6;
    const int64_t bi = wave_index();
    if (!is_batch(tg, bi)) return;
    bool valid;
    int orig;
    double t0, t1, t2;
    load_target(tg, bi, valid, orig, t0, t1, t2);
    const Batch bt = batch_of(valid, t0, t1, t2);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0;
    unsigned long long nfar = 0, nnear = 0;
    constexpr unsigned long long NT = unsigned(P * (P + 1) * (P + 2) / 6);
    int sp = 1;
    s_node[w][0] = 0;
    while (sp > 0) {
        const int n = __builtin_amdgcn_readfirstlane(s_node[w][--sp]);
        const double4 c = tv.ncent[n];
        const double dx = bt.cx - c.x, dy = bt.cy - c.y, dz = bt.cz - c.z;
        const double s = sqrt(dx * dx + dy * dy + dz * dz);
        const double rn = tv.nrad[n];
        if (uniform(rn + bt.r < theta * s)) {
            far_node<P, F32, false>(P, n, tv.nterm, t0, t1, t2, c, d2, tv.momg, tv.mom32, s0, s1, s2, v0, v1, v2);
            ++nfar;
            continue;
        }
        const int nk = tv.nkids[n];
        if (nk == 0) {
            const int b = tv.nbeg[n], e = tv.nend[n];
            // FP32: the zero-rule test only when some target of the wave lies
            // within r_n + 1e-6 of the leaf centre (FP32 offsets: 2e-6)
            bool chk = true;
            if (F32) {
                const float q0 = float(t0 - c.x), q1 = float(t1 - c.y), q2 = float(t2 - c.z);
                const float lim = float(rn) + 2e-6f;
                chk = __ballot(valid && fmaf(q0, q0, fmaf(q1, q1, q2 * q2)) <= lim * lim) != 0ull;
            }
            if (chk)
                near_leaf<F32, true>(b, e, t0, t1, t2, c, d2, eps2v, tv.spos, tv.soff, s0, s1, s2, v0, v1, v2);
            else
                near_leaf<F32, false>(b, e, t0, t1, t2, c, d2, eps2v, tv.spos, tv.soff, s0, s1, s2, v0, v1, v2);
            nnear += unsigned(e - b);
        } else {
            const int f = tv.nfirst[n];
            for (int q = nk - 1; q >= 0; --q) s_node[w][sp++] = f + q;
        }
    }
    cross_add(t0, t1, t2, v0, v1, v2, s0, s1, s2);
    if (valid) {
        ux[orig] = s0 * mult;
        uy[orig] = s1 * mult;
        uz[orig] = s2 * mult;
    }
    const unsigned long long nv = unsigned(bt.nvalid);
    flush_stats(stats, nfar * nv, nnear * nv, nfar * nv * NT, 0ull);
}

// ---------------------------------------------------------------- learned batch treecode
// Two passes.  tree_plan_learned walks the tree per batch and decides every
// pair with the order model; the result is an interaction list per batch:
// far entries (node | p << 27) and near leaves.  When a wave descends into
// node n, the children of n are classified together: lane k < nkids computes
// the features of child k, the rows go through the model in one pass, and
// each child is pushed with its class (0 = descend, also for children failing
// theta_max).  tree_eval_lists then evaluates each batch's list (near
// leaves and low-order far entries in one launch, high orders in a second)
// without the model's registers and LDS, and tree_list_finish scales and
// scatters.
constexpr int kNodeBits = 27;
// near-list entry: leaf | kNearCheck when a target of the batch may come
// within 1e-6 of one of the leaf's sources (the zero rule must be tested)
constexpr int kNearCheck = 1 << 30;

// planning-stack entry of an internal node: first child | child count << kNodeBits
__device__ inline int stack_entry(int first, int kids) { return first | (kids << kNodeBits); }

// Far entries of order <= kLowOrder fill a batch's far list from the front,
// higher orders from the back, so two evaluation kernels with different
// register footprints each read a contiguous range.
constexpr int kLowOrder = 7;

struct Lists {
    int* far;
    int* near;
    int* st;          // stereographic entries (node | P << kNodeBits)
    int* far_cnt;     // low-order far entries (front)
    int* far_hi_cnt;  // high-order far entries (back)
    int* near_cnt;
    int* st_cnt;      // stereographic entries in the batch's own projection (front)
    int* st_o_cnt;    // stereographic entries in the other projection (back)
    int far_cap, near_cap, st_cap;
    int* overflow;  // [0] a batch overflowed, [1] far, [2] near and [3] stereographic capacity it needs
    // the batches by estimated cost (vector instructions of their lists):
    // bins[k] batches in bin k, listed at queue[k * qcap ...]
    unsigned* bins;
    int* queue;
    int64_t qcap;
};

// Vector instructions per near pair, Cartesian term, Cartesian entry,
// stereographic entry and complex term: the list kernels' costs, used to run
// the dearest batches first so that the cheap ones fill the end of the launch.
constexpr unsigned kCostNear = 14, kCostTerm = 5, kCostFar = 35, kCostSt = 24, kCostStTerm = 4;

// The waves of the list kernels take the batches dearest first (an LPT
// order): the planning pass files each batch under one of 32 bins of
// log2(its cost estimate), and wave w of a list kernel takes the w-th batch
// counting from the dearest bin.  Within a bin the order is that of the
// filing atomics: it only decides which wave evaluates which batch, not any
// result.
constexpr int kCostBins = 32;
static_assert(kCostBins <= 32, "bins of the batch header");
__device__ inline int cost_bin(unsigned c) { return c == 0u ? kCostBins - 1 : min(kCostBins - 1, int(__clz(c))); }

// Batch of wave w (wave-uniform), or -1 past the last filed batch.
__device__ inline int64_t list_batch(const Lists& ls, int64_t w) {
    const int lane = threadIdx.x & 63;
    const unsigned c = lane < kCostBins ? ls.bins[lane] : 0u;
    unsigned incl = c;  // inclusive sum over the bins up to this lane's
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    // the bins that end at or before w are the first b (sums are non-decreasing)
    const int b = __popcll(__ballot(int64_t(incl) <= w));
    if (b >= kCostBins) return -1;
    const unsigned before = __shfl(incl - c, b, 64);
    return __builtin_amdgcn_readfirstlane(ls.queue[int64_t(b) * ls.qcap + (w - int64_t(before))]);
}

// Rows of one model call: the planning pass pops the top stack entries
// whose children fit in ROWS rows (at least one entry: a node has at most 8
// children) and pushes up to ROWS children.  Seen as segments (the children
// pushed by one call), every segment below the top has lost an entry (at
// most ROWS - 1 left) and the segments' smallest depths increase up the
// stack, so at most (ROWS - 1) d + ROWS entries are held for a tree of depth
// d.  The host takes the largest ROWS whose bound fits kStackPlan.
__host__ __device__ constexpr int plan_stack_bound(int rows, int depth) { return (rows - 1) * depth + rows; }
static_assert(plan_stack_bound(16, kMaxDepth) <= kStackPlan, "16 rows fit any tree");

// One lane per row: lane l computes the geometry and the ten features of
// row l's child, the rows go through the model in ROWS / 16 tiles of 16, and
// lane l files its child.  Four waves per SIMD with 64 or 32 rows (the
// register footprint of the model's row tiles; an occupancy bound spills),
// five with 16.  STEREO: accepted pairs may become stereographic entries (the
// batch's plane centre and radius in both projections are kept in scalar
// registers).
template <bool STEREO, int ROWS>
__global__ __launch_bounds__(256) void tree_plan_learned(Targets tg, TreeView tv, double theta_max, int margin,
                                                         int max_order, double u_ref,
                                                         const unsigned short* __restrict__ wts, MlpNorm norm,
                                                         Lists ls, StereoView sv,
                                                         unsigned long long* __restrict__ stats) {
    static_assert(ROWS == 16 || ROWS == 32 || ROWS == 64, "rows of a model call");
    constexpr int RT = ROWS / 16;
    __shared__ __align__(16) unsigned char s_model[kModelLds];
    __shared__ int s_node[4][kStackPlan];
    __shared__ int s_row[4][ROWS];  // child node of each row of a model call
    // the BF16 model inputs of each row, read back in the MFMA operand layout
    __shared__ __align__(16) unsigned s_feat[4][ROWS][kIn / 2];
    // the input normalization (mean, inv_std, lo, hi) of each input
    __shared__ float4 s_norm[kIn];
    mlp_stage_weights(wts, s_model);
    if (threadIdx.x < kIn)
        s_norm[threadIdx.x] = make_float4(norm.mean[threadIdx.x], norm.inv_std[threadIdx.x], norm.lo[threadIdx.x],
                                          norm.hi[threadIdx.x]);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
    const int64_t bi = wave_index();
    if (!is_batch(tg, bi)) return;
    bool valid;
    int orig;
    double t0, t1, t2;
    load_target(tg, bi, valid, orig, t0, t1, t2);
    const Batch bt = batch_of(valid, t0, t1, t2);
    StereoBatch sb{}, so{};
    if (STEREO) {
        sb = stereo_batch(bt, valid, t0, t1, t2);
        so = stereo_batch_in(1 - sb.proj, bt, valid, t0, t1, t2);
    }
    const bool st_on = STEREO && sb.ok, so_on = STEREO && so.ok;
    int* far = ls.far + bi * ls.far_cap;
    int* near = ls.near + bi * ls.near_cap;
    int* stl = ls.st + bi * ls.st_cap;
    int nfar = 0, nhi = 0, nnear = 0, nstl = 0, nso = 0;
    unsigned long long near_pairs = 0, nterms = 0, nterms_hi = 0, ncalls = 0, nst = 0, nterms_st = 0;
    unsigned l_terms = 0, l_terms_hi = 0, l_np = 0, l_terms_st = 0;
    const uint64_t below = (1ull << lane) - 1ull;
    // normalized inputs kFeatures .. kIn - 1 (always 0 before normalization)
    // as BF16 pairs, the same for every row
    unsigned pad[(kIn - kFeatures) / 2];
#pragma unroll
    for (int k = kFeatures; k < kIn; k += 2) {
        const float4 a = s_norm[k], b = s_norm[k + 1];
        pad[(k - kFeatures) / 2] = cvt_bf16x2((__builtin_amdgcn_fmed3f(0.f, a.z, a.w) - a.x) * a.y,
                                              (__builtin_amdgcn_fmed3f(0.f, b.z, b.w) - b.x) * b.y);
    }
    int sp = 0;
    const int4 root = tv.nodew[0];
    if (root.w == 0) {  // the root is a leaf
        if (lane == 0 && ls.near_cap > 0) near[0] = 0 | kNearCheck;
        nnear = 1;
        near_pairs = unsigned(root.y - root.x);
    } else {
        s_node[w][0] = stack_entry(root.z, root.w);
        sp = 1;
    }
    // The stack holds internal nodes to descend into, as their first child and
    // child count (stack_entry), so a pop needs no node-table load.  The top
    // entries whose children fit in the ROWS rows of one model call are
    // popped together, rows in stack order; lane l then files row l's child:
    // accepted -> far list (order p), leaf -> near list, else -> the stack.
    while (sp > 0) {
        // lane j < ncand holds the j-th entry from the top; incl = its
        // inclusive prefix sum of child counts (counts 0..8, summed bit by bit
        // over the lanes with ballots)
        const int ncand = min(sp, ROWS);
        int kc = 0, kf = 0;
        if (lane < ncand) {
            const int ent = s_node[w][sp - 1 - lane];
            kc = ent >> kNodeBits;
            kf = ent & ((1 << kNodeBits) - 1);
        }
        int incl = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t mb = __ballot((kc >> b) & 1);
            incl += __popcll(mb & (below | (1ull << lane))) << b;
        }
        // the taken entries are a prefix of the lanes (incl is non-decreasing)
        const bool take = lane < ncand && incl <= ROWS;
        const int m = __popcll(__ballot(take));
        const int tot = __builtin_amdgcn_readlane(incl, m - 1);
        if (take) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (k < kc) s_row[w][incl - kc + k] = kf + k;
        }
        wave_sync();
        sp -= m;
        const bool has = lane < tot;
        const int chr = has ? s_row[w][lane] : 0;
        bool prop = false;
        bool near_chk = true;  // a leaf child's sources may come within 1e-6 of a target
        // stereographic geometry test of the pair in the batch's projection,
        // else in the other one
        bool st_geo = false, so_geo = false;
        // every node-table word of the child is loaded here, in one round
        int ckids = 0, cfirst = 0, cb = 0, ce = 0;
        unsigned xin[kIn / 2];
#pragma unroll
        for (int k = 0; k < kIn / 2; ++k) xin[k] = 0u;
        if (has) {
            const int4 nw = tv.nodew[chr];
            cb = nw.x;
            ce = nw.y;
            cfirst = nw.z;
            ckids = nw.w;
            const double4 cc = tv.ncent[chr];
            const double dx = bt.cx - cc.x, dy = bt.cy - cc.y, dz = bt.cz - cc.z;
            const double sd = sqrt(dx * dx + dy * dy + dz * dz);
            const double rn = tv.nrad[chr];
            near_chk = !(sd - rn - bt.r > 1e-6);
            prop = rn + bt.r < theta_max * sd;
            if (prop) {
                if (st_on) st_geo = stereo_fits(sv, sb, chr, sd, rn, bt.r);
                if (!st_geo && so_on) so_geo = stereo_fits(sv, so, chr, sd, rn, bt.r);
                float f[kIn];
                pair_features(f, rn, bt.r, sd, tv.nstat[chr], ckids == 0, bt.nvalid, u_ref);
#pragma unroll
                for (int k = 0; k < kFeatures; k += 2) {
                    const float4 a = s_norm[k], b = s_norm[k + 1];
                    // clamp to [lo, hi] (lo <= hi): the median of the three
                    xin[k / 2] = cvt_bf16x2((__builtin_amdgcn_fmed3f(f[k], a.z, a.w) - a.x) * a.y,
                                            (__builtin_amdgcn_fmed3f(f[k + 1], b.z, b.w) - b.x) * b.y);
                }
#pragma unroll
                for (int k = kFeatures; k < kIn; k += 2) xin[k / 2] = pad[(k - kFeatures) / 2];
            }
        }
        int p = 0;
        if (__ballot(prop)) {
            if (lane < ROWS) {
                uint4* dst = reinterpret_cast<uint4*>(&s_feat[w][lane][0]);
                dst[0] = make_uint4(xin[0], xin[1], xin[2], xin[3]);
                dst[1] = make_uint4(xin[4], xin[5], xin[6], xin[7]);
            }
            wave_sync();
            // lane (g, c) of tile t: inputs 4g .. 4g+3 of row 16 t + c
            bf16x4 x[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const uint2 v = *reinterpret_cast<const uint2*>(&s_feat[w][16 * t + c16][2 * g]);
                x[t] = as_bf16x4(v);
            }
            const int q = mlp_classify_tiles<RT>(s_model, x);
            ++ncalls;
            if (prop && q > 0 && q <= max_order) p = min(q + margin, max_order);
            wave_sync();  // the next call's rows overwrite s_feat
        }
        const bool mine = has;
        const bool leaf = mine && ckids == 0;
        // an accepted pair that passes the stereographic geometry test
        // becomes a stereographic entry of order p + extra
        const bool is_st = st_geo && mine && p > 0, is_so = so_geo && mine && p > 0;
        const int pz = (is_st || is_so) ? min(p + sv.extra, sv.maxp) : 0;
        // a pair for which neither the stereographic nor the Cartesian series
        // of its order is at hand (stereographic-only moments) is refused
        const bool cart = mine && p > 0 && !is_st && !is_so && p <= sv.cart_max;
        const bool is_lo = cart && p <= kLowOrder, is_hi = cart && p > kLowOrder;
        const bool refused = mine && !cart && !is_st && !is_so;
        const bool is_near = refused && leaf, is_desc = refused && !leaf;
        const uint64_t m_lo = __ballot(is_lo), m_hi = __ballot(is_hi), m_st = __ballot(is_st), m_so = __ballot(is_so);
        const uint64_t m_near = __ballot(is_near), m_desc = __ballot(is_desc);
        const int c_lo = __popcll(m_lo), c_hi = __popcll(m_hi), c_near = __popcll(m_near), c_desc = __popcll(m_desc);
        const int c_st = __popcll(m_st), c_so = __popcll(m_so);
        // entries are written only while the batch's lists fit (else the
        // batch is reported as overflowing at the end)
        if (nfar + c_lo + nhi + c_hi <= ls.far_cap) {
            const int e = chr | (p << kNodeBits);
            if (is_lo) far[nfar + __popcll(m_lo & below)] = e;
            if (is_hi) far[ls.far_cap - 1 - nhi - __popcll(m_hi & below)] = e;
        }
        if (nstl + c_st + nso + c_so <= ls.st_cap) {
            if (is_st) stl[nstl + __popcll(m_st & below)] = chr | (pz << kNodeBits);
            if (is_so) stl[ls.st_cap - 1 - nso - __popcll(m_so & below)] = chr | (pz << kNodeBits);
        }
        if (nnear + c_near <= ls.near_cap && is_near)
            near[nnear + __popcll(m_near & below)] = chr | (near_chk ? kNearCheck : 0);
        if (is_desc) s_node[w][sp + c_desc - 1 - __popcll(m_desc & below)] = stack_entry(cfirst, ckids);
        nfar += c_lo;
        nhi += c_hi;
        nstl += c_st;
        nso += c_so;
        nnear += c_near;
        sp += c_desc;
        // terms of the accepted children and sources of the near leaves, kept
        // per lane and summed over the wave after the walk
        const unsigned t_all = cart ? unsigned(p * (p + 1) * (p + 2) / 6) : 0u;
        l_terms += t_all;
        l_terms_hi += is_hi ? t_all : 0u;
        l_np += is_near ? unsigned(ce - cb) : 0u;
        l_terms_st += (is_st || is_so) ? unsigned(pz + 1) : 0u;
        nst += unsigned(c_st + c_so);
        wave_sync();
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        l_terms += __shfl_xor(l_terms, o, 64);
        l_terms_hi += __shfl_xor(l_terms_hi, o, 64);
        l_np += __shfl_xor(l_np, o, 64);
        l_terms_st += __shfl_xor(l_terms_st, o, 64);
    }
    nterms = __builtin_amdgcn_readfirstlane(l_terms);
    nterms_hi = __builtin_amdgcn_readfirstlane(l_terms_hi);
    near_pairs += __builtin_amdgcn_readfirstlane(l_np);
    nterms_st = __builtin_amdgcn_readfirstlane(l_terms_st);
    if (lane == 0) {
        // a batch whose lists do not fit is evaluated as empty and flagged with
        // the capacities it needs; the caller plans again with larger lists
        const bool ovf = nfar + nhi > ls.far_cap || nnear > ls.near_cap || nstl + nso > ls.st_cap;
        ls.far_cnt[bi] = ovf ? 0 : nfar;
        ls.far_hi_cnt[bi] = ovf ? 0 : nhi;
        ls.near_cnt[bi] = ovf ? 0 : nnear;
        ls.st_cnt[bi] = ovf ? 0 : nstl;
        ls.st_o_cnt[bi] = ovf ? 0 : nso;
        const unsigned long long cost = kCostNear * near_pairs + kCostTerm * nterms +
                                        kCostFar * unsigned(nfar + nhi) + kCostSt * nst + kCostStTerm * nterms_st;
        const int bin = cost_bin(ovf ? 0u : unsigned(cost < 0xFFFFFFFFull ? cost : 0xFFFFFFFFull));
        ls.queue[int64_t(bin) * ls.qcap + atomicAdd(&ls.bins[bin], 1u)] = int(bi);
        if (ovf) {
            atomicOr(ls.overflow, 1);
            atomicMax(ls.overflow + 1, nfar + nhi);
            atomicMax(ls.overflow + 2, nnear);
            atomicMax(ls.overflow + 3, nstl + nso);
        }
    }
    const unsigned long long nv = unsigned(bt.nvalid);
    flush_stats(stats, (unsigned long long)(nfar + nhi + nstl + nso) * nv, near_pairs * nv, nterms * nv, ncalls,
                (unsigned long long)nhi * nv, nterms_hi * nv, nst * nv, nterms_st * nv);
}

// The stereographic entries e[0 .. n) of a batch in projection sb.proj,
// their velocity added to (s0, s1, s2): FP32 sums over runs of kStRun
// entries, each run then added in FP64.
__device__ inline void stereo_entries(const StereoView& sv, const StereoBatch& sb, const int* __restrict__ e, int n,
                                      double& s0, double& s1, double& s2) {
    constexpr int kStRun = 16;
    double Sr = 0.0, Si = 0.0, Gs = 0.0;
    float sr = 0.f, si = 0.f, gs = 0.f;
    const int64_t base = int64_t(sb.proj) * sv.nn;
    const unsigned mask = (1u << kNodeBits) - 1u;
    // the next entry's node data is requested before the current entry is
    // summed, the entry after it one step earlier still
    unsigned en = unsigned(__builtin_amdgcn_readfirstlane(e[0]));
    unsigned en1 = unsigned(__builtin_amdgcn_readfirstlane(e[n > 1 ? 1 : 0]));
    StNode cur = stereo_fetch(sv, base + (en & mask));
    for (int j = 0; j < n; ++j) {
        const unsigned en2 = unsigned(__builtin_amdgcn_readfirstlane(e[j + 2 < n ? j + 2 : n - 1]));
        const StNode nxt = stereo_fetch(sv, base + (en1 & mask));
        stereo_sum(cur, int(en >> kNodeBits), base + (en & mask), sv, sb.zr, sb.zi, sr, si, gs);
        cur = nxt;
        en = en1;
        en1 = en2;
        if ((j % kStRun) == kStRun - 1 || j + 1 == n) {
            Sr += double(sr);
            Si += double(si);
            Gs += double(gs);
            sr = si = gs = 0.f;
        }
    }
    stereo_add(sb.proj, sb.zr, sb.zi, Sr, Si, Gs, s0, s1, s2);
}

// The batch's interaction list.  LOW: near leaves and the far entries of
// order <= kLowOrder, written to acc[3][nt] (sorted positions, FP64); else
// the far entries of order kLowOrder + 1 .. MAXP, added to acc.  The order is
// wave-uniform per entry.
// Output: acc[3][nt] at the sorted positions (LOW writes, else adds): a
// wave's 64 targets are consecutive there, so its stores fill whole lines
// (tree_list_finish then scatters to the targets' own indices).
template <int MAXP, bool LOW, bool F32, bool STEREO = false>
__global__ __launch_bounds__(256) void tree_eval_lists(Targets tg, TreeView tv, Lists ls, StereoView sv, double d2,
                                                       float eps2v, double* __restrict__ acc) {
    const int64_t wi = wave_index();
    if (wi >= tg.nb) return;
    const int64_t bi = list_batch(ls, wi);
    if (bi < 0) return;
    const int* far = ls.far + bi * ls.far_cap;
    int j0 = 0, j1 = __builtin_amdgcn_readfirstlane(ls.far_cnt[bi]);
    if (!LOW) {
        j1 = ls.far_cap;
        j0 = j1 - __builtin_amdgcn_readfirstlane(ls.far_hi_cnt[bi]);
        if (j0 == j1) return;
    }
    bool valid;
    int orig;
    double t0, t1, t2;
    const int64_t pos = load_target(tg, bi, valid, orig, t0, t1, t2);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (LOW) {
        const int* near = ls.near + bi * ls.near_cap;
        const int nnear = __builtin_amdgcn_readfirstlane(ls.near_cnt[bi]);
        // the next leaf's entry, range and centre are requested before the
        // current leaf is summed
        int ent = nnear > 0 ? __builtin_amdgcn_readfirstlane(near[0]) : 0;
        int4 nw = tv.nodew[ent & ~kNearCheck];
        double4 cc = tv.ncent[ent & ~kNearCheck];
        for (int j = 0; j < nnear; ++j) {
            const int ent1 = __builtin_amdgcn_readfirstlane(near[j + 1 < nnear ? j + 1 : j]);
            const int4 nw1 = tv.nodew[ent1 & ~kNearCheck];
            const double4 cc1 = tv.ncent[ent1 & ~kNearCheck];
            bool chk = !F32 || (ent & kNearCheck);
            if (F32 && chk) {
                // a source within 1e-6 of a target needs the target within
                // r_n + 1e-6 of the leaf centre (FP32 offsets: 2e-6)
                const float q0 = float(t0 - cc.x), q1 = float(t1 - cc.y), q2 = float(t2 - cc.z);
                const float lim = float(tv.nrad[ent & ~kNearCheck]) + 2e-6f;
                chk = __ballot(valid && fmaf(q0, q0, fmaf(q1, q1, q2 * q2)) <= lim * lim) != 0ull;
            }
            if (chk)
                near_leaf<F32, true>(nw.x, nw.y, t0, t1, t2, cc, d2, eps2v, tv.spos, tv.soff, s0, s1, s2, v0, v1, v2);
            else
                near_leaf<F32, false>(nw.x, nw.y, t0, t1, t2, cc, d2, eps2v, tv.spos, tv.soff, s0, s1, s2, v0, v1,
                                      v2);
            ent = ent1;
            nw = nw1;
            cc = cc1;
        }
    }
    for (int j = j0; j < j1; ++j) {
        const unsigned e = unsigned(__builtin_amdgcn_readfirstlane(far[j]));
        const int n = int(e & ((1u << kNodeBits) - 1u));
        const int p = int((e >> kNodeBits) & 15u);
        far_node<MAXP, F32, true>(p, n, tv.nterm, t0, t1, t2, tv.ncent[n], d2, tv.momg, tv.mom32, s0, s1, s2, v0, v1,
                                  v2);
    }
    cross_add(t0, t1, t2, v0, v1, v2, s0, s1, s2);
    if (STEREO && LOW) {
        // stereographic entries: in the batch's projection from the front of
        // the list, in the other projection from the back
        const int* stl = ls.st + bi * ls.st_cap;
        const int nst = __builtin_amdgcn_readfirstlane(ls.st_cnt[bi]);
        const int nso = __builtin_amdgcn_readfirstlane(ls.st_o_cnt[bi]);
        if (nst + nso > 0) {
            const Batch bt = batch_of(valid, t0, t1, t2);
            const StereoBatch sb = stereo_batch(bt, valid, t0, t1, t2);
            if (nst > 0) stereo_entries(sv, sb, stl, nst, s0, s1, s2);
            if (nso > 0)
                stereo_entries(sv, stereo_batch_in(1 - sb.proj, bt, valid, t0, t1, t2), stl + ls.st_cap - nso, nso,
                               s0, s1, s2);
        }
    }
    if (valid) {
        if (LOW) {
            acc[pos] = s0;
            acc[tg.n + pos] = s1;
            acc[2 * tg.n + pos] = s2;
        } else {
            acc[pos] += s0;
            acc[tg.n + pos] += s1;
            acc[2 * tg.n + pos] += s2;
        }
    }
}


__global__ __launch_bounds__(256) void tree_list_finish(const int* __restrict__ perm, int64_t n,
                                                        const double* __restrict__ acc, double mult,
                                                        double* __restrict__ ux, double* __restrict__ uy,
                                                        double* __restrict__ uz) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    ux[o] = acc[i] * mult;
    uy[o] = acc[n + i] * mult;
    uz[o] = acc[2 * n + i] * mult;
}

// ---------------------------------------------------------------- streamfunction
// psi(t) = -(1/4 pi) sum_j gamma_j log(1 - t.y_j) (src/LpmScalarKernel.h:31-45)
// with the batch acceptance of tree_eval_batch.  Far nodes: the log series
// of order P about the node centre from the scalar moments, minus ln 2 times
// the node's total strength (log(1 - t.y) = log(|t - y|^2 / 2) on the unit
// sphere).  Near leaves: summed directly -- FP64: the reference's pair
// log(1 - t.y); FP32: log2 |t - y|^2 of the FP32 offsets.  Pairs at zero
// distance contribute nothing (the singular kernel has no self term).
template <int P, bool F32>
__global__ __launch_bounds__(256) void tree_eval_psi(Targets tg, TreeView tv, const double* __restrict__ pmomg,
                                                     const float* __restrict__ pmom32, double theta, double mult,
                                                     double* __restrict__ psi, unsigned long long* __restrict__ stats) {
    __shared__ int s_node[4][kStackB];
    constexpr double kLn2 = 0.69314718055994530942;
    const int w = threadIdx.x >> 6;
    const int64_t bi = wave_index();
    if (!is_batch(tg, bi)) return;
    bool valid;
    int orig;
    double t0, t1, t2;
    load_target(tg, bi, valid, orig, t0, t1, t2);
    const Batch bt = batch_of(valid, t0, t1, t2);
    double acc = 0.0;
    unsigned long long nfar = 0, nnear = 0;
    const int nterm = tv.nterm;
    int sp = 1;
    s_node[w][0] = 0;
    while (sp > 0) {
        const int n = __builtin_amdgcn_readfirstlane(s_node[w][--sp]);
        const double4 c = tv.ncent[n];
        const double dx = bt.cx - c.x, dy = bt.cy - c.y, dz = bt.cz - c.z;
        const double s = sqrt(dx * dx + dy * dy + dz * dz);
        const double rn = tv.nrad[n];
        if (uniform(rn + bt.r < theta * s)) {
            const double* mu = pmomg + int64_t(n) * nterm;
            double S;
            if constexpr (F32)
                S = far_log_scaled<P>(float(t0 - c.x), float(t1 - c.y), float(t2 - c.z), 0.f, float(c.w),
                                      pmom32 + int64_t(n) * nterm);
            else
                S = far_log<P>(t0 - c.x, t1 - c.y, t2 - c.z, 0.0, mu);
            acc += S - kLn2 * mu[0];
            ++nfar;
            continue;
        }
        const int nk = tv.nkids[n];
        if (nk == 0) {
            const int b = tv.nbeg[n], e = tv.nend[n];
            if constexpr (F32) {
                const float q0 = float(t0 - c.x), q1 = float(t1 - c.y), q2 = float(t2 - c.z);
                // a target that is one of the leaf's sources (a zero-distance
                // pair, left out) lies within r_n of the leaf centre (FP32
                // offsets: 2e-6 of margin); other waves sum every pair and
                // take the leaf's strength sum from its degree-0 moment
                const float lim = float(rn) + 2e-6f;
                if (uniform(__ballot(valid && fmaf(q0, q0, fmaf(q1, q1, q2 * q2)) <= lim * lim) != 0ull)) {
                    float l2 = 0.f, gs = 0.f;
                    for (int j = b; j < e; ++j) {
                        const float4 o = tv.soff[j];
                        const float ex = q0 - o.x, ey = q1 - o.y, ez = q2 - o.z;
                        const float r2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                        const bool ok = r2 > 0.f;
                        l2 = fmaf(ok ? o.w : 0.f, __builtin_amdgcn_logf(ok ? r2 : 1.f), l2);
                        gs += ok ? o.w : 0.f;
                    }
                    acc += kLn2 * (double(l2) - double(gs));
                } else {
                    float l2 = 0.f;
                    for (int j = b; j < e; ++j) {
                        const float4 o = tv.soff[j];
                        const float ex = q0 - o.x, ey = q1 - o.y, ez = q2 - o.z;
                        l2 = fmaf(o.w, __builtin_amdgcn_logf(fmaf(ex, ex, fmaf(ey, ey, ez * ez))), l2);
                    }
                    acc += kLn2 * (double(l2) - pmomg[int64_t(n) * nterm]);
                }
            } else {
                for (int j = b; j < e; ++j) {
                    const double4 y = tv.spos[j];
                    const double ex = t0 - y.x, ey = t1 - y.y, ez = t2 - y.z;
                    if (ex * ex + ey * ey + ez * ez > 0.0) acc += log(1.0 - (t0 * y.x + t1 * y.y + t2 * y.z)) * y.w;
                }
            }
            nnear += unsigned(e - b);
        } else {
            const int f = tv.nfirst[n];
            for (int q = nk - 1; q >= 0; --q) s_node[w][sp++] = f + q;
        }
    }
    if (valid) psi[orig] = acc * mult;
    const unsigned long long nv = unsigned(bt.nvalid);
    flush_stats(stats, nfar * nv, nnear * nv, nfar * nv * unsigned(P * (P + 1) * (P + 2) / 6), 0ull);
}

// ---------------------------------------------------------------- labels
__device__ inline uint32_t mix32(uint32_t h) {  // MurmurHash3 finalizer
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__global__ __launch_bounds__(256) void tree_labels_kernel(Targets tg, TreeView tv, int64_t nsel, int64_t bstride,
                                                          double theta_max, double d2, double u_ref,
                                                          uint32_t keep_threshold, int big, uint32_t seed, int frontier,
                                                          int64_t max_samples, unsigned long long* __restrict__ counter,
                                                          float* __restrict__ feat_out, float* __restrict__ err_out,
                                                          int* __restrict__ id_out) {
    __shared__ int s_node[4][kStackB];
    __shared__ int s_prop[4][kStackB];  // proposed strict ancestors of the entry
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t sel = wave_index();
    if (sel >= nsel) return;
    const int64_t bi = sel * bstride;
    bool valid;
    int orig;
    double t0, t1, t2;
    load_target(tg, bi, valid, orig, t0, t1, t2);
    const Batch bt = batch_of(valid, t0, t1, t2);
    const double inv = 1.0 / (4.0 * 3.14159265358979323846 * u_ref);
    int sp = 1;
    s_node[w][0] = 0;
    s_prop[w][0] = 0;
    while (sp > 0) {
        --sp;
        const int n = __builtin_amdgcn_readfirstlane(s_node[w][sp]);
        const int a = __builtin_amdgcn_readfirstlane(s_prop[w][sp]);
        const double4 c = tv.ncent[n];
        const double dx = bt.cx - c.x, dy = bt.cy - c.y, dz = bt.cz - c.z;
        const double s = sqrt(dx * dx + dy * dy + dz * dz);
        const int nk = tv.nkids[n];
        const int b = tv.nbeg[n], e = tv.nend[n];
        const double rn = tv.nrad[n];
        const bool prop = uniform(rn + bt.r < theta_max * s);
        if (prop && (frontier < 0 || a <= frontier)) {
            const uint32_t h = mix32(uint32_t(bi) * 0x9E3779B1u ^ mix32(uint32_t(n) + seed));
            const uint32_t h2 = mix32(h ^ 0x27d4eb2fu);
            const bool keep = h <= keep_threshold && (e - b <= big || int(h2 % uint32_t(e - b)) < big);
            if (keep) {
                unsigned long long slot = 0;
                if (lane == 0) slot = atomicAdd(counter, 1ull);
                slot = __shfl(slot, 0, 64);
                if (int64_t(slot) < max_samples) {
                    float f[kIn];
                    pair_features(f, rn, bt.r, s, tv.nstat[n], nk == 0, bt.nvalid, u_ref);
                    if (lane == 0) {
                        for (int k = 0; k < kIn; ++k) feat_out[slot * kIn + k] = f[k];
                        id_out[2 * slot] = int(bi);
                        id_out[2 * slot + 1] = n;
                    }
                    // exact S = sum_j gamma_j y_j / (1 - t.y_j + delta^2) over the node
                    double E0 = 0.0, E1 = 0.0, E2 = 0.0;
                    for (int j = b; j < e; ++j) {
                        const double4 y = tv.spos[j];
                        const double den = 1.0 - (t0 * y.x + t1 * y.y + t2 * y.z) + d2;
                        if (fabs(den) > 1.0e-13) {
                            const double g = y.w / den;
                            E0 = fma(g, y.x, E0);
                            E1 = fma(g, y.y, E1);
                            E2 = fma(g, y.z, E2);
                        }
                    }
                    double err[kLabelOrder];
                    auto on_shell = [&](int q, double S0, double S1, double S2) {
                        const double a0 = S0 - E0, a1 = S1 - E1, a2 = S2 - E2;
                        const double v0 = t1 * a2 - t2 * a1, v1 = t2 * a0 - t0 * a2, v2 = t0 * a1 - t1 * a0;
                        err[q] = sqrt(v0 * v0 + v1 * v1 + v2 * v2) * inv;
                    };
                    far_cartesian_shells<kLabelOrder>(t0 - c.x, t1 - c.y, t2 - c.z, 2.0 * d2,
                                                      tv.momg + int64_t(n) * tv.nterm * 3, on_shell);
#pragma unroll
                    for (int q = 0; q < kLabelOrder; ++q) {
                        const double m = wmax(valid ? err[q] : 0.0);
                        if (lane == 0) err_out[slot * kLabelOrder + q] = float(fmin(m, 3.0e38));
                    }
                }
            }
        }
        const int a2 = a + (prop ? 1 : 0);
        if (nk > 0 && (frontier < 0 || a2 <= frontier)) {
            const int f = tv.nfirst[n];
            for (int q = nk - 1; q >= 0; --q) {
                s_node[w][sp] = f + q;
                s_prop[w][sp] = a2;
                ++sp;
            }
        }
    }
}

TreeView view(const Tree* t) {
    TreeView v;
    v.ncent = t->ncent.as<double4>();
    v.nrad = t->nrad.as<double>();
    v.nfirst = t->nfirst.as<int>();
    v.nkids = t->nkids.as<int>();
    v.nbeg = t->nbeg.as<int>();
    v.nend = t->nend.as<int>();
    v.nodew = t->nodew.as<int4>();
    v.nstat = t->nstat.as<float4>();
    v.spos = t->spos.as<double4>();
    v.soff = t->soff32.as<float4>();
    v.momg = t->momg.as<double>();
    v.mom32 = t->mom32.as<float>();
    v.nterm = nterms(t->order);
    return v;
}

constexpr double kPi = 3.1415926535897932384626433832795027975;

}  // namespace

int tree_eval_batch_device(lpm_ctx* c, int prec, int order, double theta, double delta, const double* tx,
                           const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz) {
    Tree* t = c->tree;
    const int* tperm = nullptr;
    int* bstart = nullptr;
    int64_t nbatch = 0;
    const int* nbd = nullptr;
    LPM_CHECK(tree_target_batches(c, tx, ty, tz, nt, &tperm, &bstart, &nbatch, &nbd));
    if (nt == 0) return LPM_OK;
    const Targets tg{tx, ty, tz, tperm, nt, bstart, nbatch, nbd};
    const TreeView tv = view(t);
    const unsigned nb = unsigned((nbatch + 3) / 4);
    const double d2 = delta * delta, mult = -1.0 / (4.0 * kPi);
    const float eps2v = float(2.0 * d2);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, c->stream, tg, tv, theta, d2, eps2v, mult, ux, uy, uz,
                           c->stats_ptr);
    };
    const bool f32 = prec == LPM_F32C;
    mark_begin(c);
    switch (order) {
#define LPM_CASE(PP)                                  \
    case PP:                                          \
        if (f32) go(tree_eval_batch<PP, true>);       \
        else go(tree_eval_batch<PP, false>);          \
        break;
        LPM_CASE(1) LPM_CASE(2) LPM_CASE(3) LPM_CASE(4) LPM_CASE(5) LPM_CASE(6)
        LPM_CASE(7) LPM_CASE(8) LPM_CASE(9) LPM_CASE(10) LPM_CASE(11) LPM_CASE(12)
#undef LPM_CASE
        default: return fail(c, LPM_ERR_INVALID_ARG, "series order %d unsupported", order);
    }
    mark_end(c);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

int tree_eval_psi_device(lpm_ctx* c, int prec, int order, double theta, const double* tx, const double* ty,
                         const double* tz, int64_t nt, double* psi) {
    Tree* t = c->tree;
    const int* tperm = nullptr;
    int* bstart = nullptr;
    int64_t nbatch = 0;
    const int* nbd = nullptr;
    LPM_CHECK(tree_target_batches(c, tx, ty, tz, nt, &tperm, &bstart, &nbatch, &nbd));
    if (nt == 0) return LPM_OK;
    const Targets tg{tx, ty, tz, tperm, nt, bstart, nbatch, nbd};
    TreeView tv = view(t);
    tv.nterm = nterms(t->psi_order);
    const unsigned nb = unsigned((nbatch + 3) / 4);
    const double mult = -1.0 / (4.0 * kPi);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, c->stream, tg, tv, t->pmomg.as<double>(),
                           t->pmom32.as<float>(), theta, mult, psi, c->stats_ptr);
    };
    const bool f32 = prec == LPM_F32C;
    mark_begin(c);
    switch (order) {
#define LPM_CASE(PP)                                \
    case PP:                                        \
        if (f32) go(tree_eval_psi<PP, true>);       \
        else go(tree_eval_psi<PP, false>);          \
        break;
        LPM_CASE(1) LPM_CASE(2) LPM_CASE(3) LPM_CASE(4) LPM_CASE(5) LPM_CASE(6)
        LPM_CASE(7) LPM_CASE(8) LPM_CASE(9) LPM_CASE(10) LPM_CASE(11) LPM_CASE(12)
#undef LPM_CASE
        default: return fail(c, LPM_ERR_INVALID_ARG, "series order %d unsupported", order);
    }
    mark_end(c);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

int tree_eval_learned_device(lpm_ctx* c, int prec, double theta_max, int margin, int max_order, double u_ref,
                             double delta, const double* tx, const double* ty, const double* tz, int64_t nt,
                             double* ux, double* uy, double* uz) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    if (t->nn >= (int64_t(1) << kNodeBits)) return fail(c, LPM_ERR_INVALID_ARG, "too many tree nodes");
    const int* tperm = nullptr;
    int* bstart = nullptr;
    int64_t nbatch = 0;
    const int* nbd = nullptr;
    LPM_CHECK(tree_target_batches(c, tx, ty, tz, nt, &tperm, &bstart, &nbatch, &nbd));
    if (nt == 0) return LPM_OK;
    const Targets tg{tx, ty, tz, tperm, nt, bstart, nbatch, nbd};
    const TreeView tv = view(t);
    MlpNorm norm;
    for (int k = 0; k < kIn; ++k) {
        norm.mean[k] = model_norm(c, 0)[k];
        norm.inv_std[k] = model_norm(c, 1)[k];
        norm.lo[k] = model_norm(c, 2)[k];
        norm.hi[k] = model_norm(c, 3)[k];
    }
    const unsigned nb = unsigned((nbatch + 3) / 4);
    const double d2 = delta * delta, mult = -1.0 / (4.0 * kPi);
    const float eps2v = float(2.0 * d2);
    const unsigned short* w = model_weights(c);
    const bool f32 = prec == LPM_F32C;
    // stereographic far entries: FP32C with delta = 0 and the option set
    StereoView sv{t->sgeo.as<double4>(), t->smom32.as<float2>(), t->nn, 0, c->stereo_extra, t->stereo_terms - 1,
                  c->stereo_ratio, t->order};
    if (c->stereo && f32 && d2 == 0.0) {
        if (t->stereo_terms == 0)
            return fail(c, LPM_ERR_STATE, "no stereographic moments (lpm_tree_moments after setting LPM_OPT_STEREO)");
        sv.on = 1;
    }
    // plan, evaluate, and check once at the end that every batch's lists fit;
    // if one did not, grow the lists to what it needed and do it again
    for (int attempt = 0;; ++attempt) {
        const int st_cap = sv.on ? c->st_cap : 0;
        const size_t bytes =
            size_t(nbatch) * (size_t(c->far_cap) + size_t(c->near_cap) + size_t(st_cap) + 5 + kCostBins) *
                sizeof(int) + size_t(3 * nt) * sizeof(double) + 256;
        LPM_CHECK(c->tl_work.reserve(bytes));
        double* acc = c->tl_work.as<double>();
        Lists ls;
        ls.far = reinterpret_cast<int*>(acc + 3 * nt);
        ls.near = ls.far + size_t(nbatch) * c->far_cap;
        ls.st = ls.near + size_t(nbatch) * c->near_cap;
        ls.far_cnt = ls.st + size_t(nbatch) * st_cap;
        ls.far_hi_cnt = ls.far_cnt + nbatch;
        ls.near_cnt = ls.far_hi_cnt + nbatch;
        ls.st_cnt = ls.near_cnt + nbatch;
        // the overflow words and the cost bins sit in the batch header,
        // cleared with the counters
        BatchHeader* hdr = batch_header(c->stats_ptr);
        ls.overflow = hdr->overflow;
        ls.bins = hdr->bins;
        ls.st_o_cnt = ls.st_cnt + nbatch;
        ls.queue = ls.st_o_cnt + nbatch;
        ls.qcap = nbatch;
        ls.far_cap = c->far_cap;
        ls.near_cap = c->near_cap;
        ls.st_cap = st_cap;
        if (attempt > 0) {
            LPM_HIP(c, hipMemsetAsync(hdr->overflow, 0, sizeof(hdr->overflow), st));
            LPM_HIP(c, hipMemsetAsync(hdr->bins, 0, sizeof(hdr->bins), st));
            LPM_HIP(c, hipMemsetAsync(c->stats_ptr, 0, kStatBytes, st));
        }
        LPM_HIP(c, join_aux(c));  // the stereographic node data
        auto plan = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, tg, tv, theta_max, margin, max_order, u_ref, w, norm,
                               ls, sv, c->stats_ptr);
        };
        // the most rows per model call whose stack bound fits this tree
        const int rows = plan_stack_bound(64, t->depth) <= kStackPlan   ? 64
                         : plan_stack_bound(32, t->depth) <= kStackPlan ? 32
                                                                        : 16;
        if (sv.on) {
            if (rows == 64) plan(tree_plan_learned<true, 64>);
            else if (rows == 32) plan(tree_plan_learned<true, 32>);
            else plan(tree_plan_learned<true, 16>);
        } else {
            if (rows == 64) plan(tree_plan_learned<false, 64>);
            else if (rows == 32) plan(tree_plan_learned<false, 32>);
            else plan(tree_plan_learned<false, 16>);
        }
        auto lists = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, tg, tv, ls, sv, d2, eps2v, acc);
        };
        // the timed launch is the low-order list kernel, the one that dominates
        mark_begin(c);
        if (sv.on) lists(tree_eval_lists<kLowOrder, true, true, true>);
        else if (f32) lists(tree_eval_lists<kLowOrder, true, true>);
        else lists(tree_eval_lists<kLowOrder, true, false>);
        mark_end(c);
        if (max_order > kLowOrder) {
            if (max_order <= 10) {
                if (f32) lists(tree_eval_lists<10, false, true>);
                else lists(tree_eval_lists<10, false, false>);
            } else {
                if (f32) lists(tree_eval_lists<12, false, true>);
                else lists(tree_eval_lists<12, false, false>);
            }
        }
        hipLaunchKernelGGL(tree_list_finish, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st, tperm, nt, acc, mult,
                           ux, uy, uz);
        LPM_HIP(c, hipGetLastError());
        LPM_HIP(c, hipMemcpyAsync(c->pin, ls.overflow, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipStreamSynchronize(st));
        if (!c->pin[0]) break;
        if (attempt > 0) return fail(c, LPM_ERR_KERNEL, "interaction lists overflow after growing");
        c->far_cap = std::max(c->far_cap, c->pin[1]);
        c->near_cap = std::max(c->near_cap, c->pin[2]);
        c->st_cap = std::max(c->st_cap, c->pin[3]);
    }
    return LPM_OK;
}

// Sort keys of the label rows: (batch, node) as one 64-bit key per row.
__global__ __launch_bounds__(256) void label_keys(const int* __restrict__ ids, int64_t n, uint64_t* __restrict__ key,
                                                  int* __restrict__ val) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (uint64_t(unsigned(ids[2 * i])) << 32) | unsigned(ids[2 * i + 1]);
    val[i] = int(i);
}

// Row i of the sorted set = row src[i] of the unsorted one.
__global__ __launch_bounds__(256) void label_gather(const int* __restrict__ src, int64_t n,
                                                    const float* __restrict__ f_in, const float* __restrict__ e_in,
                                                    const int* __restrict__ id_in, float* __restrict__ f_out,
                                                    float* __restrict__ e_out, int* __restrict__ id_out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = src[i];
    for (int k = 0; k < kIn; ++k) f_out[i * kIn + k] = f_in[r * kIn + k];
    for (int k = 0; k < kLabelOrder; ++k) e_out[i * kLabelOrder + k] = e_in[r * kLabelOrder + k];
    id_out[2 * i] = id_in[2 * r];
    id_out[2 * i + 1] = id_in[2 * r + 1];
}

static int tree_labels_device(lpm_ctx* c, double theta_max, double delta, double u_ref, int64_t bstride,
                              double keep_rate, int frontier, uint32_t seed, int64_t max_samples, const double* tx,
                              const double* ty, const double* tz, int64_t nt, float* feat, float* err, int* ids,
                              int64_t* n_out) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    const int* tperm = nullptr;
    int* bstart = nullptr;
    int64_t nbatch = 0;
    const int* nbd = nullptr;
    LPM_CHECK(tree_target_batches(c, tx, ty, tz, nt, &tperm, &bstart, &nbatch, &nbd));
    if (nt > 0) LPM_CHECK(read_batch_count(c, nbd, &nbatch));
    const Targets tg{tx, ty, tz, tperm, nt, bstart, nbatch, nbd};
    const TreeView tv = view(t);
    const int64_t nsel = (nbatch + bstride - 1) / bstride;
    // the first evaluation counter serves as the sample counter
    unsigned long long* counter = c->stats_ptr;
    const uint32_t thr = keep_rate >= 1.0 ? 0xFFFFFFFFu : uint32_t(keep_rate * 4294967295.0);
    if (nsel > 0)
        hipLaunchKernelGGL(tree_labels_kernel, dim3(unsigned((nsel + 3) / 4)), dim3(256), 0, st, tg, tv, nsel,
                           bstride, theta_max, delta * delta, u_ref, thr, 1 << 16, seed, frontier, max_samples,
                           counter, feat, err, ids);
    LPM_HIP(c, hipGetLastError());
    unsigned long long h = 0;
    LPM_HIP(c, hipMemcpyAsync(&h, counter, sizeof(h), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    const int64_t n = std::min<int64_t>(int64_t(h), max_samples);
    *n_out = n;
    if (n < 2) return LPM_OK;
    // The rows arrive in the order of an atomic counter, which varies from
    // run to run; sorted by (batch, node) they are the same set in the same
    // order every time, so training (which shuffles with a fixed seed) is
    // reproducible.  Every (batch, node) pair is labelled at most once.
    size_t sort_bytes = 0;
    uint64_t* k0 = nullptr;
    int* v0 = nullptr;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k0, v0, v0, int(n), 0, 64, st));
    const size_t rows = size_t(n) * (kIn + kLabelOrder + 2) * 4;
    LPM_CHECK(c->lab_work.reserve(size_t(n) * 24 + rows + sort_bytes + 1024));
    uint64_t* kin = c->lab_work.as<uint64_t>();
    uint64_t* kout = kin + n;
    int* vin = reinterpret_cast<int*>(kout + n);
    int* vout = vin + n;
    float* f2 = reinterpret_cast<float*>((reinterpret_cast<uintptr_t>(vout + n) + 255) & ~uintptr_t(255));
    float* e2 = f2 + size_t(n) * kIn;
    int* i2 = reinterpret_cast<int*>(e2 + size_t(n) * kLabelOrder);
    void* tmp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(i2 + 2 * n) + 255) & ~uintptr_t(255));
    const unsigned g = unsigned((n + 255) / 256);
    hipLaunchKernelGGL(label_keys, dim3(g), dim3(256), 0, st, ids, n, kin, vin);
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, kin, kout, vin, vout, int(n), 0, 64, st));
    hipLaunchKernelGGL(label_gather, dim3(g), dim3(256), 0, st, vout, n, feat, err, ids, f2, e2, i2);
    LPM_HIP(c, hipGetLastError());
    LPM_HIP(c, hipMemcpyAsync(feat, f2, size_t(n) * kIn * 4, hipMemcpyDeviceToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(err, e2, size_t(n) * kLabelOrder * 4, hipMemcpyDeviceToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(ids, i2, size_t(n) * 2 * 4, hipMemcpyDeviceToDevice, st));
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

namespace {

static int check_targets(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt, double* ux,
                         double* uy, double* uz) {
    if (nt < 0 || nt > (int64_t(1) << 30)) return fail(c, LPM_ERR_INVALID_ARG, "invalid target count");
    if (nt > 0 && (!tx || !ty || !tz || !ux || !uy || !uz)) return fail(c, LPM_ERR_NULL_PTR, "NULL target/output");
    return LPM_OK;
}

static int check_tree_for_batch(lpm_ctx* c, int order) {
    if (!c->tree) return fail(c, LPM_ERR_STATE, "no tree (lpm_tree_build first)");
    if (c->tree->order < order) return fail(c, LPM_ERR_STATE, "moments of order >= %d needed", order);
    if (c->tree->moment_sign != 1) return fail(c, LPM_ERR_STATE, "the Cartesian series needs moment_sign +1");
    return LPM_OK;
}

// Host-pointer wrapper: stage the targets, run `fn` on device copies, copy the velocity back.
template <class F>
static int with_staged(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt, double* ux,
                       double* uy, double* uz, F fn) {
    hipStream_t st = c->stream;
    LPM_CHECK(c->h_in.reserve(6 * size_t(nt) * sizeof(double)));
    double* d = c->h_in.as<double>();
    LPM_HIP(c, hipMemcpyAsync(d, tx, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + nt, ty, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + 2 * nt, tz, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_CHECK(fn(d, d + nt, d + 2 * nt, d + 3 * nt, d + 4 * nt, d + 5 * nt));
    LPM_HIP(c, hipMemcpyAsync(ux, d + 3 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(uy, d + 4 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(uz, d + 5 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

static int check_batch(lpm_ctx* c, int prec, int order, double theta, double delta, const double* tx,
                       const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (prec != LPM_F64 && prec != LPM_F32C) return fail(c, LPM_ERR_INVALID_ARG, "invalid precision %d", prec);
    if (order < 1 || order > kMaxLearnedOrder) return fail(c, LPM_ERR_INVALID_ARG, "order %d unsupported", order);
    if (!(theta >= 0.0) || !(delta >= 0.0)) return fail(c, LPM_ERR_INVALID_ARG, "theta and delta must be >= 0");
    LPM_CHECK(check_tree_for_batch(c, order));
    return check_targets(c, tx, ty, tz, nt, ux, uy, uz);
}

static int check_learned(lpm_ctx* c, int prec, double theta_max, int margin, int max_order, double u_ref,
                         double delta, const double* tx, const double* ty, const double* tz, int64_t nt, double* ux,
                         double* uy, double* uz) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (prec != LPM_F64 && prec != LPM_F32C) return fail(c, LPM_ERR_INVALID_ARG, "invalid precision %d", prec);
    if (max_order < 1 || max_order > kMaxLearnedOrder)
        return fail(c, LPM_ERR_INVALID_ARG, "max_order must be in 1..%d", kMaxLearnedOrder);
    if (margin < 0 || !(u_ref > 0.0) || !(delta >= 0.0) || !(theta_max >= 0.0) || !std::isfinite(theta_max))
        return fail(c, LPM_ERR_INVALID_ARG, "invalid learned-evaluation arguments");
    // the far field of a stereographic FP32C evaluation may come from the
    // stereographic moments alone (LPM_OPT_STEREO 2)
    const Tree* t = c->tree;
    const bool st_far = t && c->stereo && t->stereo_terms > 0 && prec == LPM_F32C && delta == 0.0;
    LPM_CHECK(check_tree_for_batch(c, st_far ? std::min(max_order, t->order) : max_order));
    if (st_far && std::max(t->order, t->stereo_order) < max_order)
        return fail(c, LPM_ERR_STATE, "moments of order >= %d needed", max_order);
    if (!model_weights(c)) return fail(c, LPM_ERR_STATE, "no order model (lpm_model_init or lpm_model_set_params)");
    return check_targets(c, tx, ty, tz, nt, ux, uy, uz);
}

}  // namespace

extern "C" {

int lpm_tree_eval_batch_dev(lpm_ctx* c, int prec, int order, double theta, double delta, const double* tx,
                            const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz) {
    LPM_CHECK(check_batch(c, prec, order, theta, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_CHECK(check_tree_sphere(c, tx, ty, tz, nt, "lpm_tree_eval_batch"));
    return tree_eval_batch_device(c, prec, order, theta, delta, tx, ty, tz, nt, ux, uy, uz);
}

int lpm_tree_eval_batch(lpm_ctx* c, int prec, int order, double theta, double delta, const double* tx,
                        const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz) {
    LPM_CHECK(check_batch(c, prec, order, theta, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    if (nt == 0) return tree_eval_batch_device(c, prec, order, theta, delta, tx, ty, tz, 0, ux, uy, uz);
    return with_staged(c, tx, ty, tz, nt, ux, uy, uz,
                       [&](const double* x, const double* y, const double* z, double* u, double* v, double* w) {
                           LPM_CHECK(check_tree_sphere(c, x, y, z, nt, "lpm_tree_eval_batch"));
                           return tree_eval_batch_device(c, prec, order, theta, delta, x, y, z, nt, u, v, w);
                       });
}

static int check_psi(lpm_ctx* c, int prec, int order, double theta, const double* tx, const double* ty,
                     const double* tz, int64_t nt, double* psi) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (prec != LPM_F64 && prec != LPM_F32C) return fail(c, LPM_ERR_INVALID_ARG, "invalid precision %d", prec);
    if (order < 1 || order > kMaxLearnedOrder) return fail(c, LPM_ERR_INVALID_ARG, "order %d unsupported", order);
    if (!(theta >= 0.0)) return fail(c, LPM_ERR_INVALID_ARG, "theta must be >= 0");
    if (!c->tree) return fail(c, LPM_ERR_STATE, "no tree (lpm_tree_build first)");
    if (c->tree->psi_order < order)
        return fail(c, LPM_ERR_STATE, "streamfunction moments of order >= %d needed (lpm_tree_moments_psi)", order);
    if (nt < 0 || nt > (int64_t(1) << 30)) return fail(c, LPM_ERR_INVALID_ARG, "invalid target count");
    if (nt > 0 && (!tx || !ty || !tz || !psi)) return fail(c, LPM_ERR_NULL_PTR, "NULL target/output");
    return LPM_OK;
}

int lpm_tree_eval_psi_dev(lpm_ctx* c, int prec, int order, double theta, const double* tx, const double* ty,
                          const double* tz, int64_t nt, double* psi) {
    LPM_CHECK(check_psi(c, prec, order, theta, tx, ty, tz, nt, psi));
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_CHECK(check_tree_sphere(c, tx, ty, tz, nt, "lpm_tree_eval_psi"));
    return tree_eval_psi_device(c, prec, order, theta, tx, ty, tz, nt, psi);
}

int lpm_tree_eval_psi(lpm_ctx* c, int prec, int order, double theta, const double* tx, const double* ty,
                      const double* tz, int64_t nt, double* psi) {
    LPM_CHECK(check_psi(c, prec, order, theta, tx, ty, tz, nt, psi));
    LPM_HIP(c, hipSetDevice(c->device));
    if (nt == 0) return LPM_OK;
    hipStream_t st = c->stream;
    LPM_CHECK(c->h_in.reserve(4 * size_t(nt) * sizeof(double)));
    double* d = c->h_in.as<double>();
    LPM_HIP(c, hipMemcpyAsync(d, tx, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + nt, ty, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + 2 * nt, tz, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_CHECK(check_tree_sphere(c, d, d + nt, d + 2 * nt, nt, "lpm_tree_eval_psi"));
    LPM_CHECK(tree_eval_psi_device(c, prec, order, theta, d, d + nt, d + 2 * nt, nt, d + 3 * nt));
    LPM_HIP(c, hipMemcpyAsync(psi, d + 3 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

int lpm_tree_eval_learned_dev(lpm_ctx* c, int prec, double theta_max, int margin, int max_order, double u_ref,
                              double delta, const double* tx, const double* ty, const double* tz, int64_t nt,
                              double* ux, double* uy, double* uz) {
    LPM_CHECK(check_learned(c, prec, theta_max, margin, max_order, u_ref, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_CHECK(check_tree_sphere(c, tx, ty, tz, nt, "lpm_tree_eval_learned"));
    return tree_eval_learned_device(c, prec, theta_max, margin, max_order, u_ref, delta, tx, ty, tz, nt, ux, uy, uz);
}

int lpm_tree_eval_learned(lpm_ctx* c, int prec, double theta_max, int margin, int max_order, double u_ref,
                          double delta, const double* tx, const double* ty, const double* tz, int64_t nt, double* ux,
                          double* uy, double* uz) {
    LPM_CHECK(check_learned(c, prec, theta_max, margin, max_order, u_ref, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    if (nt == 0)
        return tree_eval_learned_device(c, prec, theta_max, margin, max_order, u_ref, delta, tx, ty, tz, 0, ux, uy, uz);
    return with_staged(c, tx, ty, tz, nt, ux, uy, uz,
                       [&](const double* x, const double* y, const double* z, double* u, double* v, double* w) {
                           LPM_CHECK(check_tree_sphere(c, x, y, z, nt, "lpm_tree_eval_learned"));
                           return tree_eval_learned_device(c, prec, theta_max, margin, max_order, u_ref, delta, x, y,
                                                           z, nt, u, v, w);
                       });
}

int lpm_tree_labels(lpm_ctx* c, double theta_max, double delta, double u_ref, int64_t batch_stride,
                    double keep_rate, int frontier_levels, uint64_t seed, int64_t max_samples, const double* tx, const double* ty,
                    const double* tz, int64_t nt, float* features, float* errors, int32_t* ids,
                    int64_t* n_samples) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    LPM_CHECK(check_tree_for_batch(c, kLabelOrder));
    if (!tx || !ty || !tz || !features || !errors || !n_samples) return fail(c, LPM_ERR_NULL_PTR, "NULL argument");
    if (nt < 1 || nt > (int64_t(1) << 30) || batch_stride < 1 || max_samples < 1 || !(keep_rate > 0.0) ||
        !(u_ref > 0.0) || !(delta >= 0.0) || !(theta_max > 0.0))
        return fail(c, LPM_ERR_INVALID_ARG, "invalid label arguments");
    LPM_HIP(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    LPM_CHECK(c->h_in.reserve(3 * size_t(nt) * sizeof(double)));
    LPM_CHECK(c->h_out.reserve(size_t(max_samples) * (kIn + kLabelOrder + 2) * sizeof(float)));
    double* d = c->h_in.as<double>();
    float* df = c->h_out.as<float>();
    float* de = df + size_t(max_samples) * kIn;
    int* di = reinterpret_cast<int*>(de + size_t(max_samples) * kLabelOrder);
    LPM_HIP(c, hipMemcpyAsync(d, tx, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + nt, ty, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + 2 * nt, tz, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_CHECK(check_tree_sphere(c, d, d + nt, d + 2 * nt, nt, "lpm_tree_labels"));
    int64_t ns = 0;
    LPM_CHECK(tree_labels_device(c, theta_max, delta, u_ref, batch_stride, keep_rate, frontier_levels,
                                 uint32_t(seed ^ (seed >> 32)),
                                 max_samples, d, d + nt, d + 2 * nt, nt, df, de, di, &ns));
    if (ns > 0) {
        LPM_HIP(c, hipMemcpyAsync(features, df, size_t(ns) * kIn * sizeof(float), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipMemcpyAsync(errors, de, size_t(ns) * kLabelOrder * sizeof(float), hipMemcpyDeviceToHost, st));
        if (ids) LPM_HIP(c, hipMemcpyAsync(ids, di, size_t(ns) * 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    }
    LPM_HIP(c, hipStreamSynchronize(st));
    *n_samples = ns;
    return LPM_OK;
}

int lpm_tree_stats(lpm_ctx* c, int64_t* out) {
    if (!c || !out) return LPM_ERR_NULL_PTR;
    if (!c->stats_ptr) return fail(c, LPM_ERR_STATE, "no tree evaluation yet");
    unsigned long long h[kStats];
    LPM_CHECK(read_stats(c, h));
    for (int k = 0; k < kStats; ++k) out[k] = int64_t(h[k]);
    return LPM_OK;
}

}  // extern "C"
