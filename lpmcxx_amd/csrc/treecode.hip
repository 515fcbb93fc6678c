// Treecode moments and evaluation on gfx950 (TreeSum, src/LpmTreeSum.cpp).
//
// K7 tree_moments  per node, for every particle of the node (no
//                  moment-to-moment shifts, as in the reference):
//                  m_k[c] += (gamma_j y_j[c]) h^k, |k| < p, lex order of k,
//                  h = moment_sign * (y_j - centroid).  The reference uses
//                  h = centroid - y (moment_sign = -1), which flips the sign of
//                  every odd-order term relative to the coefficients; +1 is the
//                  consistent expansion.  Leaves: work item = (leaf, chunk of
//                  128 particles), the chunk's powers h^a staged in LDS, each
//                  multi-index summed by 256 / nterm thread groups over parts of
//                  the chunk, partials added in group and chunk order
//                  (deterministic).  Internal nodes: the children's moments
//                  translated and summed (upward pass, tree_m2m).
// K8-K11 tree_eval one wave per 64 Morton-sorted targets, depth-first over the
//                  node table with a wave-shared stack of (node, lane mask):
//                  every lane evaluates the reference MAC
//                  maxRadius / |t - c| < theta for itself; accepting lanes run
//                  the far field, the remaining lanes descend, and at a leaf
//                  they sum the kernel over the leaf's particles.  Each lane
//                  therefore visits exactly the nodes the reference recursion
//                  visits for that target, in the same order, and the MAC
//                  count is exact.
//                  Far field: coefficients of 1/(1 + delta^2 - t.c - t.h)
//                  a_k = |k|!/k! t^k / D^{|k|+1} by the reference's recurrence
//                  (each a_k from its last lex-order writer: k - e3, else
//                  k - e2, else k - e1), streamed with three running values so
//                  no coefficient array is stored, contracted with the node
//                  moments (wave-uniform scalar loads) as S = sum_k a_k m_k,
//                  then t x S.  FP64 throughout.
//                  Near field: LPM_F64 uses the reference's pair formula;
//                  LPM_F32C uses FP32 offsets from the leaf centroid,
//                  |d|^2/2 denominators and an FP32 per-leaf sum; LPM_F64F
//                  keeps the reference's denominator and zero rule, takes
//                  gamma / den from v_rcp_f64 refined to about an ulp and
//                  sums w_j (y_j - t) with FMA, applying t x once per
//                  target (t x y_j = t x (y_j - t)).
#include "lpm_tree.hpp"
#include "lpm_mlp.hpp"
#include "lpm_series.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <utility>
#include <vector>

namespace lpm {

namespace {

constexpr int kStack = 160;   // >= 7 * kMaxDepth + 1

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

// ---------------------------------------------------------------- moments
// Moments of NC components per multi-index: NC = 3, m_k = sum_j gamma_j y_j
// h_j^k (velocity); NC = 1, m_k = sum_j gamma_j h_j^k (streamfunction).

// lex index t of order p -> (k1, k2, k3)
__device__ inline void lex_decode(int p, int t, int& k1, int& k2, int& k3) {
    int r = t;
    k1 = 0;
    k2 = 0;
    while (r >= (p - k1) * (p - k1 + 1) / 2) { r -= (p - k1) * (p - k1 + 1) / 2; ++k1; }
    while (r >= p - k1 - k2) { r -= p - k1 - k2; ++k2; }
    k3 = r;
}

// The three stored forms of a node's moments: lex FP64 (the reference's
// order, exported and used by the reference-series traversal), graded FP64
// (Cartesian series) and graded FP32 scaled by rho^-|k| (rho = the node's box
// radius; FP32 far field).
struct MomOut {
    double* lex;
    double* graded;
    float* scaled;
};

// store component cc of lex term t of `node`; irho = 1 / rho
template <int NC>
__device__ inline void store_moment(const MomOut& o, int64_t node, int p, int nterm, int t, int cc, double v,
                                    double irho) {
    int k1, k2, k3;
    lex_decode(p, t, k1, k2, k3);
    const int64_t base = node * nterm * NC;
    const int64_t g = base + NC * graded_index(k1, k2, k3) + cc;
    o.lex[base + NC * t + cc] = v;
    o.graded[g] = v;
    double sc = 1.0;
    for (int d = k1 + k2 + k3; d > 0; --d) sc *= irho;
    o.scaled[g] = float(v * sc);
}

// Graded position and degree of every lex term |k| < P (packed g | deg << 16),
// for the kernels that know P at compile time.
template <int P>
struct LexInfo {
    static constexpr int NT = P * (P + 1) * (P + 2) / 6;
    int gd[NT] = {};
    constexpr LexInfo() {
        int t = 0;
        for (int a = 0; a < P; ++a)
            for (int b = 0; a + b < P; ++b)
                for (int c = 0; a + b + c < P; ++c) gd[t++] = graded_index(a, b, c) | ((a + b + c) << 16);
    }
};
template <int P>
inline constexpr LexInfo<P> kLexInfo{};

// store_moment with P known: ipow[d] = irho^d (d < P) formed by repeated
// multiplication, the same products as store_moment's loop
template <int NC, int P>
__device__ inline void store_moment_p(const MomOut& o, int64_t node, int t, int cc, double v, const double* ipow) {
    constexpr int NT = P * (P + 1) * (P + 2) / 6;
    const int gd = kLexInfo<P>.gd[t];
    const int64_t base = node * NT * NC;
    const int64_t g = base + NC * (gd & 0xffff) + cc;
    o.lex[base + NC * t + cc] = v;
    o.graded[g] = v;
    o.scaled[g] = float(v * ipow[gd >> 16]);
}

// ipow[d] = irho^d, d < P (one thread)
template <int P>
__device__ inline void inverse_powers(double irho, double* ipow) {
    double x = 1.0;
#pragma unroll
    for (int d = 0; d < P; ++d) {
        ipow[d] = x;
        x *= irho;
    }
}

constexpr int kMomThreads = 256;
// LDS row of the leaf moment kernel: a chunk's particles plus two pad
// doubles, so that rows read at the same particle by different lanes start
// four banks apart
constexpr int kMomRow = kChunk + 2;

// P2M of one leaf (or one leaf chunk), order P known at compile time: the
// particles' powers h^a (a < P, per axis) and gamma y (or gamma) staged in
// LDS as rows over up to kChunk particles at a time, then G groups of nterm
// threads (G = 256 / nterm, at least 1) each sum their term over a contiguous
// part of the staged particles, four per step (ds_read2 pairs of consecutive
// particles), in registers across the leaf's chunks; the group partials are
// added in group order.
// chunk_node == nullptr: one block per node, a leaf's particles staged chunk
// by chunk (all three forms of its moments are stored); otherwise one block
// per entry of the chunk list, whose partial goes to `part`.
template <int NC, int P>
__global__ __launch_bounds__(kMomThreads) void tree_moments_chunk(const int* __restrict__ chunk_node,
                                                                  const int* __restrict__ chunk_beg,
                                                                  const int* __restrict__ chunk_end,
                                                                  const int* __restrict__ nkids,
                                                                  const int* __restrict__ nbeg,
                                                                  const int* __restrict__ nend,
                                                                  const double4* __restrict__ ncent,
                                                                  const double4* __restrict__ spos, double sign,
                                                                  double* __restrict__ part, MomOut out) {
    constexpr int nterm = P * (P + 1) * (P + 2) / 6;
    constexpr int G = kMomThreads / nterm > 1 ? kMomThreads / nterm : 1;
    constexpr int KT = (G * nterm + kMomThreads - 1) / kMomThreads;  // tasks per thread
    extern __shared__ double lds[];
    double* pw = lds;                       // [3][P][kMomRow] powers
    double* gy = lds + 3 * P * kMomRow;     // [NC][kMomRow]   gamma * y, or gamma
    double* red = gy + NC * kMomRow;        // [G][nterm][NC]  group partials (G > 1)
    const int ch = blockIdx.x;
    const bool direct = chunk_node == nullptr;
    int node = ch, b, e;
    double4 c;
    if (direct) {
        const int nk = nkids[ch];
        b = nbeg[ch];
        e = nend[ch];
        c = ncent[ch];
        if (nk != 0) return;  // the whole block
    } else {
        node = chunk_node[ch];
        b = chunk_beg[ch];
        e = chunk_end[ch];
        c = ncent[node];
    }
    __shared__ double s_ipow[P];
    if (threadIdx.x == 0) inverse_powers<P>(1.0 / c.w, s_ipow);
    double m[KT][NC] = {};
    for (int c0 = b; c0 < e; c0 += kChunk) {
        const int cnt = min(kChunk, e - c0);
        if (c0 != b) __syncthreads();  // the previous chunk's rows are read
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {  // the sums read rows q < cnt only
            const double4 s = spos[c0 + q];
            const double yx = s.x, yy = s.y, yz = s.z, g = s.w;
            const double hx = sign * (s.x - c.x), hy = sign * (s.y - c.y), hz = sign * (s.z - c.z);
            double ax = 1.0, ay = 1.0, az = 1.0;
#pragma unroll
            for (int a = 0; a < P; ++a) {
                pw[a * kMomRow + q] = ax;
                pw[(P + a) * kMomRow + q] = ay;
                pw[(2 * P + a) * kMomRow + q] = az;
                ax *= hx; ay *= hy; az *= hz;
            }
            if (NC == 3) {
                gy[q] = g * yx;
                gy[kMomRow + q] = g * yy;
                gy[2 * kMomRow + q] = g * yz;
            } else {
                gy[q] = g;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int it = threadIdx.x + k * kMomThreads;
            if (it >= G * nterm) break;
            const int grp = it / nterm, t = it - grp * nterm;
            int k1, k2, k3;
            lex_decode(P, t, k1, k2, k3);
            const int q0 = grp * cnt / G, q1 = (grp + 1) * cnt / G;
            const double* wx = pw + k1 * kMomRow + q0;
            const double* wy = pw + (P + k2) * kMomRow + q0;
            const double* wz = pw + (2 * P + k3) * kMomRow + q0;
            const double* gq = gy + q0;
            int n = q1 - q0;
            for (; n >= 4; n -= 4, wx += 4, wy += 4, wz += 4, gq += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double pv = wx[u] * wy[u] * wz[u];
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) m[k][cc] = __builtin_fma(gq[cc * kMomRow + u], pv, m[k][cc]);
                }
            }
            for (int u = 0; u < n; ++u) {
                const double pv = wx[u] * wy[u] * wz[u];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) m[k][cc] = __builtin_fma(gq[cc * kMomRow + u], pv, m[k][cc]);
            }
        }
    }
    if (G > 1) {
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int it = threadIdx.x + k * kMomThreads;
            if (it >= G * nterm) break;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) red[it * NC + cc] = m[k][cc];
        }
        __syncthreads();
        for (int it = threadIdx.x; it < nterm * NC; it += blockDim.x) {
            double v = red[it];
            for (int grp = 1; grp < G; ++grp) v += red[grp * nterm * NC + it];
            if (direct) store_moment_p<NC, P>(out, node, it / NC, it % NC, v, s_ipow);
            else part[int64_t(ch) * nterm * NC + it] = v;
        }
    } else {
        __syncthreads();  // s_ipow
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int t = threadIdx.x + k * kMomThreads;
            if (t >= nterm) break;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                if (direct) store_moment_p<NC, P>(out, node, t, cc, m[k][cc], s_ipow);
                else part[(int64_t(ch) * nterm + t) * NC + cc] = m[k][cc];
            }
        }
    }
}

// nch blocks: one per chunk of the list cn/cb/ce, or (cn == nullptr) one per node
template <int NC, int P>
void launch_moments_chunk(hipStream_t st, int64_t nch, const int* cn, const int* cb, const int* ce, const int* nkids,
                          const int* nbeg, const int* nend, const double4* ncent, const double4* spos, double sign,
                          double* part, MomOut out) {
    constexpr int nterm = P * (P + 1) * (P + 2) / 6;
    constexpr int G = kMomThreads / nterm > 1 ? kMomThreads / nterm : 1;
    const size_t lds = (size_t(kMomRow) * (3 * P + NC) + (G > 1 ? size_t(G) * nterm * NC : 0)) * sizeof(double);
    hipLaunchKernelGGL((tree_moments_chunk<NC, P>), dim3(unsigned(nch)), dim3(kMomThreads), lds, st, cn, cb, ce,
                       nkids, nbeg, nend, ncent, spos, sign, part, out);
}

// chunk partials of every leaf summed in chunk order (leaves of more than one
// chunk); internal nodes are left to the upward pass
template <int NC>
__global__ __launch_bounds__(256) void tree_moments_reduce(int nn, int p, int nterm, const int* __restrict__ nkids,
                                                           const int* __restrict__ node_chunk0,
                                                           const int* __restrict__ node_nchunk,
                                                           const double4* __restrict__ ncent,
                                                           const double* __restrict__ part, MomOut out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(nn) * nterm * NC) return;
    const int node = int(i / (int64_t(nterm) * NC));
    if (nkids[node] != 0) return;
    const int rem = int(i - int64_t(node) * nterm * NC);
    const int c0 = node_chunk0[node], nc = node_nchunk[node];
    double s = 0.0;
    for (int k = 0; k < nc; ++k) s += part[(int64_t(c0 + k) * nterm) * NC + rem];
    store_moment<NC>(out, node, p, nterm, rem / NC, rem % NC, s, 1.0 / ncent[node].w);
}

// Upward pass for the internal nodes of one level (moments about the box
// centroid, h = sign (y - c)): with d = sign (c_child - c),
//   m_k(node) = sum_children sum_{j <= k} C(k, j) d^(k - j) m_j(child),
// the same sums as the particle-by-particle definition of the leaves,
// regrouped.  Since (h + d)^k factorizes over the axes, the translation is
// applied as three one-dimensional shifts (x, then y, then z).  Along one
// axis the terms with the other two exponents fixed form a line m[0..L-1]
// (L <= P), and m'[e] = sum_{a <= e} C(e, a) d^(e - a) m[a] is L - 1 sweeps
// of m[e] += d m[e - 1] (e from the top down), Pascal's rule, done in
// registers.  One block per node: up to G children staged in LDS at once,
// one thread per (child, line, component), the children then added to the
// node's sum in child order.  The node's source radius (max over children
// of |c_child - c| + r_child) is set on the way.
template <int P>
__device__ __forceinline__ int lex_at(int a, int b, int c) {  // lex position of (a, b, c), |k| < P
    return P * (P + 1) * (P + 2) / 6 - (P - a) * (P - a + 1) * (P - a + 2) / 6 + b * (P - a) - b * (b - 1) / 2 + c;
}

// line l (l < P (P + 1) / 2) of the pairs (u, v) with u + v < P, u major
template <int P>
__device__ __forceinline__ void line_pair(int l, int& u, int& v) {
    u = 0;
    while (l >= P - u) { l -= P - u; ++u; }
    v = l;
}

// children staged per round of tree_m2m: G S doubles plus the node's sum within 64 KB
template <int NC, int P>
constexpr int m2m_group() {
    constexpr int S = NC * P * (P + 1) * (P + 2) / 6;
    constexpr int g = (8192 - S - 32) / S;
    return g > 8 ? 8 : (g < 1 ? 1 : g);
}

template <int NC, int P, int AXIS>
__device__ __forceinline__ void shift_lines(double* A, const double* dd, int ng) {
    constexpr int S = NC * P * (P + 1) * (P + 2) / 6;
    constexpr int NL = P * (P + 1) / 2;
    for (int task = threadIdx.x; task < ng * NL * NC; task += blockDim.x) {
        const int k = task / (NL * NC), rem = task - k * NL * NC;
        const int l = rem / NC, cc = rem - l * NC;
        int u, v;
        line_pair<P>(l, u, v);
        const int L = P - u - v;
        const double d = dd[3 * k + AXIS];
        double* buf = A + k * S;
        // position of exponent e along AXIS (the other two are u, v in axis order)
        auto pos = [&](int e) {
            return AXIS == 0 ? lex_at<P>(e, u, v) : AXIS == 1 ? lex_at<P>(u, e, v) : lex_at<P>(u, v, e);
        };
        double m[P];
#pragma unroll
        for (int e = 0; e < P; ++e) m[e] = e < L ? buf[NC * pos(e) + cc] : 0.0;
#pragma unroll
        for (int r = 1; r < P; ++r)
#pragma unroll
            for (int e = P - 1; e >= r; --e)
                if (e < L) m[e] = __builtin_fma(d, m[e - 1], m[e]);
#pragma unroll
        for (int e = 0; e < P; ++e)
            if (e < L) buf[NC * pos(e) + cc] = m[e];
    }
}

template <int NC, int P>
__global__ __launch_bounds__(256) void tree_m2m(int first, double sign, const int* __restrict__ nfirst,
                                                const int* __restrict__ nkids, const double4* __restrict__ ncent,
                                                double* __restrict__ nrad, MomOut out) {
    constexpr int S = NC * P * (P + 1) * (P + 2) / 6;
    constexpr int G = m2m_group<NC, P>();
    __shared__ double acc[S];
    __shared__ double A[G * S];
    __shared__ double dd[3 * G];
    static_assert(3 * G >= P, "dd also holds the P inverse powers of the stores");
    const int i = first + blockIdx.x;
    const int nk = nkids[i];
    if (nk == 0) return;  // the whole block
    const double4 c = ncent[i];
    const int f = nfirst[i];
    for (int q = threadIdx.x; q < S; q += blockDim.x) acc[q] = 0.0;
    for (int c0 = 0; c0 < nk; c0 += G) {
        const int ng = min(G, nk - c0);
        __syncthreads();
        // children are consecutive nodes: their moments are one contiguous block
        const double* mc = out.lex + int64_t(f + c0) * S;
        for (int q = threadIdx.x; q < ng * S; q += blockDim.x) A[q] = mc[q];
        if (threadIdx.x < 3 * ng) {
            const int k = threadIdx.x / 3, ax = threadIdx.x - 3 * k;
            const double4 cc = ncent[f + c0 + k];
            dd[threadIdx.x] = sign * (ax == 0 ? cc.x - c.x : (ax == 1 ? cc.y - c.y : cc.z - c.z));
        }
        __syncthreads();
        shift_lines<NC, P, 0>(A, dd, ng);
        __syncthreads();
        shift_lines<NC, P, 1>(A, dd, ng);
        __syncthreads();
        shift_lines<NC, P, 2>(A, dd, ng);
        __syncthreads();
        for (int q = threadIdx.x; q < S; q += blockDim.x) {
            double v = acc[q];
            for (int k = 0; k < ng; ++k) v += A[k * S + q];
            acc[q] = v;
        }
    }
    if (threadIdx.x == 0) {
        inverse_powers<P>(1.0 / c.w, dd);  // dd is free now: it holds irho^d for the stores
        double rad = 0.0;
        for (int k = f; k < f + nk; ++k) {
            const double4 ck = ncent[k];
            const double dx = ck.x - c.x, dy = ck.y - c.y, dz = ck.z - c.z;
            rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz) + nrad[k]);
        }
        nrad[i] = rad;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < S; q += blockDim.x) store_moment_p<NC, P>(out, i, q / NC, q % NC, acc[q], dd);
}

template <int NC, int P>
void launch_m2m(hipStream_t st, int64_t f, int64_t cnt, double sign, const int* nfirst, const int* nkids,
                const double4* ncent, double* nrad, MomOut out) {
    hipLaunchKernelGGL((tree_m2m<NC, P>), dim3(unsigned(cnt)), dim3(256), 0, st, int(f), sign, nfirst, nkids, ncent,
                       nrad, out);
}

// Norms of the degree-0, 1 and 2 moment shells (graded positions 0, 1..3,
// 4..9) and the particle count of every node.
__global__ __launch_bounds__(256) void tree_node_stats(int nn, int p, int nterm, const double* __restrict__ momg,
                                                       const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                       float4* __restrict__ nstat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const double* m = momg + int64_t(i) * nterm * 3;
    double sh[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < nterm && t < 10; ++t) {
        const int d = t == 0 ? 0 : (t < 4 ? 1 : 2);
        sh[d] += m[3 * t] * m[3 * t] + m[3 * t + 1] * m[3 * t + 1] + m[3 * t + 2] * m[3 * t + 2];
    }
    nstat[i] = make_float4(float(sqrt(sh[0])), float(sqrt(sh[1])), float(sqrt(sh[2])), float(nend[i] - nbeg[i]));
}

// Strengths of the sorted particles (gamma[perm[q]] into spos[q].w), their
// FP32 offsets from their leaf's centroid (FP32C near field) and the leaf's
// source radius (largest distance of its particles from the centroid), one
// wave per leaf: every particle belongs to exactly one leaf.
__global__ __launch_bounds__(64) void tree_leaf_strengths(int nn, const int* __restrict__ nkids,
                                                           const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                           const double4* __restrict__ ncent,
                                                           const double* __restrict__ gamma,
                                                           const int* __restrict__ perm, double4* __restrict__ spos,
                                                           float4* __restrict__ soff, double* __restrict__ nrad) {
    const int i = blockIdx.x;
    if (i >= nn || nkids[i] != 0) return;  // the whole block
    const double4 c = ncent[i];
    double r2 = 0.0;
    for (int q = nbeg[i] + threadIdx.x; q < nend[i]; q += blockDim.x) {
        double4 s = spos[q];
        s.w = gamma[perm[q]];
        spos[q] = s;
        soff[q] = make_float4(float(s.x - c.x), float(s.y - c.y), float(s.z - c.z), float(s.w));
        const double dx = s.x - c.x, dy = s.y - c.y, dz = s.z - c.z;
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if (threadIdx.x == 0) nrad[i] = sqrt(r2);
}

// ---------------------------------------------------------------- target ordering
__global__ __launch_bounds__(256) void target_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, int64_t n, double x0, double y0,
                                                   double z0, double x1, double y1, double z1,
                                                   uint64_t* __restrict__ keys, int* __restrict__ vals) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // clamp into the root box: the order only groups nearby targets
    const double px = fmin(fmax(x[i], x0), x1), py = fmin(fmax(y[i], y0), y1), pz = fmin(fmax(z[i], z0), z1);
    double xl = x0, xh = x1, yl = y0, yh = y1, zl = z0, zh = z1;
    uint64_t key = 0;
    for (int l = 0; l < kMaxDepth; ++l) {
        const double mx = 0.5 * (xl + xh), my = 0.5 * (yl + yh), mz = 0.5 * (zl + zh);
        const int bx = px >= mx, by = py >= my, bz = pz >= mz;
        if (bx) xl = mx; else xh = mx;
        if (by) yl = my; else yh = my;
        if (bz) zl = mz; else zh = mz;
        key = (key << 3) | uint64_t(bx | (by << 1) | (bz << 2));
    }
    keys[i] = key;
    vals[i] = int(i);
}

// ---------------------------------------------------------------- target batches
// Number of leading equal octree levels of two 63-bit keys.
__device__ inline int key_lcp(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0) return kMaxDepth;
    return (62 - (63 - __clzll(x))) / 3;
}

// hist[l] += adjacent sorted keys with a common prefix of l levels.  Sorted
// keys mostly share deep prefixes, so the few busy bins would queue every
// lane's add: the lanes of a wave are counted per value by ballots (one LDS
// add per value and wave), and a grid of at most kHistBlocks blocks strides
// over the keys, so that few global adds reach the bins.
constexpr int kHistBlocks = 512;
__global__ __launch_bounds__(256) void batch_lcp_hist(const uint64_t* __restrict__ k, int64_t n,
                                                      unsigned* __restrict__ hist) {
    __shared__ unsigned h[kMaxDepth + 1];
    if (threadIdx.x <= kMaxDepth) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    // the loop bound is wave-uniform: every lane of a wave takes the same trips
    const int64_t base = int64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~63);
    for (int64_t w0 = base; w0 < n; w0 += stride) {
        const int64_t i = w0 + (threadIdx.x & 63);
        const int v = (i >= 1 && i < n) ? key_lcp(k[i - 1], k[i]) : -1;
        uint64_t left = __ballot(v >= 0);
        while (left) {  // one pass per distinct value in the wave
            const int u = __builtin_amdgcn_readlane(v, __builtin_ctzll(left));
            const uint64_t m = __ballot(v == u);
            if ((threadIdx.x & 63) == 0) atomicAdd(&h[u], unsigned(__popcll(m)));
            left &= ~m;
        }
    }
    __syncthreads();
    if (threadIdx.x <= kMaxDepth && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// The batching level L: the shallowest level with at least `want` occupied
// cells (hist[l] = adjacent sorted keys whose common prefix is l levels), and
// at most max_level, the key levels by which the keys are sorted.  Up to
// max_level every cell is one run of the sorted keys, so the cells of level L
// number fewer than 8 want (fewer than want when max_level caps L).  Below
// it, neighbouring keys are in input order and a cell would fall apart into
// runs of single targets.
__device__ inline int batching_level(const unsigned* hist, int64_t want, int max_level) {
    int L = max_level;
    int64_t cells = 1;
    for (int l = 0; l < max_level; ++l) {
        if (cells >= want) { L = l; break; }
        cells += hist[l];
    }
    return L;
}

// Top L levels of a 63-bit key (L <= kMaxDepth).
__device__ inline uint64_t key_prefix(uint64_t k, int L) { return L == 0 ? 0 : k >> (3 * (kMaxDepth - L)); }

// Batch starts of one tile of kBatchTile sorted positions, kBatchPer
// consecutive positions per thread: a position starts a batch when it lies a
// multiple of 64 after the start of its level-L cell.  The cell start carried
// into the tile is the first position with the tile's first prefix (prefixes
// of at most the sorted levels are non-decreasing), found by a 256-way search
// (a few rounds of one key per thread); inside the tile a running maximum of
// the cell starts continues it.  Returns the thread's flags (bit j: position
// first + j).
constexpr int kBatchPer = 8;
constexpr int kBatchTile = 256 * kBatchPer;
using BatchScan = hipcub::BlockScan<int, 256>;

__device__ inline unsigned batch_tile_flags(const uint64_t* __restrict__ k, int n, int L, int& first,
                                            BatchScan::TempStorage& tmp) {
    const int t0 = blockIdx.x * kBatchTile;
    first = t0 + threadIdx.x * kBatchPer;
    // the first position in [lo, hi] whose prefix is not below p0 (t0 is one);
    // every thread holds the same lo and hi, so the loop is uniform
    const uint64_t p0 = key_prefix(k[t0], L);
    int lo = 0, hi = t0;
    while (lo < hi) {
        const int step = (hi - lo + 255) / 256;
        const int m = lo + int(threadIdx.x) * step;
        // the probes below p0 are the first cnt (prefixes are non-decreasing)
        const int cnt = __syncthreads_count(m < hi && key_prefix(k[m], L) < p0);
        if (cnt == 0) {
            hi = lo;
        } else {
            hi = min(hi, lo + cnt * step);
            lo = lo + (cnt - 1) * step + 1;
        }
    }
    const int carry = lo;
    int st[kBatchPer];
    int last = -1;
    uint64_t prev = (first >= 1 && first <= n) ? key_prefix(k[first - 1], L) : 0;
#pragma unroll
    for (int j = 0; j < kBatchPer; ++j) {
        const int i = first + j;
        st[j] = -1;
        if (i < n) {
            const uint64_t cur = key_prefix(k[i], L);
            if (i == 0 || cur != prev) st[j] = i;
            prev = cur;
        }
        last = max(last, st[j]);
    }
    int run = 0;
    BatchScan(tmp).ExclusiveScan(last, run, carry, hipcub::Max());
    unsigned flags = 0;
#pragma unroll
    for (int j = 0; j < kBatchPer; ++j) {
        const int i = first + j;
        run = max(run, st[j]);
        if (i < n && ((i - run) & (kWave - 1)) == 0) flags |= 1u << j;
    }
    return flags;
}

// Batch count of every tile.
__global__ __launch_bounds__(256) void batch_tile_count(const uint64_t* __restrict__ k, int n, int64_t want,
                                                        int max_level, const BatchHeader* __restrict__ hdr,
                                                        int* __restrict__ tcnt) {
    __shared__ BatchScan::TempStorage tmp;
    const int L = batching_level(hdr->hist, want, max_level);
    int first;
    const unsigned flags = batch_tile_flags(k, n, L, first, tmp);
    __syncthreads();  // tmp is reused
    int ex = 0, tot = 0;
    BatchScan(tmp).ExclusiveSum(__popc(flags), ex, tot);
    if (threadIdx.x == 0) tcnt[blockIdx.x] = tot;
}

// Exclusive sums of the tile counts (in place), the batch count, and
// bstart[nb] = n.  One block.
__global__ __launch_bounds__(256) void batch_tile_scan(int ntile, int n, BatchHeader* __restrict__ hdr,
                                                       int* __restrict__ tcnt, int* __restrict__ bs) {
    __shared__ BatchScan::TempStorage tmp;
    int run = 0;
    for (int b0 = 0; b0 < ntile; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const int v = b < ntile ? tcnt[b] : 0;
        int ex = 0, sum = 0;
        BatchScan(tmp).ExclusiveSum(v, ex, sum);
        if (b < ntile) tcnt[b] = run + ex;
        run += sum;
        __syncthreads();  // tmp is reused
    }
    if (threadIdx.x == 0) {
        hdr->nsel = run;
        bs[run] = n;
    }
}

// Writes the batch starts of every tile at its offset.
__global__ __launch_bounds__(256) void batch_tile_emit(const uint64_t* __restrict__ k, int n, int64_t want,
                                                       int max_level, const BatchHeader* __restrict__ hdr,
                                                       const int* __restrict__ tcnt, int* __restrict__ bs) {
    __shared__ BatchScan::TempStorage tmp;
    const int L = batching_level(hdr->hist, want, max_level);
    int first;
    const unsigned flags = batch_tile_flags(k, n, L, first, tmp);
    __syncthreads();  // tmp is reused
    int ex = 0;
    BatchScan(tmp).ExclusiveSum(__popc(flags), ex);
    int o = tcnt[blockIdx.x] + ex;
#pragma unroll
    for (int j = 0; j < kBatchPer; ++j)
        if (flags & (1u << j)) bs[o++] = first + j;
}

// ---------------------------------------------------------------- evaluation
__device__ inline uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return (uint64_t(hi) << 32) | lo;
}

// near-field forms of tree_eval
constexpr int kNearF64 = 0;   // the reference's pair formula, IEEE division
constexpr int kNearF32 = 1;   // FP32 offsets from the leaf centroid, |d|^2/2 denominators
constexpr int kNearF64F = 2;  // the reference's denominator, refined hardware reciprocal, t x sum w (y - t)

template <int P, int NEAR>
__global__ __launch_bounds__(256) void tree_eval(const double* __restrict__ tx, const double* __restrict__ ty,
                                                 const double* __restrict__ tz, const int* __restrict__ tperm,
                                                 int64_t nt, const double4* __restrict__ ncent,
                                                 const int* __restrict__ nfirst, const int* __restrict__ nkids,
                                                 const int* __restrict__ nbeg, const int* __restrict__ nend,
                                                 const double4* __restrict__ spos, const float4* __restrict__ soff,
                                                 const double* __restrict__ mom, const double* __restrict__ momg,
                                                 int p, int nterm, double theta,
                                                 double delta, double R2, double d2, float eps2v, double mult,
                                                 double* __restrict__ ux, double* __restrict__ uy,
                                                 double* __restrict__ uz, unsigned long long* __restrict__ mac_total) {
    // mac_total[0] += accepted (target, node) pairs, mac_total[1] += near-field (target, source) pairs
    __shared__ int s_node[4][kStack];
    __shared__ uint64_t s_mask[4][kStack];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t it = (int64_t(blockIdx.x) * 4 + w) * 64 + lane;
    const bool valid = it < nt;
    const int orig = valid ? tperm[it] : 0;
    const double t0 = valid ? tx[orig] : 0.0, t1 = valid ? ty[orig] : 0.0, t2 = valid ? tz[orig] : 0.0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;  // kNearF64F: sum_j w_j (y_j - t) over the near sources
    unsigned mac = 0;
    unsigned long long npairs = 0;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    int sp = 0;
    s_node[w][0] = 0;
    s_mask[w][0] = vmask;
    sp = 1;
    while (sp > 0) {
        --sp;
        const int n = __builtin_amdgcn_readfirstlane(s_node[w][sp]);
        const uint64_t m = uniform64(s_mask[w][sp]);
        const bool act = (m >> lane) & 1ull;
        const double4 c = ncent[n];
        bool acc = false;
        if (act) {
            const double dist = sqrt((t0 - c.x) * (t0 - c.x) + (t1 - c.y) * (t1 - c.y) + (t2 - c.z) * (t2 - c.z));
            acc = c.w / dist < theta;
        }
        const uint64_t am = __ballot(acc);
        if (am) {
            if (acc) {
                double S0, S1, S2;
                if (P == 0) far_reference(p, t0, t1, t2, c.x, c.y, c.z, d2, mom + int64_t(n) * nterm * 3, S0, S1, S2);
                else far_cartesian<(P > 0 ? P : 1)>(t0 - c.x, t1 - c.y, t2 - c.z, 2.0 * d2,
                                                    momg + int64_t(n) * nterm * 3, S0, S1, S2);
                s0 += t1 * S2 - t2 * S1;
                s1 += t2 * S0 - t0 * S2;
                s2 += t0 * S1 - t1 * S0;
                ++mac;
            }
        }
        const uint64_t rest = m & ~am;
        if (rest == 0) continue;
        const int nk = nkids[n];
        if (nk == 0) {
            const bool mine = (rest >> lane) & 1ull;
            const int b = nbeg[n], e = nend[n];
            npairs += mine ? (unsigned long long)(e - b) : 0ull;
            if (NEAR == kNearF32) {
                const float q0 = float(t0 - c.x), q1 = float(t1 - c.y), q2 = float(t2 - c.z);
                float a0 = 0.f, a1 = 0.f, a2 = 0.f;
                if (mine) {
                    for (int j = b; j < e; ++j) {
                        const float4 o = soff[j];
                        const float dx = q0 - o.x, dy = q1 - o.y, dz = q2 - o.z;
                        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2v)));
                        // the reference's |den| <= 1e-13 rule: |d|^2 + 2 delta^2 <= 2e-13 contributes 0
                        const float wgt = (r2 > 2e-13f) ? o.w * __builtin_amdgcn_rcpf(r2) : 0.f;
                        a0 = fmaf(wgt, dx, a0);
                        a1 = fmaf(wgt, dy, a1);
                        a2 = fmaf(wgt, dz, a2);
                    }
                    // (t x y) gamma / den = -2 t x d gamma / (|d|^2 + 2 delta^2)
                    const double b0 = a0, b1 = a1, b2 = a2;
                    s0 -= 2.0 * (t1 * b2 - t2 * b1);
                    s1 -= 2.0 * (t2 * b0 - t0 * b2);
                    s2 -= 2.0 * (t0 * b1 - t1 * b0);
                }
            } else if (NEAR == kNearF64F) {
                if (mine) {
                    for (int j = b; j < e; ++j) {
                        const double4 s = spos[j];
                        const double dot = t0 * s.x + t1 * s.y + t2 * s.z;
                        const double den = R2 - dot + d2;
                        const double wgt = fabs(den) > 1.0e-13 ? weight_f64f(s.w, den) : 0.0;
                        e0 = __builtin_fma(wgt, s.x - t0, e0);
                        e1 = __builtin_fma(wgt, s.y - t1, e1);
                        e2 = __builtin_fma(wgt, s.z - t2, e2);
                    }
                }
            } else if (mine) {
                for (int j = b; j < e; ++j) {
                    const double4 s = spos[j];
                    const double cx = t1 * s.z - t2 * s.y;
                    const double cy = t2 * s.x - t0 * s.z;
                    const double cz = t0 * s.y - t1 * s.x;
                    const double den = R2 - (t0 * s.x + t1 * s.y + t2 * s.z) + d2;
                    if (fabs(den) > 1.0e-13) {
                        const double inv = 1.0 / den;
                        s0 += (cx * inv) * s.w;
                        s1 += (cy * inv) * s.w;
                        s2 += (cz * inv) * s.w;
                    }
                }
            }
        } else {
            const int f = nfirst[n];
            for (int q = nk - 1; q >= 0; --q) {
                s_node[w][sp] = f + q;
                s_mask[w][sp] = rest;
                ++sp;
            }
        }
    }
    if (NEAR == kNearF64F) {
        s0 += t1 * e2 - t2 * e1;
        s1 += t2 * e0 - t0 * e2;
        s2 += t0 * e1 - t1 * e0;
    }
    if (valid) {
        ux[orig] = s0 * mult;
        uy[orig] = s1 * mult;
        uz[orig] = s2 * mult;
    }
    // wave-reduce the acceptance count, one atomic per wave
    unsigned long long tot = mac;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
    for (int o = 32; o > 0; o >>= 1) npairs += __shfl_xor(npairs, o, 64);
    if (lane == 0) {
        unsigned long long* m = mac_total + (blockIdx.x % kStatSlots) * kStatStride;
        atomicAdd(m, tot);
        atomicAdd(m + 1, npairs);
    }
}

}  // namespace

// P2M over the build's leaf chunks, chunk reduction (leaves of several
// chunks), upward pass level by level; every kernel stores the lex, graded
// and scaled FP32 forms of the moments it makes.  NC components.
// between_passes (may be empty) is called after the leaf moments are
// enqueued and before the upward pass (the stereographic moments go to the
// second stream there, behind the first stream's longest kernel).
template <int NC, class Hook>
static int moments_pass(lpm_ctx* c, int p, int sign, DevBuf& mom, DevBuf& momg, DevBuf& mom32,
                        Hook between_passes) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    const int nterm = nterms(p);
    const int64_t nn = t->nn;
    // chunk list of the leaves, made by the build when a leaf holds more than
    // one chunk (internal nodes come from the upward pass)
    const int64_t nch = t->nch;
    // one block per node when every leaf is one chunk or the leaves are small
    // enough for a block to stage chunk after chunk
    const int direct = (nch == t->nleaves || (t->leaf_max > 0 && t->leaf_max <= kDirectLeafMax)) ? 1 : 0;
    const int* d_cn = direct ? nullptr : t->chunks.as<int>();
    const int* d_cb = direct ? nullptr : d_cn + nch;
    const int* d_ce = direct ? nullptr : d_cb + nch;
    const int* d_n0 = direct ? nullptr : d_ce + nch;
    const int* d_nc = direct ? nullptr : d_n0 + nn;
    if (!direct) LPM_CHECK(t->mpart.reserve(size_t(std::max<int64_t>(nch, 1)) * nterm * NC * sizeof(double)));
    LPM_CHECK(mom.reserve(size_t(nn) * nterm * NC * sizeof(double)));
    LPM_CHECK(momg.reserve(size_t(nn) * nterm * NC * sizeof(double)));
    LPM_CHECK(mom32.reserve(size_t(nn) * nterm * NC * sizeof(float)));
    const MomOut out{mom.as<double>(), momg.as<double>(), mom32.as<float>()};
    const int64_t nblk = direct ? nn : nch;
    if (nch > 0) {
        switch (p) {
#define LPM_CASE(PP)                                                                                          \
    case PP:                                                                                                  \
        launch_moments_chunk<NC, PP>(st, nblk, d_cn, d_cb, d_ce, t->nkids.as<int>(), t->nbeg.as<int>(),       \
                                     t->nend.as<int>(), t->ncent.as<double4>(), t->spos.as<double4>(),        \
                                     double(sign), t->mpart.as<double>(), out);                              \
        break;
            LPM_CASE(1) LPM_CASE(2) LPM_CASE(3) LPM_CASE(4) LPM_CASE(5) LPM_CASE(6)
            LPM_CASE(7) LPM_CASE(8) LPM_CASE(9) LPM_CASE(10) LPM_CASE(11) LPM_CASE(12)
#undef LPM_CASE
            default: return fail(c, LPM_ERR_INVALID_ARG, "moment order %d unsupported", p);
        }
    }
    if (!direct)
        hipLaunchKernelGGL(tree_moments_reduce<NC>, dim3(blocks(nn * nterm * NC)), dim3(256), 0, st, int(nn), p, nterm,
                           t->nkids.as<int>(), d_n0, d_nc, t->ncent.as<double4>(), t->mpart.as<double>(), out);
    LPM_CHECK(between_passes());
    // upward pass, deepest internal level first
    for (int l = int(t->level_start.size()) - 3; l >= 0; --l) {
        const int64_t f = t->level_start[l], cnt = t->level_start[l + 1] - f;
        if (cnt <= 0) continue;
        switch (p) {
#define LPM_CASE(PP)                                                                                      \
    case PP:                                                                                              \
        launch_m2m<NC, PP>(st, f, cnt, double(sign), t->nfirst.as<int>(), t->nkids.as<int>(),             \
                           t->ncent.as<double4>(), t->nrad.as<double>(), out);                            \
        break;
            LPM_CASE(1) LPM_CASE(2) LPM_CASE(3) LPM_CASE(4) LPM_CASE(5) LPM_CASE(6)
            LPM_CASE(7) LPM_CASE(8) LPM_CASE(9) LPM_CASE(10) LPM_CASE(11) LPM_CASE(12)
#undef LPM_CASE
            default: return fail(c, LPM_ERR_INVALID_ARG, "moment order %d unsupported", p);
        }
    }
    return LPM_OK;
}

static int set_strengths(lpm_ctx* c, const double* gamma) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    LPM_CHECK(t->soff32.reserve(size_t(std::max<int64_t>(t->ns, 1)) * sizeof(float4)));
    hipLaunchKernelGGL(tree_leaf_strengths, dim3(unsigned(t->nn)), dim3(64), 0, st, int(t->nn), t->nkids.as<int>(),
                       t->nbeg.as<int>(), t->nend.as<int>(), t->ncent.as<double4>(), gamma, t->perm.as<int>(),
                       t->spos.as<double4>(), t->soff32.as<float4>(), t->nrad.as<double>());
    return LPM_OK;
}

// Cartesian order of the moments made under LPM_OPT_STEREO 2: the degrees 0 to
// 2 that the order model's moment-shell features read.
constexpr int kFeatureOrder = 3;

int tree_moments_device(lpm_ctx* c, const double* gamma, int p_far, int sign) {
    Tree* t = c->tree;
    LPM_HIP(c, join_aux(c));
    // stereographic only: the far field of order p_far comes from the
    // stereographic moments, the Cartesian ones serve the features
    const int p = c->stereo == 2 ? std::min(p_far, kFeatureOrder) : p_far;
    const int nterm = nterms(p);
    LPM_CHECK(t->nstat.reserve(size_t(t->nn) * sizeof(float4)));
    LPM_CHECK(set_strengths(c, gamma));
    t->stereo_terms = 0;
    t->stereo_order = 0;
    if (c->stereo && !c->aux) {
        LPM_HIP(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        LPM_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        LPM_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    // the stereographic moments on the second stream, beside the Cartesian
    // ones: forked once the strengths are set, enqueued after the leaf
    // moments so that those start first
    if (c->stereo) LPM_HIP(c, hipEventRecord(c->ev_fork, c->stream));
    LPM_CHECK(moments_pass<3>(c, p, sign, t->mom, t->momg, t->mom32, [&]() -> int {
        if (!c->stereo) return LPM_OK;
        LPM_HIP(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
        LPM_CHECK(tree_stereo_moments(c, p_far, c->aux));
        LPM_HIP(c, hipEventRecord(c->ev_join, c->aux));
        return LPM_OK;
    }));
    hipLaunchKernelGGL(tree_node_stats, dim3(blocks(t->nn)), dim3(256), 0, c->stream, int(t->nn), p, nterm,
                       t->momg.as<double>(), t->nbeg.as<int>(), t->nend.as<int>(), t->nstat.as<float4>());
    LPM_HIP(c, hipGetLastError());
    // only the learned evaluation reads the stereographic data: it joins the
    // second stream after its target batching, which runs beside it
    if (c->stereo) c->aux_pending = true;
    t->order = p;
    t->moment_sign = sign;
    return LPM_OK;
}

int tree_moments_psi_device(lpm_ctx* c, const double* gamma, int p) {
    Tree* t = c->tree;
    LPM_HIP(c, join_aux(c));
    LPM_CHECK(set_strengths(c, gamma));
    LPM_CHECK(moments_pass<1>(c, p, 1, t->pmom, t->pmomg, t->pmom32, []() -> int { return LPM_OK; }));
    LPM_HIP(c, hipGetLastError());
    t->psi_order = p;
    return LPM_OK;
}

constexpr int kTargetLevels = 12;  // key levels that order the targets

int sort_targets_for_tree(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                          const int** tperm, int* levels, const uint64_t** keys) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    const int64_t na = std::max<int64_t>(nt, 1);
    LPM_CHECK(c->ds_tgt.reserve(size_t(na) * (sizeof(uint64_t) * 2 + sizeof(int) * 2) + kStatBytes + kHeaderBytes +
                                128));
    uint64_t* k_in = c->ds_tgt.as<uint64_t>();
    uint64_t* k_out = k_in + na;
    int* v_in = reinterpret_cast<int*>(k_out + na);
    int* v_out = v_in + na;
    const uintptr_t sa = (reinterpret_cast<uintptr_t>(v_out + na) + 127) & ~uintptr_t(127);
    c->stats_ptr = reinterpret_cast<unsigned long long*>(sa);
    // the counters and the batch header in one fill
    LPM_HIP(c, hipMemsetAsync(c->stats_ptr, 0, kStatBytes + kHeaderBytes, st));
    *tperm = v_out;
    if (keys) *keys = k_out;
    if (levels) *levels = kTargetLevels;
    if (nt == 0) return LPM_OK;
    if (nt == t->ns && tx == t->sx && ty == t->sy && tz == t->sz) {
        // the targets are the tree's sources: the build's keys are sorted already
        if (levels) *levels = t->sorted_levels;
        *tperm = t->kperm.as<int>();
        if (keys) *keys = t->keys.as<uint64_t>();
        return LPM_OK;
    }
    hipLaunchKernelGGL(target_keys, dim3(blocks(nt)), dim3(256), 0, st, tx, ty, tz, nt, t->root_box[0], t->root_box[1],
                       t->root_box[2], t->root_box[3], t->root_box[4], t->root_box[5], k_in, v_in);
    // the top kTargetLevels levels order the targets (cells of about 5e-4 on
    // the unit sphere); the batch kernels only compare neighbouring keys, so
    // lower levels left unsorted can only split a batch, never mix two far apart
    const int begin_bit = 3 * (kMaxDepth - kTargetLevels);
    size_t bytes = 0;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in, v_out, int(nt), begin_bit,
                                                  3 * kMaxDepth, st));
    LPM_CHECK(t->sort_tmp.reserve(bytes + 256));
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(t->sort_tmp.ptr, bytes, k_in, k_out, v_in, v_out, int(nt), begin_bit,
                                                  3 * kMaxDepth, st));
    return LPM_OK;
}

int tree_target_batches(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                        const int** tperm, int** bstart, int64_t* nb_bound, const int** nb_dev) {
    int sorted_levels = 0;
    const uint64_t* keys = nullptr;
    LPM_CHECK(sort_targets_for_tree(c, tx, ty, tz, nt, tperm, &sorted_levels, &keys));
    *nb_bound = 0;
    BatchHeader* hdr = batch_header(c->stats_ptr);
    *nb_dev = &hdr->nsel;
    if (nt == 0) return LPM_OK;
    hipStream_t st = c->stream;
    // workspace: per-tile batch counts (then their exclusive sums) | bstart[nt + 1]
    const int ntile = int((nt + kBatchTile - 1) / kBatchTile);
    const size_t off_b = (size_t(ntile) * sizeof(int) + 255) & ~size_t(255);
    LPM_CHECK(c->tb_work.reserve(off_b + size_t(nt + 1) * sizeof(int)));
    int* tcnt = c->tb_work.as<int>();
    int* bs = reinterpret_cast<int*>(c->tb_work.as<char>() + off_b);
    const int64_t want = std::max<int64_t>(1, nt / c->batch_cell_targets);
    hipLaunchKernelGGL(batch_lcp_hist, dim3(unsigned(std::min<int64_t>(kHistBlocks, blocks(nt)))), dim3(256), 0, st,
                       keys, nt, hdr->hist);
    hipLaunchKernelGGL(batch_tile_count, dim3(unsigned(ntile)), dim3(256), 0, st, keys, int(nt), want, sorted_levels,
                       static_cast<const BatchHeader*>(hdr), tcnt);
    hipLaunchKernelGGL(batch_tile_scan, dim3(1), dim3(256), 0, st, ntile, int(nt), hdr, tcnt, bs);
    hipLaunchKernelGGL(batch_tile_emit, dim3(unsigned(ntile)), dim3(256), 0, st, keys, int(nt), want, sorted_levels,
                       static_cast<const BatchHeader*>(hdr), static_cast<const int*>(tcnt), bs);
    LPM_HIP(c, hipGetLastError());
    // every level-L cell of s targets makes ceil(s / 64) batches; the level
    // above L has fewer than `want` cells, so L has fewer than 8 want (L is
    // never below the sorted key levels, where that would not hold)
    *bstart = bs;
    *nb_bound = std::min<int64_t>(nt, (nt + kWave - 1) / kWave + 8 * want);
    return LPM_OK;
}

int check_tree_sphere(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                      const char* what) {
    const Tree* t = c->tree;
    if (t->ns > 0 && !r2_on_sphere(t->r2min, t->r2max, 1.0))
        return fail(c, LPM_ERR_INVALID_ARG,
                    "%s: this form needs the tree's sources on the unit sphere (|x|^2 in [%.17g, %.17g]); use "
                    "LPM_F64 with LPM_SERIES_REFERENCE for other points", what, t->r2min, t->r2max);
    if (nt > 0 && !(nt == t->ns && tx == t->sx && ty == t->sy && tz == t->sz))
        return check_on_sphere(c, tx, ty, tz, nt, 1.0, what);
    return LPM_OK;
}

int read_stats(lpm_ctx* c, unsigned long long* out) {
    unsigned long long h[kStatSlots * kStatStride];
    LPM_HIP(c, hipMemcpyAsync(h, c->stats_ptr, kStatBytes, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < kStats; ++k) {
        out[k] = 0;
        for (int s = 0; s < kStatSlots; ++s) out[k] += h[s * kStatStride + k];
    }
    return LPM_OK;
}

int read_batch_count(lpm_ctx* c, const int* nb_dev, int64_t* nb) {
    LPM_HIP(c, hipMemcpyAsync(c->pin, nb_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    *nb = c->pin[0];
    return LPM_OK;
}

int tree_eval_device(lpm_ctx* c, int prec, int series, double theta, double delta, const double* tx,
                     const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz,
                     int64_t* mac_out, double radius) {
    Tree* t = c->tree;
    hipStream_t st = c->stream;
    const int* v_out = nullptr;
    LPM_CHECK(sort_targets_for_tree(c, tx, ty, tz, nt, &v_out));
    if (nt == 0) { if (mac_out) *mac_out = 0; return LPM_OK; }
    unsigned long long* d_mac = c->stats_ptr;
    const double pi = 3.1415926535897932384626433832795027975;
    const double mult = -1.0 / (4.0 * pi * radius);
    const float eps2v = float(2.0 * delta * delta);
    const int p = t->order, nterm = nterms(p);
    const unsigned nb = unsigned((nt + 255) / 256);
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, tx, ty, tz, v_out, nt, t->ncent.as<double4>(),
                           t->nfirst.as<int>(), t->nkids.as<int>(), t->nbeg.as<int>(), t->nend.as<int>(),
                           t->spos.as<double4>(), t->soff32.as<float4>(), t->mom.as<double>(), t->momg.as<double>(), p, nterm, theta, delta,
                           radius * radius, delta * delta, eps2v, mult, ux, uy, uz, d_mac);
    };
    mark_begin(c);
    if (series == LPM_SERIES_REFERENCE) {
        if (prec == LPM_F32C) args(tree_eval<0, kNearF32>);
        else if (prec == LPM_F64F) args(tree_eval<0, kNearF64F>);
        else args(tree_eval<0, kNearF64>);
    } else {
        switch (p) {
#define LPM_CASE(PP)                                              \
    case PP:                                                      \
        if (prec == LPM_F32C) args(tree_eval<PP, kNearF32>);      \
        else if (prec == LPM_F64F) args(tree_eval<PP, kNearF64F>); \
        else args(tree_eval<PP, kNearF64>);                       \
        break;
            LPM_CASE(1) LPM_CASE(2) LPM_CASE(3) LPM_CASE(4) LPM_CASE(5) LPM_CASE(6)
            LPM_CASE(7) LPM_CASE(8) LPM_CASE(9) LPM_CASE(10) LPM_CASE(11) LPM_CASE(12)
#undef LPM_CASE
            default: return fail(c, LPM_ERR_INVALID_ARG, "series order %d unsupported", p);
        }
    }
    mark_end(c);
    LPM_HIP(c, hipGetLastError());
    if (mac_out) {
        unsigned long long h[kStats];
        LPM_CHECK(read_stats(c, h));
        *mac_out = int64_t(h[0]);
    }
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

extern "C" {

static int check_moments(lpm_ctx* c, const double* gamma, int p, int sign) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (!c->tree) return fail(c, LPM_ERR_STATE, "lpm_tree_moments before lpm_tree_build");
    if (p < 1 || p > 12) return fail(c, LPM_ERR_INVALID_ARG, "series order must be in [1, 12]");
    if (sign != 1 && sign != -1) return fail(c, LPM_ERR_INVALID_ARG, "moment_sign must be +1 or -1");
    if (c->tree->ns > 0 && !gamma) return fail(c, LPM_ERR_NULL_PTR, "gamma is NULL");
    return LPM_OK;
}

int lpm_tree_moments_dev(lpm_ctx* c, const double* gamma, int p, int sign) {
    LPM_CHECK(check_moments(c, gamma, p, sign));
    LPM_HIP(c, hipSetDevice(c->device));
    return tree_moments_device(c, gamma, p, sign);
}

int lpm_tree_moments(lpm_ctx* c, const double* gamma, int p, int sign) {
    LPM_CHECK(check_moments(c, gamma, p, sign));
    LPM_HIP(c, hipSetDevice(c->device));
    const int64_t n = std::max<int64_t>(c->tree->ns, 1);
    LPM_CHECK(c->h_out.reserve(n * sizeof(double)));
    if (c->tree->ns > 0)
        LPM_HIP(c, hipMemcpyAsync(c->h_out.ptr, gamma, c->tree->ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    LPM_CHECK(tree_moments_device(c, c->h_out.as<double>(), p, sign));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_tree_moments_psi_dev(lpm_ctx* c, const double* gamma, int p) {
    LPM_CHECK(check_moments(c, gamma, p, 1));
    LPM_HIP(c, hipSetDevice(c->device));
    return tree_moments_psi_device(c, gamma, p);
}

int lpm_tree_moments_psi(lpm_ctx* c, const double* gamma, int p) {
    LPM_CHECK(check_moments(c, gamma, p, 1));
    LPM_HIP(c, hipSetDevice(c->device));
    const int64_t n = std::max<int64_t>(c->tree->ns, 1);
    LPM_CHECK(c->h_out.reserve(n * sizeof(double)));
    if (c->tree->ns > 0)
        LPM_HIP(c, hipMemcpyAsync(c->h_out.ptr, gamma, c->tree->ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    LPM_CHECK(tree_moments_psi_device(c, c->h_out.as<double>(), p));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_tree_last_stats(lpm_ctx* c, int64_t* far_pairs, int64_t* near_pairs) {
    if (!c) return LPM_ERR_NULL_PTR;
    if (!c->stats_ptr) return fail(c, LPM_ERR_STATE, "no tree evaluation yet");
    unsigned long long h[kStats];
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_CHECK(read_stats(c, h));
    if (far_pairs) *far_pairs = int64_t(h[0]);
    if (near_pairs) *near_pairs = int64_t(h[1]);
    return LPM_OK;
}

int lpm_tree_export_moments(lpm_ctx* c, double* out) {
    if (!c) return LPM_ERR_NULL_PTR;
    if (!c->tree || c->tree->order == 0) return fail(c, LPM_ERR_STATE, "no moments computed");
    if (!out) return fail(c, LPM_ERR_NULL_PTR, "output is NULL");
    LPM_HIP(c, hipSetDevice(c->device));
    const size_t n = size_t(c->tree->nn) * nterms(c->tree->order) * 3;
    LPM_HIP(c, hipMemcpyAsync(out, c->tree->mom.ptr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

static int check_eval(lpm_ctx* c, int prec, int series, double theta, double delta, const double* tx,
                      const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (!c->tree) return fail(c, LPM_ERR_STATE, "lpm_tree_eval before lpm_tree_build");
    if (c->tree->order == 0) return fail(c, LPM_ERR_STATE, "lpm_tree_eval before lpm_tree_moments");
    if (prec != LPM_F64 && prec != LPM_F32C && prec != LPM_F64F)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid precision");
    if (series != LPM_SERIES_REFERENCE && series != LPM_SERIES_CARTESIAN)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid series %d", series);
    if (series == LPM_SERIES_CARTESIAN && c->tree->moment_sign != 1)
        return fail(c, LPM_ERR_INVALID_ARG, "the Cartesian series needs moments with moment_sign = +1");
    if (nt < 0 || nt > (int64_t(1) << 30)) return fail(c, LPM_ERR_INVALID_ARG, "invalid target count");
    if (!(delta >= 0.0) || !std::isfinite(theta)) return fail(c, LPM_ERR_INVALID_ARG, "invalid theta/delta");
    if (nt > 0 && (!tx || !ty || !tz || !ux || !uy || !uz)) return fail(c, LPM_ERR_NULL_PTR, "NULL target/output");
    return LPM_OK;
}

int lpm_tree_eval_dev(lpm_ctx* c, int prec, int series, double theta, double delta, const double* tx,
                      const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz,
                      int64_t* mac_count) {
    LPM_CHECK(check_eval(c, prec, series, theta, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    if (prec == LPM_F32C || series == LPM_SERIES_CARTESIAN)
        LPM_CHECK(check_tree_sphere(c, tx, ty, tz, nt, "lpm_tree_eval"));
    return tree_eval_device(c, prec, series, theta, delta, tx, ty, tz, nt, ux, uy, uz, mac_count, 1.0);
}

int lpm_tree_eval(lpm_ctx* c, int prec, int series, double theta, double delta, const double* tx, const double* ty,
                  const double* tz, int64_t nt, double* ux, double* uy, double* uz, int64_t* mac_count) {
    LPM_CHECK(check_eval(c, prec, series, theta, delta, tx, ty, tz, nt, ux, uy, uz));
    LPM_HIP(c, hipSetDevice(c->device));
    if (nt == 0) { if (mac_count) *mac_count = 0; return LPM_OK; }
    LPM_CHECK(c->h_in.reserve(6 * size_t(nt) * sizeof(double)));
    double* d = c->h_in.as<double>();
    hipStream_t st = c->stream;
    LPM_HIP(c, hipMemcpyAsync(d, tx, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + nt, ty, nt * sizeof(double), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + 2 * nt, tz, nt * sizeof(double), hipMemcpyHostToDevice, st));
    if (prec == LPM_F32C || series == LPM_SERIES_CARTESIAN)
        LPM_CHECK(check_tree_sphere(c, d, d + nt, d + 2 * nt, nt, "lpm_tree_eval"));
    int64_t mac = 0;
    LPM_CHECK(tree_eval_device(c, prec, series, theta, delta, d, d + nt, d + 2 * nt, nt, d + 3 * nt, d + 4 * nt,
                               d + 5 * nt, &mac, 1.0));
    LPM_HIP(c, hipMemcpyAsync(ux, d + 3 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(uy, d + 4 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(uz, d + 5 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    if (mac_count) *mac_count = mac;
    return LPM_OK;
}

}  // extern "C"
