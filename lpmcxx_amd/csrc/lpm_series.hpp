// Far-field series of the treecode (device code shared by the traversal
// kernels): the Cartesian expansion generated at compile time and the
// reference's lex-order recurrence.
#pragma once

#include "lpm_tree.hpp"

#include <utility>

namespace lpm {
namespace {

__constant__ double c_rinv[kMaxOrder + 1] = {0.0,       1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,
                                             1.0 / 6,   1.0 / 7,    1.0 / 8,  1.0 / 9,  1.0 / 10, 1.0 / 11,
                                             1.0 / 12,  1.0 / 13,   1.0 / 14, 1.0 / 15, 1.0 / 16};


// Lex index of the multi-index (a, b, c), |k| < P: terms with first index < a,
// then with first index a and second < b, then c.
__host__ __device__ constexpr int lex_index(int P, int a, int b, int c) {
    int idx = 0;
    for (int i = 0; i < a; ++i) idx += (P - i) * (P - i + 1) / 2;
    for (int j = 0; j < b; ++j) idx += P - a - j;
    return idx + c;
}

// degree of the graded position I (number of multi-indices with |k| <= n is (n+1)(n+2)(n+3)/6)
__host__ __device__ constexpr int shell_of(int I) {
    int n = 0;
    while ((n + 1) * (n + 2) * (n + 3) / 6 <= I) ++n;
    return n;
}

// Graded position of the multi-index (a, b, c), n = a + b + c: the shells
// |k| = 0, 1, ... in turn, inside a shell a from n down to 0, then b from
// n - a down to 0.  The terms |k| < p are the first p(p+1)(p+2)/6 positions
// for every p, so moments stored in this order serve every lower order.
__host__ __device__ constexpr int graded_index(int a, int b, int c) {
    const int n = a + b + c;
    return n * (n + 1) * (n + 2) / 6 + (n - a) * (n - a + 1) / 2 + (n - a - b);
}

// Graded visiting order of the multi-indices |k| < P with the graded
// positions of each term's k - e_i and k - 2 e_i neighbours, built at
// compile time so every coefficient lives in a named register.
template <int P>
struct CartTab {
    static constexpr int NT = P * (P + 1) * (P + 2) / 6;
    int p1[NT][3] = {};
    int p2[NT][3] = {};
    constexpr CartTab() {
        int pos = 0;
        for (int n = 0; n < P; ++n)
            for (int k1 = n; k1 >= 0; --k1)
                for (int k2 = n - k1; k2 >= 0; --k2) {
                    const int k3 = n - k1 - k2;
                    p1[pos][0] = k1 >= 1 ? graded_index(k1 - 1, k2, k3) : -1;
                    p1[pos][1] = k2 >= 1 ? graded_index(k1, k2 - 1, k3) : -1;
                    p1[pos][2] = k3 >= 1 ? graded_index(k1, k2, k3 - 1) : -1;
                    p2[pos][0] = k1 >= 2 ? graded_index(k1 - 2, k2, k3) : -1;
                    p2[pos][1] = k2 >= 2 ? graded_index(k1, k2 - 2, k3) : -1;
                    p2[pos][2] = k3 >= 2 ? graded_index(k1, k2, k3 - 2) : -1;
                    ++pos;
                }
    }
};
template <int P>
inline constexpr CartTab<P> kCartTab{};

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int P, int I, typename T>
__device__ __forceinline__ void cart_step(T* b, T e0, T e1, T e2, T iR2, const T* __restrict__ mm, T& S0, T& S1,
                                          T& S2) {
    constexpr int a0 = kCartTab<P>.p1[I][0], a1 = kCartTab<P>.p1[I][1], a2 = kCartTab<P>.p1[I][2];
    constexpr int c0 = kCartTab<P>.p2[I][0], c1 = kCartTab<P>.p2[I][1], c2 = kCartTab<P>.p2[I][2];
    T s = T(0);
    if constexpr (a0 >= 0) s = fmadd(e0, b[a0], s);
    if constexpr (a1 >= 0) s = fmadd(e1, b[a1], s);
    if constexpr (a2 >= 0) s = fmadd(e2, b[a2], s);
    if constexpr (c0 >= 0) s -= b[c0];
    if constexpr (c1 >= 0) s -= b[c1];
    if constexpr (c2 >= 0) s -= b[c2];
    b[I] = s * iR2;
    S0 = fmadd(b[I], mm[3 * I + 0], S0);
    S1 = fmadd(b[I], mm[3 * I + 1], S1);
    S2 = fmadd(b[I], mm[3 * I + 2], S2);
}

template <int P, typename T, int... I>
__device__ __forceinline__ void cart_all(std::integer_sequence<int, I...>, T* b, T e0, T e1, T e2, T iR2,
                                         const T* __restrict__ mm, T& S0, T& S1, T& S2) {
    (cart_step<P, I + 1>(b, e0, e1, e2, iR2, mm, S0, S1, S2), ...);
}

// Cartesian far field: with d = t - c and R2 = |d|^2 + 2 delta^2 the
// coefficients of 2 / (|d - h|^2 + 2 delta^2) = sum_k b_k h^k satisfy
// b_0 = 2/R2 and b_k = (2 sum_i d_i b_{k-e_i} - sum_i b_{k-2e_i}) / R2
// (from Q grad F = -F grad Q for F = 1/Q).  Computed degree shell by degree
// shell and contracted with the node moments m_k = sum_j gamma_j y_j h_j^k
// (graded order): S = sum_k b_k m_k.  Converges for maxRadius / |d| < 1.
// T = double, or float with FP32 moments.
template <int P, typename T>
__device__ __forceinline__ void far_cartesian(T d0, T d1, T d2, T delta2x2, const T* __restrict__ mm, T& S0, T& S1,
                                              T& S2) {
    constexpr int NT = CartTab<P>::NT;
    const T R2 = d0 * d0 + d1 * d1 + d2 * d2 + delta2x2;
    const T iR2 = T(1) / R2;
    T b[NT];
    b[0] = T(2) * iR2;
    S0 = b[0] * mm[0];
    S1 = b[0] * mm[1];
    S2 = b[0] * mm[2];
    cart_all<P>(std::make_integer_sequence<int, NT - 1>{}, b, T(2) * d0, T(2) * d1, T(2) * d2, iR2, mm, S0, S1, S2);
}

// FP32 far field with node-scaled coefficients and moments.  With rho the
// node's box radius, the dimensionless coefficients c_k = b_k rho^|k| R2
// satisfy c_0 = 2 and c_k = sum_i a_i c_{k-e_i} - q sum_i c_{k-2e_i} with
// a_i = 2 d_i rho / R2 and q = rho^2 / R2 (no multiply per term), and
// S = (1 / R2) sum_k c_k mh_k with the scaled moments mh_k = m_k / rho^|k|.
// Every quantity stays O(1) whatever the node size, so FP32 neither
// overflows nor underflows at high order.
template <int P, int I>
__device__ __forceinline__ void cart_step_scaled(float* c, float a0, float a1, float a2, float q,
                                                 const float* __restrict__ mm, float& S0, float& S1, float& S2) {
    constexpr int i0 = kCartTab<P>.p1[I][0], i1 = kCartTab<P>.p1[I][1], i2 = kCartTab<P>.p1[I][2];
    constexpr int j0 = kCartTab<P>.p2[I][0], j1 = kCartTab<P>.p2[I][1], j2 = kCartTab<P>.p2[I][2];
    float s = 0.f;
    if constexpr (i0 >= 0) s = fmadd(a0, c[i0], s);
    if constexpr (i1 >= 0) s = fmadd(a1, c[i1], s);
    if constexpr (i2 >= 0) s = fmadd(a2, c[i2], s);
    if constexpr (j0 >= 0 || j1 >= 0 || j2 >= 0) {
        float t = 0.f;
        if constexpr (j0 >= 0) t += c[j0];
        if constexpr (j1 >= 0) t += c[j1];
        if constexpr (j2 >= 0) t += c[j2];
        s = fmadd(-q, t, s);
    }
    c[I] = s;
    S0 = fmadd(s, mm[3 * I + 0], S0);
    S1 = fmadd(s, mm[3 * I + 1], S1);
    S2 = fmadd(s, mm[3 * I + 2], S2);
}

template <int P, int... I>
__device__ __forceinline__ void cart_all_scaled(std::integer_sequence<int, I...>, float* c, float a0, float a1,
                                                float a2, float q, const float* __restrict__ mm, float& S0,
                                                float& S1, float& S2) {
    (cart_step_scaled<P, I + 1>(c, a0, a1, a2, q, mm, S0, S1, S2), ...);
}

template <int P>
__device__ __forceinline__ void far_cartesian_scaled(float d0, float d1, float d2, float delta2x2, float rho,
                                                     const float* __restrict__ mm, float& S0, float& S1,
                                                     float& S2) {
    constexpr int NT = CartTab<P>::NT;
    const float R2 = fmadd(d0, d0, fmadd(d1, d1, fmadd(d2, d2, delta2x2)));
    const float iR2 = __builtin_amdgcn_rcpf(R2);
    const float rs = rho * iR2;
    const float q = rho * rs, rs2 = 2.f * rs;
    float c[NT];
    c[0] = 2.f;
    S0 = 2.f * mm[0];
    S1 = 2.f * mm[1];
    S2 = 2.f * mm[2];
    cart_all_scaled<P>(std::make_integer_sequence<int, NT - 1>{}, c, rs2 * d0, rs2 * d1, rs2 * d2, q, mm, S0, S1,
                       S2);
    S0 *= iR2;
    S1 *= iR2;
    S2 *= iR2;
}

// ---------------------------------------------------------------- log kernel
// Streamfunction far field: log(|d - h|^2 + eps) = sum_k a_k h^k about the
// node centre.  From grad_h log Q = (grad_h Q) / Q with Q = |d - h|^2 + eps:
// k_i a_k = -2 d_i g_{k-e_i} + 2 g_{k-2e_i} for any i with k_i > 0, where g_k
// are the coefficients of 1/Q (the b_k above are those of 2/Q, so
// a_k = (-d_i b_{k-e_i} + b_{k-2e_i}) / k_i), and a_0 = log R2.  The first
// non-zero component of k is used.  Terms |k| < P need b up to degree P - 2.
template <int P>
struct LogTab {
    static constexpr int NT = P * (P + 1) * (P + 2) / 6;
    int first[NT] = {};  // first non-zero component of k (-1 for k = 0)
    int kf[NT] = {};     // its value
    constexpr LogTab() {
        int pos = 0;
        for (int n = 0; n < P; ++n)
            for (int k1 = n; k1 >= 0; --k1)
                for (int k2 = n - k1; k2 >= 0; --k2) {
                    const int k3 = n - k1 - k2;
                    first[pos] = k1 > 0 ? 0 : (k2 > 0 ? 1 : (k3 > 0 ? 2 : -1));
                    kf[pos] = k1 > 0 ? k1 : (k2 > 0 ? k2 : k3);
                    ++pos;
                }
    }
};
template <int P>
inline constexpr LogTab<P> kLogTab{};

// coefficient recurrence only (b_I from lower degrees), FP64 form of cart_step
template <int P, int I, typename T>
__device__ __forceinline__ void coef_step(T* b, T e0, T e1, T e2, T iR2) {
    constexpr int a0 = kCartTab<P>.p1[I][0], a1 = kCartTab<P>.p1[I][1], a2 = kCartTab<P>.p1[I][2];
    constexpr int c0 = kCartTab<P>.p2[I][0], c1 = kCartTab<P>.p2[I][1], c2 = kCartTab<P>.p2[I][2];
    T s = T(0);
    if constexpr (a0 >= 0) s = fmadd(e0, b[a0], s);
    if constexpr (a1 >= 0) s = fmadd(e1, b[a1], s);
    if constexpr (a2 >= 0) s = fmadd(e2, b[a2], s);
    if constexpr (c0 >= 0) s -= b[c0];
    if constexpr (c1 >= 0) s -= b[c1];
    if constexpr (c2 >= 0) s -= b[c2];
    b[I] = s * iR2;
}

// term I of the log series: a_I mu_I added to S (FP64, unscaled moments)
template <int P, int I>
__device__ __forceinline__ void log_term(double* b, double e0, double e1, double e2, double iR2, double d0, double d1,
                                         double d2, const double* __restrict__ mu, double& S) {
    constexpr int NTb = CartTab<P - 1>::NT;  // coefficients of degree <= P - 2
    if constexpr (I < NTb) coef_step<P, I>(b, e0, e1, e2, iR2);
    constexpr int i = kLogTab<P>.first[I];
    constexpr int j1 = kCartTab<P>.p1[I][i], j2 = kCartTab<P>.p2[I][i];
    constexpr double inv = 1.0 / double(kLogTab<P>.kf[I]);
    const double di = i == 0 ? d0 : (i == 1 ? d1 : d2);
    double a = -di * b[j1];
    if constexpr (j2 >= 0) a += b[j2];
    S = fma(a * inv, mu[I], S);
}

template <int P, int... I>
__device__ __forceinline__ void log_all(std::integer_sequence<int, I...>, double* b, double e0, double e1, double e2,
                                        double iR2, double d0, double d1, double d2, const double* __restrict__ mu,
                                        double& S) {
    (log_term<P, I + 1>(b, e0, e1, e2, iR2, d0, d1, d2, mu, S), ...);
}

// sum_{|k| < P} a_k mu_k for the target offset d = t - c (FP64)
template <int P>
__device__ __forceinline__ double far_log(double d0, double d1, double d2, double eps, const double* __restrict__ mu) {
    constexpr int NT = CartTab<P>::NT;
    constexpr int NTb = CartTab<P - 1>::NT > 0 ? CartTab<P - 1>::NT : 1;
    const double R2 = d0 * d0 + d1 * d1 + d2 * d2 + eps;
    const double iR2 = 1.0 / R2;
    double b[NTb];
    b[0] = 2.0 * iR2;
    double S = log(R2) * mu[0];
    if constexpr (NT > 1)
        log_all<P>(std::make_integer_sequence<int, NT - 1>{}, b, 2.0 * d0, 2.0 * d1, 2.0 * d2, iR2, d0, d1, d2, mu, S);
    return S;
}

// FP32 with node-scaled coefficients c_k = b_k rho^|k| R2 and moments
// mu_k / rho^|k|: a_k rho^|k| = (-h_i c_{k-e_i} + q c_{k-2e_i}) / k_i with
// h_i = d_i rho / R2 and q = rho^2 / R2.
template <int P, int I>
__device__ __forceinline__ void log_term_scaled(float* c, float a0, float a1, float a2, float q, float h0, float h1,
                                                float h2, const float* __restrict__ mu, float& S) {
    constexpr int NTb = CartTab<P - 1>::NT;
    if constexpr (I < NTb) {
        constexpr int i0 = kCartTab<P>.p1[I][0], i1 = kCartTab<P>.p1[I][1], i2 = kCartTab<P>.p1[I][2];
        constexpr int k0 = kCartTab<P>.p2[I][0], k1 = kCartTab<P>.p2[I][1], k2 = kCartTab<P>.p2[I][2];
        float s = 0.f;
        if constexpr (i0 >= 0) s = fmadd(a0, c[i0], s);
        if constexpr (i1 >= 0) s = fmadd(a1, c[i1], s);
        if constexpr (i2 >= 0) s = fmadd(a2, c[i2], s);
        if constexpr (k0 >= 0 || k1 >= 0 || k2 >= 0) {
            float t = 0.f;
            if constexpr (k0 >= 0) t += c[k0];
            if constexpr (k1 >= 0) t += c[k1];
            if constexpr (k2 >= 0) t += c[k2];
            s = fmadd(-q, t, s);
        }
        c[I] = s;
    }
    constexpr int i = kLogTab<P>.first[I];
    constexpr int j1 = kCartTab<P>.p1[I][i], j2 = kCartTab<P>.p2[I][i];
    constexpr float inv = 1.f / float(kLogTab<P>.kf[I]);
    const float hi = i == 0 ? h0 : (i == 1 ? h1 : h2);
    float a = -hi * c[j1];
    if constexpr (j2 >= 0) a = fmadd(q, c[j2], a);
    S = fmadd(a * inv, mu[I], S);
}

template <int P, int... I>
__device__ __forceinline__ void log_all_scaled(std::integer_sequence<int, I...>, float* c, float a0, float a1,
                                               float a2, float q, float h0, float h1, float h2,
                                               const float* __restrict__ mu, float& S) {
    (log_term_scaled<P, I + 1>(c, a0, a1, a2, q, h0, h1, h2, mu, S), ...);
}

template <int P>
__device__ __forceinline__ float far_log_scaled(float d0, float d1, float d2, float eps, float rho,
                                                const float* __restrict__ mu) {
    constexpr int NT = CartTab<P>::NT;
    constexpr int NTb = CartTab<P - 1>::NT > 0 ? CartTab<P - 1>::NT : 1;
    const float R2 = fmadd(d0, d0, fmadd(d1, d1, fmadd(d2, d2, eps)));
    const float iR2 = __builtin_amdgcn_rcpf(R2);
    const float rs = rho * iR2, q = rho * rs;
    float c[NTb];
    c[0] = 2.f;
    float S = 0.69314718055994531f * __builtin_amdgcn_logf(R2) * mu[0];
    if constexpr (NT > 1)
        log_all_scaled<P>(std::make_integer_sequence<int, NT - 1>{}, c, 2.f * rs * d0, 2.f * rs * d1, 2.f * rs * d2,
                          q, rs * d0, rs * d1, rs * d2, mu, S);
    return S;
}

// Same series in FP64, calling on_shell(n, S0, S1, S2) after every degree
// shell n: S is then the sum over |k| <= n, the order-(n + 1) far field.
template <int P, int I, class F>
__device__ __forceinline__ void cart_step_shell(double* b, double e0, double e1, double e2, double iR2,
                                                const double* __restrict__ mm, double& S0, double& S1, double& S2,
                                                F& on_shell) {
    cart_step<P, I>(b, e0, e1, e2, iR2, mm, S0, S1, S2);
    constexpr int n = shell_of(I);
    if constexpr (I + 1 == (n + 1) * (n + 2) * (n + 3) / 6) on_shell(n, S0, S1, S2);
}

template <int P, class F, int... I>
__device__ __forceinline__ void cart_all_shells(std::integer_sequence<int, I...>, double* b, double e0, double e1,
                                                double e2, double iR2, const double* __restrict__ mm, double& S0,
                                                double& S1, double& S2, F& on_shell) {
    (cart_step_shell<P, I + 1>(b, e0, e1, e2, iR2, mm, S0, S1, S2, on_shell), ...);
}

template <int P, class F>
__device__ __forceinline__ void far_cartesian_shells(double d0, double d1, double d2, double delta2x2,
                                                     const double* __restrict__ mm, F& on_shell) {
    constexpr int NT = CartTab<P>::NT;
    const double R2 = d0 * d0 + d1 * d1 + d2 * d2 + delta2x2;
    const double iR2 = 1.0 / R2;
    double b[NT];
    b[0] = 2.0 * iR2;
    double S0 = b[0] * mm[0], S1 = b[0] * mm[1], S2 = b[0] * mm[2];
    on_shell(0, S0, S1, S2);
    cart_all_shells<P>(std::make_integer_sequence<int, NT - 1>{}, b, 2.0 * d0, 2.0 * d1, 2.0 * d2, iR2, mm, S0, S1,
                       S2, on_shell);
}

// Reference far field (SumNode::calc_coeffs, src/LpmTreeSum.cpp:149-171):
// coefficients of 1/(1 + delta^2 - t.c - t.h) streamed in lex order.
__device__ __forceinline__ void far_reference(int p, double t0, double t1, double t2, double cx, double cy,
                                              double cz, double d2, const double* __restrict__ mm, double& S0,
                                              double& S1, double& S2) {
    const double D = 1.0 + d2 - (t0 * cx + t1 * cy + t2 * cz);
    const double invD = 1.0 / D;
    S0 = S1 = S2 = 0.0;
    double A1 = invD;
    int idx = 0;
    for (int k1 = 0; k1 < p; ++k1) {
        if (k1 > 0) A1 = ((double(k1) * invD) * t0) * A1 * c_rinv[k1];
        double A2 = A1;
        for (int k2 = 0; k2 < p - k1; ++k2) {
            if (k2 > 0) A2 = ((double(k1 + k2) * invD) * t1) * A2 * c_rinv[k2];
            double A3 = A2;
            for (int k3 = 0; k3 < p - k1 - k2; ++k3) {
                if (k3 > 0) A3 = ((double(k1 + k2 + k3) * invD) * t2) * A3 * c_rinv[k3];
                S0 = fma(A3, mm[3 * idx + 0], S0);
                S1 = fma(A3, mm[3 * idx + 1], S1);
                S2 = fma(A3, mm[3 * idx + 2], S2);
                ++idx;
            }
        }
    }
}

}  // namespace
}  // namespace lpm
