// TaylorSeries3d / SphereGreensSeries on the GPU
// (src/LpmTaylorSeries3d.h:20-63, src/LpmTaylorSeries3d.cpp:9-154).
//
// The series keeps the multi-indices |k| <= maxP (src/LpmTaylorSeries3d.cpp:9-50),
// stored here in the lex order of MultiIndex (src/LpmMultiIndex.h:34-43), the
// iteration order of the reference's std::map.
//
//   series_coeffs   one thread per (target, expansion centre) pair: the
//                   coefficients of 1 / (1 - t.(c + h)) in h, by the
//                   reference's loop order and FP64 operations
//                   (src/LpmTaylorSeries3d.cpp:63-117): a_0 = 1/D,
//                   D = 1 - t.c, each other a_k from one predecessor times
//                   t_i (k1+k2+k3)/k_i / D.  The library is built with
//                   -ffp-contract=off, so every coefficient is the
//                   reference's bit pattern.
//   series_moments  one thread per multi-index: m_k = sum_i (y_i - c)^k s_i
//                   over the points in index order (src/LpmTaylorSeries3d.cpp:119-154),
//                   the power as pow(h_x, k1) pow(h_y, k2) pow(h_z, k3) like
//                   MultiIndex::vectorPower (src/LpmMultiIndex.h:25-27).
// Both are tiny, latency-bound workloads; they give the host API's series
// objects the same device arithmetic as the rest of the library.
#include "lpm_internal.hpp"

#include <algorithm>
#include <cmath>

namespace lpm {
namespace {

constexpr int kSeriesMaxP = 20;

// lex position of (a, b, c) among the multi-indices |k| <= P
__host__ __device__ inline int series_index(int P, int a, int b, int c) {
    int idx = 0;
    for (int i = 0; i < a; ++i) idx += (P - i + 1) * (P - i + 2) / 2;
    for (int j = 0; j < b; ++j) idx += P - a - j + 1;
    return idx + c;
}

__global__ __launch_bounds__(64) void series_coeffs(int P, int64_t n, const double* __restrict__ tx,
                                                    const double* __restrict__ ty, const double* __restrict__ tz,
                                                    const double* __restrict__ cx, const double* __restrict__ cy,
                                                    const double* __restrict__ cz, double* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nt = (P + 1) * (P + 2) * (P + 3) / 6;
    double* A = out + i * nt;
    auto at = [&](int a, int b, int c) -> double& { return A[series_index(P, a, b, c)]; };
    const double t0 = tx[i], t1 = ty[i], t2 = tz[i];
    const double denom = 1.0 - (t0 * cx[i] + t1 * cy[i] + t2 * cz[i]);
    at(0, 0, 0) = 1.0 / denom;
    for (int k1 = 1; k1 <= P; ++k1) at(k1, 0, 0) = t0 * at(k1 - 1, 0, 0) / denom;
    for (int k2 = 1; k2 <= P; ++k2) at(0, k2, 0) = t1 * at(0, k2 - 1, 0) / denom;
    for (int k1 = 1; k1 <= P; ++k1)
        for (int k2 = 1; k2 <= P - k1; ++k2)
            at(k1, k2, 0) = t1 * at(k1, k2 - 1, 0) * (double(k1 + k2) / k2) / denom;
    for (int k3 = 1; k3 <= P; ++k3) {
        at(0, 0, k3) = t2 * at(0, 0, k3 - 1) / denom;
        for (int k1 = 1; k1 <= P - k3; ++k1) at(k1, 0, k3) = t0 * at(k1 - 1, 0, k3) * (double(k1 + k3) / k1) / denom;
        for (int k2 = 1; k2 <= P - k3; ++k2) at(0, k2, k3) = t1 * at(0, k2 - 1, k3) * (double(k2 + k3) / k2) / denom;
        for (int k1 = 1; k1 <= P - k3; ++k1)
            for (int k2 = 1; k2 <= P - k3 - k1; ++k2)
                at(k1, k2, k3) = t1 * at(k1, k2 - 1, k3) * (double(k1 + k2 + k3) / k2) / denom;
    }
}

// weight_order 0: pow * (s w) (srcVals x srcWeights, src/LpmTaylorSeries3d.cpp:119-131);
// 1: (pow * s) * w (faces: values x face area, :145-154); w == nullptr: pow * s
__global__ __launch_bounds__(64) void series_moments(int P, int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ z,
                                                     const double* __restrict__ s, const double* __restrict__ w,
                                                     int weight_order, double c0, double c1, double c2,
                                                     double* __restrict__ out) {
    const int nt = (P + 1) * (P + 2) * (P + 3) / 6;
    const int t = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int a = 0, b = 0, c = 0, idx = 0;
    for (int i = 0; i <= P; ++i)
        for (int j = 0; i + j <= P; ++j)
            for (int k = 0; i + j + k <= P; ++k, ++idx)
                if (idx == t) { a = i; b = j; c = k; }
    double m = 0.0;
    for (int64_t q = 0; q < n; ++q) {
        const double h0 = x[q] - c0, h1 = y[q] - c1, h2 = z[q] - c2;
        const double vp = pow(h0, double(a)) * pow(h1, double(b)) * pow(h2, double(c));
        if (!w) m += vp * s[q];
        else if (weight_order == 0) m += vp * (s[q] * w[q]);
        else m += vp * s[q] * w[q];
    }
    out[t] = m;
}

}  // namespace
}  // namespace lpm

using namespace lpm;

extern "C" {

int lpm_series_coeffs(lpm_ctx* c, int max_p, int64_t n, const double* tx, const double* ty, const double* tz,
                      const double* cx, const double* cy, const double* cz, double* coeffs) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (max_p < 0 || max_p > kSeriesMaxP) return fail(c, LPM_ERR_INVALID_ARG, "max_p must be in [0, %d]", kSeriesMaxP);
    if (n < 0 || n > (int64_t(1) << 24)) return fail(c, LPM_ERR_INVALID_ARG, "invalid pair count");
    if (n == 0) return LPM_OK;
    if (!tx || !ty || !tz || !cx || !cy || !cz || !coeffs) return fail(c, LPM_ERR_NULL_PTR, "NULL argument");
    LPM_HIP(c, hipSetDevice(c->device));
    const int nt = (max_p + 1) * (max_p + 2) * (max_p + 3) / 6;
    const size_t in = size_t(n) * sizeof(double), outb = size_t(n) * nt * sizeof(double);
    LPM_CHECK(c->h_in.reserve(6 * in));
    LPM_CHECK(c->h_out.reserve(outb));
    double* d = c->h_in.as<double>();
    hipStream_t st = c->stream;
    const double* src[6] = {tx, ty, tz, cx, cy, cz};
    for (int k = 0; k < 6; ++k) LPM_HIP(c, hipMemcpyAsync(d + k * n, src[k], in, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(series_coeffs, dim3(unsigned((n + 63) / 64)), dim3(64), 0, st, max_p, n, d, d + n, d + 2 * n,
                       d + 3 * n, d + 4 * n, d + 5 * n, c->h_out.as<double>());
    LPM_HIP(c, hipGetLastError());
    LPM_HIP(c, hipMemcpyAsync(coeffs, c->h_out.ptr, outb, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

int lpm_series_moments(lpm_ctx* c, int max_p, int64_t n, const double* x, const double* y, const double* z,
                       const double* s, const double* w, int weight_order, double cx, double cy, double cz,
                       double* moments) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (max_p < 0 || max_p > kSeriesMaxP) return fail(c, LPM_ERR_INVALID_ARG, "max_p must be in [0, %d]", kSeriesMaxP);
    if (n < 0 || n > (int64_t(1) << 24)) return fail(c, LPM_ERR_INVALID_ARG, "invalid point count");
    if (weight_order != 0 && weight_order != 1) return fail(c, LPM_ERR_INVALID_ARG, "weight_order must be 0 or 1");
    if (!moments) return fail(c, LPM_ERR_NULL_PTR, "moments is NULL");
    if (n > 0 && (!x || !y || !z || !s)) return fail(c, LPM_ERR_NULL_PTR, "NULL point or strength array");
    LPM_HIP(c, hipSetDevice(c->device));
    const int nt = (max_p + 1) * (max_p + 2) * (max_p + 3) / 6;
    const size_t in = size_t(std::max<int64_t>(n, 1)) * sizeof(double);
    LPM_CHECK(c->h_in.reserve(5 * in));
    LPM_CHECK(c->h_out.reserve(size_t(nt) * sizeof(double)));
    double* d = c->h_in.as<double>();
    const int64_t na = std::max<int64_t>(n, 1);
    hipStream_t st = c->stream;
    if (n > 0) {
        const double* src[4] = {x, y, z, s};
        for (int k = 0; k < 4; ++k)
            LPM_HIP(c, hipMemcpyAsync(d + k * na, src[k], size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
        if (w) LPM_HIP(c, hipMemcpyAsync(d + 4 * na, w, size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(series_moments, dim3(unsigned((nt + 63) / 64)), dim3(64), 0, st, max_p, n, d, d + na,
                       d + 2 * na, d + 3 * na, w ? d + 4 * na : nullptr, weight_order, cx, cy, cz,
                       c->h_out.as<double>());
    LPM_HIP(c, hipGetLastError());
    LPM_HIP(c, hipMemcpyAsync(moments, c->h_out.ptr, size_t(nt) * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

}  // extern "C"
