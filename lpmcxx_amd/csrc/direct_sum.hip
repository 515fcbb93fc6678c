// Direct O(N*M) spherical Biot-Savart / Green's-function sums on gfx950.
//
// Replaces the pair loops of DirectSum::meshSolve (src/LpmDirectSum.cpp:25-173)
// with the kernels BiotSavartSphere (src/LpmVectorKernel.h:19-39) and
// SphereGreensFn (src/LpmScalarKernel.h:31-45).
//
// Layout and mapping
//   * Targets stay in registers: each lane owns TPL targets, a wave owns
//     64*TPL consecutive targets.
//   * Sources are read with wave-uniform indices, so every source value is a
//     scalar (SGPR) operand fetched once per wave through the scalar cache;
//     no LDS and no barriers are needed.
//   * Grid = (target groups) x (source splits).  Each (group, split) task
//     writes one FP64 partial per target; ds_reduce adds the splits in
//     ascending order in FP64 and applies the cross product and the
//     -1/(4 pi R) multiplier.  Splits exist so that small problems still fill
//     256 CUs.
//
// FP64 form (LPM_F64): the reference's per-pair arithmetic in its operation
// order -- cross product x*y, denominator (R*R - x.y) + delta*delta, zero when
// |den| <= 1e-13, reciprocal by IEEE division, then scaled by the strength.
// The library is compiled with -ffp-contract=off so each pair value is
// bit-identical to the reference's; only the summation order differs.
//
// FP64F form (LPM_F64F): the FP64 pair regrouped for the FP64 units --
// u_i = -1/(4 pi R) x_i x sum_j w_j (y_j - x_i), w_j = gamma_j / den_j with
// the reference's denominator and zero rule, a refined hardware
// reciprocal and FMA accumulation (21 VALU instructions per pair against
// about 45 for the IEEE-division form).  It agrees with LPM_F64 to rounding.
//
// FP32C form (LPM_F32C): R^2 - x.y is formed as |d|^2/2 with d = x - y,
// u = 1/(4 pi R) x * sum_j [2 gamma_j / (|d|^2 + 2 delta^2)] d_j.
// Positions enter as FP32 offsets from FP64 centers (per wave of targets and
// per 256-source tile); d = (fp32(c_t - c_s) + o_t) - o_s keeps the relative
// error of near-pair separations at the offset scale, not at |x| ~ 1.
// Plain FP32 sums run inside a tile; tile sums enter a TwoSum-compensated
// running sum; partials leave the kernel in FP64.
// Tiles whose bounding sphere comes within 1e-6 of the wave's targets ("near"
// tiles, a fraction of a percent on mesh-ordered inputs) run their pairs in
// FP64 from an FP64 copy of the tile: d = x - y is then exactly zero for a
// coincident pair (the face-target self pair), close pairs keep full accuracy,
// and the reference's zero rule |R^2 - x.y + delta^2| <= 1e-13, i.e.
// |d|^2 + 2 delta^2 <= 2e-13, applies.  All other tiles run the plain
// 11-instruction FP32 pair loop.  Inputs whose order is not spatially coherent
// would make every tile near (accurate, at FP64 speed): direct_sum_device
// therefore sums large FP32C problems on copies sorted along the sphere (cube
// face, then Hilbert index), whatever the caller's order.  The no-allocation
// entry keeps the caller's order.
#include "lpm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace lpm {
namespace {

constexpr int kTile = 256;          // sources per tile
constexpr int kWavesPerBlock = 4;   // 256 threads
constexpr double kNearPad = 1e-6;   // tiles closer than this to a wave run the tested loop

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline void two_sum_acc(float& S, float& C, float a) {
    const float s = S + a;
    const float bp = s - S;
    const float e = (S - (s - bp)) + (a - bp);
    S = s;
    C += e;
}

// ---------------------------------------------------------------- source prep
// One block per 256-source tile: FP64 mean center and FP32 offsets + strength.
__global__ __launch_bounds__(256) void ds_prep_src_f32c(const double* __restrict__ sx, const double* __restrict__ sy,
                                                        const double* __restrict__ sz, const double* __restrict__ g,
                                                        int64_t ns, float4* __restrict__ out,
                                                        double* __restrict__ centers) {
    __shared__ double red[4][kWavesPerBlock];
    const int t = threadIdx.x;
    const int64_t j = int64_t(blockIdx.x) * kTile + t;
    const bool v = j < ns;
    double x = 0.0, y = 0.0, z = 0.0;
    if (v) { x = sx[j]; y = sy[j]; z = sz[j]; }
    double a = wave_sum(x), b = wave_sum(y), c = wave_sum(z), n = wave_sum(v ? 1.0 : 0.0);
    if ((t & 63) == 0) { red[0][t >> 6] = a; red[1][t >> 6] = b; red[2][t >> 6] = c; red[3][t >> 6] = n; }
    __syncthreads();
    double cx = 0.0, cy = 0.0, cz = 0.0, cn = 0.0;
    for (int w = 0; w < kWavesPerBlock; ++w) { cx += red[0][w]; cy += red[1][w]; cz += red[2][w]; cn += red[3][w]; }
    cx /= cn; cy /= cn; cz /= cn;
    // tile radius: max distance of a source from the center
    double r = v ? sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) : 0.0;
    r = wave_max(r);
    __syncthreads();
    if ((t & 63) == 0) red[0][t >> 6] = r;
    __syncthreads();
    if (t == 0) {
        double rm = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) rm = fmax(rm, red[0][w]);
        centers[4 * blockIdx.x + 0] = cx;
        centers[4 * blockIdx.x + 1] = cy;
        centers[4 * blockIdx.x + 2] = cz;
        centers[4 * blockIdx.x + 3] = rm;
    }
    out[j] = v ? make_float4(float(x - cx), float(y - cy), float(z - cz), float(g[j])) : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void ds_prep_src_f64(const double* __restrict__ sx, const double* __restrict__ sy,
                                                       const double* __restrict__ sz, const double* __restrict__ g,
                                                       int64_t ns, int64_t nsp, double4* __restrict__ out) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nsp) return;
    out[j] = (j < ns) ? make_double4(sx[j], sy[j], sz[j], g[j]) : make_double4(0.0, 0.0, 0.0, 0.0);
}

// Sum of all strengths (potential correction term), one block.
__global__ __launch_bounds__(1024) void ds_gamma_total(const double* __restrict__ g, int64_t ns, double* out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < ns; j += 1024) s += g[j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; ++w) r += red[w];
        *out = r;
    }
}

// ---------------------------------------------------------------- FP32C pair kernel
template <int TPL, bool VEL, bool PSI>
__global__ __launch_bounds__(256) void ds_f32c(const double* __restrict__ tx, const double* __restrict__ ty,
                                               const double* __restrict__ tz, int64_t nt,
                                               const float4* __restrict__ src, const double* __restrict__ scen,
                                               const double4* __restrict__ src64, double d2x2,
                                               int ntiles, int tiles_per_split, float eps2v,
                                               const int64_t* __restrict__ skip, double* __restrict__ part,
                                               int64_t ntp) {
    const int lane = threadIdx.x & 63;
    const int64_t group = int64_t(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
    const int64_t base = group * (64 * TPL);

    double px[TPL], py[TPL], pz[TPL];
    double sxw = 0.0, syw = 0.0, szw = 0.0, cnt = 0.0;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int64_t i = base + k * 64 + lane;
        const bool v = i < nt;
        px[k] = v ? tx[i] : 0.0;
        py[k] = v ? ty[i] : 0.0;
        pz[k] = v ? tz[i] : 0.0;
        sxw += px[k]; syw += py[k]; szw += pz[k]; cnt += v ? 1.0 : 0.0;
    }
    cnt = wave_sum(cnt);
    if (cnt == 0.0) return;  // wave entirely past the end (uniform)
    const double cwx = wave_sum(sxw) / cnt, cwy = wave_sum(syw) / cnt, cwz = wave_sum(szw) / cnt;

    float ox[TPL], oy[TPL], oz[TPL];
    float Sx[TPL], Sy[TPL], Sz[TPL], Cx[TPL], Cy[TPL], Cz[TPL], Sp[TPL], Cp[TPL];
    int sk[TPL];
    double rw = 0.0;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int64_t i = base + k * 64 + lane;
        if (i < nt)
            rw = fmax(rw, sqrt((px[k] - cwx) * (px[k] - cwx) + (py[k] - cwy) * (py[k] - cwy) +
                               (pz[k] - cwz) * (pz[k] - cwz)));
    }
    rw = wave_max(rw);
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        ox[k] = float(px[k] - cwx); oy[k] = float(py[k] - cwy); oz[k] = float(pz[k] - cwz);
        Sx[k] = Sy[k] = Sz[k] = Cx[k] = Cy[k] = Cz[k] = Sp[k] = Cp[k] = 0.f;
        sk[k] = -1;
        if (PSI && skip) {
            const int64_t i = base + k * 64 + lane;
            if (i < nt) sk[k] = int(skip[i]);
        }
    }

    const int t0 = blockIdx.y * tiles_per_split;
    const int t1 = min(t0 + tiles_per_split, ntiles);
    for (int tile = t0; tile < t1; ++tile) {
        const float dcx = float(cwx - scen[4 * tile + 0]);
        const float dcy = float(cwy - scen[4 * tile + 1]);
        const float dcz = float(cwz - scen[4 * tile + 2]);
        float qx[TPL], qy[TPL], qz[TPL], ax[TPL], ay[TPL], az[TPL], ap[TPL];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
            qx[k] = dcx + ox[k]; qy[k] = dcy + oy[k]; qz[k] = dcz + oz[k];
            ax[k] = ay[k] = az[k] = ap[k] = 0.f;
        }
        const float4* s = src + int64_t(tile) * kTile;
        const int jbase = tile * kTile;
        const double ex = cwx - scen[4 * tile + 0], ey = cwy - scen[4 * tile + 1], ez = cwz - scen[4 * tile + 2];
        const bool near = sqrt(ex * ex + ey * ey + ez * ez) <= rw + scen[4 * tile + 3] + kNearPad;
        if (!near) {
#pragma unroll 8
            for (int j = 0; j < kTile; ++j) {
                const float4 q = s[j];
#pragma unroll
                for (int k = 0; k < TPL; ++k) {
                    const float dx = qx[k] - q.x;
                    const float dy = qy[k] - q.y;
                    const float dz = qz[k] - q.z;
                    if (VEL) {
                        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2v)));
                        const float w = q.w * __builtin_amdgcn_rcpf(r2);
                        ax[k] = fmaf(w, dx, ax[k]);
                        ay[k] = fmaf(w, dy, ay[k]);
                        az[k] = fmaf(w, dz, az[k]);
                    }
                    if (PSI) {
                        const float r2p = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        const float l = q.w * __builtin_amdgcn_logf(r2p);
                        ap[k] += (jbase + j == sk[k]) ? 0.f : l;
                    }
                }
            }
        } else {
            const double4* s64 = src64 + int64_t(tile) * kTile;
            for (int j = 0; j < kTile; ++j) {
                const double4 q = s64[j];
#pragma unroll
                for (int k = 0; k < TPL; ++k) {
                    const double dx = px[k] - q.x, dy = py[k] - q.y, dz = pz[k] - q.z;
                    const double dd = dx * dx + dy * dy + dz * dz;
                    if (VEL) {
                        const double r2 = dd + d2x2;
                        const double w = (r2 > 2e-13) ? q.w / r2 : 0.0;
                        ax[k] += float(w * dx);
                        ay[k] += float(w * dy);
                        az[k] += float(w * dz);
                    }
                    if (PSI) {
                        const float l = float(q.w) * __builtin_amdgcn_logf(float(dd));
                        ap[k] += (jbase + j == sk[k]) ? 0.f : l;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
            if (VEL) {
                two_sum_acc(Sx[k], Cx[k], ax[k]);
                two_sum_acc(Sy[k], Cy[k], ay[k]);
                two_sum_acc(Sz[k], Cz[k], az[k]);
            }
            if (PSI) two_sum_acc(Sp[k], Cp[k], ap[k]);
        }
    }

    double* p = part + int64_t(blockIdx.y) * 4 * ntp;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int64_t i = base + k * 64 + lane;
        if (i >= nt) continue;
        if (VEL) {
            p[i] = double(Sx[k]) + double(Cx[k]);
            p[ntp + i] = double(Sy[k]) + double(Cy[k]);
            p[2 * ntp + i] = double(Sz[k]) + double(Cz[k]);
        }
        if (PSI) p[3 * ntp + i] = double(Sp[k]) + double(Cp[k]);
    }
}

// ---------------------------------------------------------------- FP64 pair kernel
// FAST (LPM_F64F): the same pair value regrouped.  x_i x y_j = x_i x (y_j - x_i),
// so sum_j w_j (x_i x y_j) = x_i x sum_j w_j e_j with e_j = y_j - x_i
// (exactly zero for a coincident pair; e_j is exact when y_j is within a
// factor 2 of x_i, so close pairs add no cancellation);
// the cross product is applied once per target by ds_reduce.  The
// denominator is formed as in the reference (same dot product order, same
// zero rule); the weight gamma / den comes from v_rcp_f64 refined to about
// an ulp (weight_f64f); the accumulation uses FMA.
template <int TPL, bool VEL, bool PSI, bool FAST>
__global__ __launch_bounds__(256) void ds_f64(const double* __restrict__ tx, const double* __restrict__ ty,
                                              const double* __restrict__ tz, int64_t nt,
                                              const double4* __restrict__ src, int64_t nsp, int64_t per_split,
                                              double R2, double d2, const int64_t* __restrict__ skip,
                                              double* __restrict__ part, int64_t ntp) {
    const int lane = threadIdx.x & 63;
    const int64_t group = int64_t(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
    const int64_t base = group * (64 * TPL);
    if (base >= nt) return;

    double px[TPL], py[TPL], pz[TPL], ax[TPL], ay[TPL], az[TPL], ap[TPL];
    int64_t sk[TPL];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int64_t i = base + k * 64 + lane;
        const bool v = i < nt;
        px[k] = v ? tx[i] : 0.0;
        py[k] = v ? ty[i] : 0.0;
        pz[k] = v ? tz[i] : 0.0;
        ax[k] = ay[k] = az[k] = ap[k] = 0.0;
        sk[k] = (PSI && skip && v) ? skip[i] : -1;
    }
    const int64_t j0 = int64_t(blockIdx.y) * per_split;
    const int64_t j1 = min(j0 + per_split, nsp);
#pragma unroll 4
    for (int64_t j = j0; j < j1; ++j) {
        const double4 s = src[j];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
            if (VEL && FAST) {
                const double ex = s.x - px[k], ey = s.y - py[k], ez = s.z - pz[k];
                const double dot = px[k] * s.x + py[k] * s.y + pz[k] * s.z;
                const double den = R2 - dot + d2;
                const double w = fabs(den) > 1.0e-13 ? weight_f64f(s.w, den) : 0.0;
                ax[k] = __builtin_fma(w, ex, ax[k]);
                ay[k] = __builtin_fma(w, ey, ay[k]);
                az[k] = __builtin_fma(w, ez, az[k]);
            } else if (VEL) {
                const double cx = py[k] * s.z - pz[k] * s.y;
                const double cy = pz[k] * s.x - px[k] * s.z;
                const double cz = px[k] * s.y - py[k] * s.x;
                const double dot = px[k] * s.x + py[k] * s.y + pz[k] * s.z;
                const double den = R2 - dot + d2;
                const double inv = 1.0 / den;
                const bool ok = fabs(den) > 1.0e-13;
                ax[k] += ok ? (cx * inv) * s.w : 0.0;
                ay[k] += ok ? (cy * inv) * s.w : 0.0;
                az[k] += ok ? (cz * inv) * s.w : 0.0;
            }
            if (PSI) {
                const double dot = s.x * px[k] + s.y * py[k] + s.z * pz[k];
                const double l = log(R2 - dot) * s.w;
                ap[k] += (j == sk[k]) ? 0.0 : l;
            }
        }
    }
    double* p = part + int64_t(blockIdx.y) * 4 * ntp;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
        const int64_t i = base + k * 64 + lane;
        if (i >= nt) continue;
        if (VEL) { p[i] = ax[k]; p[ntp + i] = ay[k]; p[2 * ntp + i] = az[k]; }
        if (PSI) p[3 * ntp + i] = ap[k];
    }
}

// ---------------------------------------------------------------- reduction over splits
// f32c: u = vel_scale * (x * S), psi = psi_scale * (A - (Gtot - gamma[skip]))
// f64 : u = vel_scale * S,       psi = psi_scale * A
// f64f: u = vel_scale * (x * S), psi = psi_scale * A
template <bool CROSS, bool F32C>
__global__ __launch_bounds__(256) void ds_reduce(const double* __restrict__ part, int nsplit, int64_t ntp, int64_t nt,
                                                 const double* __restrict__ tx, const double* __restrict__ ty,
                                                 const double* __restrict__ tz, int vel, int psi, double vel_scale,
                                                 double psi_scale, const double* __restrict__ gtot,
                                                 const double* __restrict__ gamma, const int64_t* __restrict__ skip,
                                                 double* ux, double* uy, double* uz, double* upsi) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
    for (int s = 0; s < nsplit; ++s) {
        const double* p = part + int64_t(s) * 4 * ntp;
        if (vel) { sx += p[i]; sy += p[ntp + i]; sz += p[2 * ntp + i]; }
        if (psi) sp += p[3 * ntp + i];
    }
    if (vel) {
        if (CROSS) {
            const double x = tx[i], y = ty[i], z = tz[i];
            ux[i] = vel_scale * (y * sz - z * sy);
            uy[i] = vel_scale * (z * sx - x * sz);
            uz[i] = vel_scale * (x * sy - y * sx);
        } else {
            ux[i] = sx * vel_scale;
            uy[i] = sy * vel_scale;
            uz[i] = sz * vel_scale;
        }
    }
    if (psi) {
        if (F32C) {
            double b = *gtot;
            if (skip && skip[i] >= 0) b -= gamma[skip[i]];
            upsi[i] = psi_scale * (sp - b);
        } else {
            upsi[i] = sp * psi_scale;
        }
    }
}

template <int TPL, bool VEL, bool PSI>
void launch_f32c(dim3 g, hipStream_t st, const double* tx, const double* ty, const double* tz, int64_t nt,
                 const float4* src, const double* scen, const double4* src64, double d2x2, int ntiles, int tps,
                 float eps2v, const int64_t* skip, double* part, int64_t ntp) {
    hipLaunchKernelGGL((ds_f32c<TPL, VEL, PSI>), g, dim3(256), 0, st, tx, ty, tz, nt, src, scen, src64, d2x2, ntiles,
                       tps, eps2v, skip, part, ntp);
}

template <int TPL, bool VEL, bool PSI>
void launch_f64(dim3 g, hipStream_t st, bool fast, const double* tx, const double* ty, const double* tz, int64_t nt,
                const double4* src, int64_t nsp, int64_t per, double R2, double d2, const int64_t* skip, double* part,
                int64_t ntp) {
    if constexpr (VEL) {  // the fused form changes the velocity pair only
        if (fast) {
            hipLaunchKernelGGL((ds_f64<TPL, VEL, PSI, true>), g, dim3(256), 0, st, tx, ty, tz, nt, src, nsp, per, R2,
                               d2, skip, part, ntp);
            return;
        }
    }
    hipLaunchKernelGGL((ds_f64<TPL, VEL, PSI, false>), g, dim3(256), 0, st, tx, ty, tz, nt, src, nsp, per, R2, d2,
                       skip, part, ntp);
}

constexpr int kTplF32 = 4;
constexpr int kTplF64 = 2;

__global__ __launch_bounds__(256) void fill_zero(double* __restrict__ a, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// Launch shape and workspace of one direct sum.
struct DsPlan {
    int ntiles = 0, tps = 0, nsplit = 0;
    int64_t nsp = 0, gx = 0;
    size_t src_bytes = 0;   // prepared sources: FP64 double4 [nsp] (+ FP32C float4 [nsp] and tile centres)
    size_t part_bytes = 0;  // FP64 partials [nsplit][4][nt]
};

// max_part_bytes > 0 lowers the split count until the partials fit (at
// least one split); 0 leaves it at the count that fills the chip.
DsPlan ds_plan(int prec, int64_t nt, int64_t ns, size_t max_part_bytes) {
    DsPlan P;
    P.ntiles = int((ns + kTile - 1) / kTile);
    P.nsp = int64_t(P.ntiles) * kTile;
    if (P.ntiles == 0 || nt == 0) return P;  // no sources or targets: no launch, no workspace
    const int tpl = (prec == LPM_F32C) ? kTplF32 : kTplF64;
    const int64_t per_block = int64_t(kWavesPerBlock) * 64 * tpl;
    P.gx = (nt + per_block - 1) / per_block;
    // enough (group, split) tasks for 8 blocks per CU on 256 CUs
    int nsplit = int(std::min<int64_t>(P.ntiles, std::max<int64_t>(1, (2048 + P.gx - 1) / P.gx)));
    if (max_part_bytes > 0)
        nsplit = int(std::max<int64_t>(1, std::min<int64_t>(nsplit, int64_t(max_part_bytes / (32 * size_t(nt))))));
    P.tps = (P.ntiles + nsplit - 1) / nsplit;
    P.nsplit = (P.ntiles + P.tps - 1) / P.tps;
    P.src_bytes = prec == LPM_F32C
                      ? size_t(P.nsp) * (sizeof(float4) + sizeof(double4)) + size_t(P.ntiles) * 4 * sizeof(double)
                      : size_t(P.nsp) * sizeof(double4);
    P.part_bytes = size_t(P.nsplit) * 4 * size_t(nt) * sizeof(double);
    return P;
}

// The kernels of one direct sum on prepared workspace: src (src_bytes),
// part (part_bytes) and gtot (one double).  nt > 0.
int ds_launch(lpm_ctx* c, const DsPlan& P, int kernels, int prec, double radius, double delta, const double* tx,
              const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy, const double* sz,
              const double* gamma, int64_t ns, const int64_t* skip, double* ux, double* uy, double* uz, double* psi,
              void* srcbuf, double* part, double* gtot) {
    const bool vel = (kernels & LPM_K_BIOT_SAVART_SPHERE) != 0;
    const bool pot = (kernels & LPM_K_GREENS_SPHERE) != 0;
    hipStream_t st = c->stream;
    if (ns == 0) {  // empty source set: zero outputs
        for (double* o : {vel ? ux : nullptr, vel ? uy : nullptr, vel ? uz : nullptr, pot ? psi : nullptr})
            if (o) hipLaunchKernelGGL(fill_zero, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st, o, nt);
        LPM_HIP(c, hipGetLastError());
        return LPM_OK;
    }
    const int64_t ntp = nt;
    const double pi = 3.1415926535897932384626433832795027975;
    const dim3 g{unsigned(P.gx), unsigned(P.nsplit), 1u};
    const int ntiles = P.ntiles, tps = P.tps;
    const int64_t nsp = P.nsp;
    if (prec == LPM_F32C) {
        double4* src64 = static_cast<double4*>(srcbuf);
        float4* src = reinterpret_cast<float4*>(src64 + nsp);
        double* scen = reinterpret_cast<double*>(src + nsp);
        hipLaunchKernelGGL(ds_prep_src_f32c, dim3(ntiles), dim3(256), 0, st, sx, sy, sz, gamma, ns, src, scen);
        hipLaunchKernelGGL(ds_prep_src_f64, dim3(unsigned((nsp + 255) / 256)), dim3(256), 0, st, sx, sy, sz, gamma,
                           ns, nsp, src64);
        const double d2x2 = 2.0 * delta * delta;
        if (pot) hipLaunchKernelGGL(ds_gamma_total, dim3(1), dim3(1024), 0, st, gamma, ns, gtot);
        const float eps2v = float(2.0 * delta * delta);
        mark_begin(c);
        if (vel && pot) launch_f32c<kTplF32, true, true>(g, st, tx, ty, tz, nt, src, scen, src64, d2x2, ntiles, tps, eps2v, skip, part, ntp);
        else if (vel) launch_f32c<kTplF32, true, false>(g, st, tx, ty, tz, nt, src, scen, src64, d2x2, ntiles, tps, eps2v, skip, part, ntp);
        else launch_f32c<kTplF32, false, true>(g, st, tx, ty, tz, nt, src, scen, src64, d2x2, ntiles, tps, eps2v, skip, part, ntp);
        mark_end(c);
        LPM_HIP(c, hipGetLastError());
        hipLaunchKernelGGL((ds_reduce<true, true>), dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st, part, P.nsplit, ntp,
                           nt, tx, ty, tz, int(vel), int(pot), 2.0 / (4.0 * pi * radius),
                           -std::log(2.0) / (4.0 * pi * radius), gtot, gamma, skip, ux, uy, uz, psi);
    } else {
        double4* src = static_cast<double4*>(srcbuf);
        hipLaunchKernelGGL(ds_prep_src_f64, dim3(unsigned((nsp + 255) / 256)), dim3(256), 0, st, sx, sy, sz, gamma,
                           ns, nsp, src);
        const int64_t per = int64_t(tps) * kTile;
        const double R2 = radius * radius, d2 = delta * delta;
        const bool fast = prec == LPM_F64F;
        mark_begin(c);
        if (vel && pot)
            launch_f64<kTplF64, true, true>(g, st, fast, tx, ty, tz, nt, src, nsp, per, R2, d2, skip, part, ntp);
        else if (vel)
            launch_f64<kTplF64, true, false>(g, st, fast, tx, ty, tz, nt, src, nsp, per, R2, d2, skip, part, ntp);
        else launch_f64<kTplF64, false, true>(g, st, fast, tx, ty, tz, nt, src, nsp, per, R2, d2, skip, part, ntp);
        mark_end(c);
        LPM_HIP(c, hipGetLastError());
        auto reduce = fast ? ds_reduce<true, false> : ds_reduce<false, false>;
        hipLaunchKernelGGL(reduce, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st, part, P.nsplit, ntp, nt, tx,
                           ty, tz, int(vel), int(pot), -1.0 / (4.0 * pi * radius), -1.0 / (4.0 * pi * radius), gtot,
                           gamma, skip, ux, uy, uz, psi);
    }
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

// ---------------------------------------------------------------- spatial order (FP32C)
// The FP32C loop needs waves of targets and tiles of sources that are small
// in space; in an order that is not (shuffled particles), every tile is near
// every wave and runs the FP64 loop.  Large FP32C sums therefore run on
// copies of the targets and sources sorted by a key that follows the sphere:
// the cube face the point projects to (largest coordinate) and the 2D
// Hilbert index of its gnomonic coordinates on that face (13 bits a side);
// the outputs are scattered back.  A Hilbert curve has no jumps, so runs of
// consecutive keys stay compact on the sphere.

constexpr int kHilbertBits = 13;

__global__ __launch_bounds__(256) void ds_order_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int64_t n,
                                                     uint32_t* __restrict__ key, int* __restrict__ val) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double X = x[i], Y = y[i], Z = z[i];
    const double ax = fabs(X), ay = fabs(Y), az = fabs(Z);
    uint32_t face;
    double a, b, m;
    if (ax >= ay && ax >= az) {
        face = X >= 0.0 ? 0u : 1u;
        a = Y, b = Z, m = ax;
    } else if (ay >= az) {
        face = Y >= 0.0 ? 2u : 3u;
        a = Z, b = X, m = ay;
    } else {
        face = Z >= 0.0 ? 4u : 5u;
        a = X, b = Y, m = az;
    }
    constexpr uint32_t side = 1u << kHilbertBits;
    auto q = [&](double v) {
        const double u = m > 0.0 ? (v / m + 1.0) * (0.5 * side) : 0.0;  // [0, side] on the face
        return uint32_t(u < 0.0 ? 0.0 : (u > side - 1 ? double(side - 1) : u));
    };
    uint32_t hx = q(a), hy = q(b), d = 0;
    for (uint32_t s = side >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (hx & s) ? 1u : 0u, ry = (hy & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {  // rotate the quadrant
            if (rx == 1) {
                hx = side - 1 - hx;
                hy = side - 1 - hy;
            }
            const uint32_t t = hx;
            hx = hy;
            hy = t;
        }
    }
    key[i] = (face << (2 * kHilbertBits)) | d;
    val[i] = int(i);
}

// position of every source in the sorted order
__global__ __launch_bounds__(256) void ds_order_inverse(const int* __restrict__ perm, int64_t n, int* __restrict__ inv) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) inv[perm[j]] = int(j);
}

__global__ __launch_bounds__(256) void ds_order_gather(const int* __restrict__ perm, int64_t n,
                                                       const double* __restrict__ a, const double* __restrict__ b,
                                                       const double* __restrict__ c, const double* __restrict__ d,
                                                       double* __restrict__ as, double* __restrict__ bs,
                                                       double* __restrict__ cs, double* __restrict__ ds,
                                                       const int64_t* __restrict__ skip, const int* __restrict__ inv,
                                                       int64_t ns, int64_t* __restrict__ skip_s) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = perm[i];
    as[i] = a[k];
    bs[i] = b[k];
    cs[i] = c[k];
    if (d) ds[i] = d[k];
    if (skip) {
        const int64_t q = skip[k];
        skip_s[i] = (q >= 0 && q < ns) ? int64_t(inv[q]) : int64_t(-1);  // a source index, or none
    }
}

__global__ __launch_bounds__(256) void ds_order_scatter(const int* __restrict__ perm, int64_t n,
                                                        const double* __restrict__ u, const double* __restrict__ v,
                                                        const double* __restrict__ w, const double* __restrict__ p,
                                                        double* ux, double* uy, double* uz, double* psi) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = perm[i];
    if (ux) {
        ux[k] = u[i];
        uy[k] = v[i];
        uz[k] = w[i];
    }
    if (psi) psi[k] = p[i];
}

// sums with at least this many targets or sources take the sorted copies
// (the sort's counts are 32-bit: larger sums keep the caller's order)
constexpr int64_t kOrderMin = 65536;

constexpr int kKeyBits = 2 * kHilbertBits + 3;

int sort_order(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, uint32_t* kin,
               uint32_t* kout, int* vin, int* vout, void* tmp, size_t tmp_bytes) {
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(ds_order_keys, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, x, y, z, n, kin, vin);
    size_t b = tmp_bytes;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, b, kin, kout, vin, vout, int(n), 0, kKeyBits, st));
    return LPM_OK;
}

}  // namespace

int direct_sum_device(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                      const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                      const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux, double* uy,
                      double* uz, double* psi) {
    if (nt == 0) return LPM_OK;
    const DsPlan P = ds_plan(prec, nt, ns, 0);
    if (ns > 0) {
        LPM_CHECK(c->ds_part.reserve(P.part_bytes));
        LPM_CHECK(c->ds_misc.reserve(64));
        LPM_CHECK(c->ds_src.reserve(P.src_bytes));
    }
    if (prec != LPM_F32C || ns == 0 || std::max(nt, ns) < kOrderMin || std::max(nt, ns) > INT32_MAX)
        return ds_launch(c, P, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, skip, ux, uy, uz,
                         psi, c->ds_src.ptr, c->ds_part.as<double>(), c->ds_misc.as<double>());
    // sorted copies: keys and values of targets and sources, the source
    // positions, the copies, the outputs, the sort's temporary storage
    hipStream_t st = c->stream;
    const int64_t nm = std::max(nt, ns);
    size_t tmp_bytes = 0;
    LPM_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, static_cast<uint32_t*>(nullptr),
                                                  static_cast<uint32_t*>(nullptr), static_cast<int*>(nullptr),
                                                  static_cast<int*>(nullptr), int(nm), 0, kKeyBits, st));
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_kin = 0, o_kout = o_kin + al(4 * nm), o_vin = o_kout + al(4 * nm), o_pt = o_vin + al(4 * nm),
                 o_ps = o_pt + al(4 * nt), o_inv = o_ps + al(4 * ns), o_t = o_inv + al(4 * ns),
                 o_s = o_t + al(3 * 8 * nt), o_sk = o_s + al(4 * 8 * ns), o_out = o_sk + al(8 * nt),
                 o_tmp = o_out + al(4 * 8 * nt), total = o_tmp + al(tmp_bytes);
    LPM_CHECK(c->ds_ord.reserve(total));
    char* w = c->ds_ord.as<char>();
    uint32_t *kin = reinterpret_cast<uint32_t*>(w + o_kin), *kout = reinterpret_cast<uint32_t*>(w + o_kout);
    int *vin = reinterpret_cast<int*>(w + o_vin), *pt = reinterpret_cast<int*>(w + o_pt);
    int *ps = reinterpret_cast<int*>(w + o_ps), *inv = reinterpret_cast<int*>(w + o_inv);
    double* T = reinterpret_cast<double*>(w + o_t);
    double* S = reinterpret_cast<double*>(w + o_s);
    int64_t* sk = skip ? reinterpret_cast<int64_t*>(w + o_sk) : nullptr;
    double* O = reinterpret_cast<double*>(w + o_out);
    LPM_CHECK(sort_order(c, tx, ty, tz, nt, kin, kout, vin, pt, w + o_tmp, tmp_bytes));
    LPM_CHECK(sort_order(c, sx, sy, sz, ns, kin, kout, vin, ps, w + o_tmp, tmp_bytes));
    const unsigned bs = unsigned((ns + 255) / 256), bt = unsigned((nt + 255) / 256);
    hipLaunchKernelGGL(ds_order_inverse, dim3(bs), dim3(256), 0, st, static_cast<const int*>(ps), ns, inv);
    hipLaunchKernelGGL(ds_order_gather, dim3(bs), dim3(256), 0, st, static_cast<const int*>(ps), ns, sx, sy, sz, gamma,
                       S, S + ns, S + 2 * ns, S + 3 * ns, static_cast<const int64_t*>(nullptr),
                       static_cast<const int*>(nullptr), ns, static_cast<int64_t*>(nullptr));
    hipLaunchKernelGGL(ds_order_gather, dim3(bt), dim3(256), 0, st, static_cast<const int*>(pt), nt, tx, ty, tz,
                       static_cast<const double*>(nullptr), T, T + nt, T + 2 * nt, static_cast<double*>(nullptr), skip,
                       static_cast<const int*>(inv), ns, sk);
    const bool vel = (kernels & LPM_K_BIOT_SAVART_SPHERE) != 0;
    const bool pot = (kernels & LPM_K_GREENS_SPHERE) != 0;
    LPM_CHECK(ds_launch(c, P, kernels, prec, radius, delta, T, T + nt, T + 2 * nt, nt, S, S + ns, S + 2 * ns, S + 3 * ns,
                        ns, sk, vel ? O : nullptr, vel ? O + nt : nullptr, vel ? O + 2 * nt : nullptr,
                        pot ? O + 3 * nt : nullptr, c->ds_src.ptr, c->ds_part.as<double>(), c->ds_misc.as<double>()));
    hipLaunchKernelGGL(ds_order_scatter, dim3(bt), dim3(256), 0, st, static_cast<const int*>(pt), nt,
                       static_cast<const double*>(O), static_cast<const double*>(O + nt),
                       static_cast<const double*>(O + 2 * nt), static_cast<const double*>(O + 3 * nt),
                       vel ? ux : nullptr, vel ? uy : nullptr, vel ? uz : nullptr, pot ? psi : nullptr);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

// ---------------------------------------------------------------- static scratch
// Workspace of lpm_direct_sum_dev_noalloc: one array per GPU, made when the
// library's code object is loaded on it and held until the process ends.
// Layout: [0, 256) the total strength and the two words of the sphere check,
// then the prepared sources, then the partials (both 256-byte aligned).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device
}  // namespace

int scratch_init(lpm_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, LPM_ERR_INVALID_ARG, "device %d beyond %d", c->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;
        LPM_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the device is current (lpm_create)
        g_scratch_ptr[c->device] = p;
    }
    c->scratch = g_scratch_ptr[c->device];
    return LPM_OK;
}

int direct_sum_device_noalloc(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                              const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                              const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux,
                              double* uy, double* uz, double* psi) {
    if (!c->scratch) return fail(c, LPM_ERR_STATE, "no static scratch address for device %d", c->device);
    if (nt == 0) return LPM_OK;
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    double* gtot = reinterpret_cast<double*>(base);
    unsigned long long* r2w = reinterpret_cast<unsigned long long*>(base + 64);
    const size_t head = 256;
    const DsPlan P0 = ds_plan(prec, nt, ns, 0);
    const size_t src_bytes = (P0.src_bytes + 255) & ~size_t(255);
    const size_t need_min = head + src_bytes + 32 * size_t(nt);  // one split
    if (ns > 0 && need_min > kScratchBytes)
        return fail(c, LPM_ERR_INVALID_ARG,
                    "lpm_direct_sum_dev_noalloc: %lld targets x %lld sources need %zu bytes of workspace; the static "
                    "scratch holds %zu (use lpm_direct_sum_dev)", (long long)nt, (long long)ns, need_min,
                    size_t(kScratchBytes));
    if (prec == LPM_F32C) {
        LPM_CHECK(check_on_sphere_words(c, tx, ty, tz, nt, radius, r2w, "lpm_direct_sum_dev_noalloc (LPM_F32C) targets"));
        LPM_CHECK(check_on_sphere_words(c, sx, sy, sz, ns, radius, r2w, "lpm_direct_sum_dev_noalloc (LPM_F32C) sources"));
    }
    const DsPlan P = ds_plan(prec, nt, ns, kScratchBytes - head - src_bytes);
    return ds_launch(c, P, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, skip, ux, uy, uz, psi,
                     base + head, reinterpret_cast<double*>(base + head + src_bytes), gtot);
}

}  // namespace lpm
