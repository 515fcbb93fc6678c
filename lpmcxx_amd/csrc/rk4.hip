// Lagrangian RK4 step of the barotropic vorticity equation on the rotating
// unit sphere (the reference C++ has no time stepping; its Fortran solvers
// advance the same Lagrangian particles).  Every particle carries a conserved
// absolute vorticity zeta_abs and a conserved area A; at each stage
//     gamma_j = (zeta_abs_j - 2 Omega z_j) A_j           (relative circulation)
//     k       = u(x_stage; gamma)                        (direct sum or treecode)
// and
//     x1 = |x0 + dt/2 k1|^-1 (x0 + dt/2 k1), x2 = ... (dt/2 k2), x3 = ... (dt k3)
//     x_new = normalize(x0 + dt/6 (k1 + 2 k2 + 2 k3 + k4)).
// Stage and combine kernels are FP64 and HBM-bound (about 100 B per particle
// per stage).
#include "lpm_internal.hpp"

#include <cmath>

namespace lpm {

namespace {

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

__global__ __launch_bounds__(256) void rk4_gamma(int64_t n, const double* __restrict__ z,
                                                 const double* __restrict__ zeta_abs, const double* __restrict__ area,
                                                 double two_omega, double* __restrict__ gamma) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) gamma[i] = (zeta_abs[i] - two_omega * z[i]) * area[i];
}

// xs = normalize(x0 + h k); acc (+)= w k
__global__ __launch_bounds__(256) void rk4_stage(int64_t n, const double* __restrict__ x0, const double* __restrict__ y0,
                                                 const double* __restrict__ z0, const double* __restrict__ kx,
                                                 const double* __restrict__ ky, const double* __restrict__ kz, double h,
                                                 double w, int first, double* __restrict__ ax, double* __restrict__ ay,
                                                 double* __restrict__ az, double* __restrict__ xs,
                                                 double* __restrict__ ys, double* __restrict__ zs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vx = kx[i], vy = ky[i], vz = kz[i];
    if (first) { ax[i] = w * vx; ay[i] = w * vy; az[i] = w * vz; }
    else { ax[i] += w * vx; ay[i] += w * vy; az[i] += w * vz; }
    if (xs) {
        const double px = x0[i] + h * vx, py = y0[i] + h * vy, pz = z0[i] + h * vz;
        const double r = sqrt(px * px + py * py + pz * pz);
        xs[i] = px / r; ys[i] = py / r; zs[i] = pz / r;
    }
}

__global__ __launch_bounds__(256) void rk4_final(int64_t n, double* __restrict__ x, double* __restrict__ y,
                                                 double* __restrict__ z, const double* __restrict__ ax,
                                                 const double* __restrict__ ay, const double* __restrict__ az,
                                                 double h) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i] + h * ax[i], py = y[i] + h * ay[i], pz = z[i] + h * az[i];
    const double r = sqrt(px * px + py * py + pz * pz);
    x[i] = px / r; y[i] = py / r; z[i] = pz / r;
}

}  // namespace

// velocity at all n particles (x, y, z) induced by the first ns of them with
// circulation gamma[ns], into (ux, uy, uz)
static int velocity(lpm_ctx* c, const lpm_solver_config& cfg, int64_t n, int64_t ns, const double* x,
                    const double* y, const double* z, const double* gamma, double* ux, double* uy, double* uz) {
    if (cfg.method == LPM_METHOD_DIRECT)
        return direct_sum_device(c, LPM_K_BIOT_SAVART_SPHERE, cfg.prec, 1.0, cfg.delta, x, y, z, n, x, y, z, gamma, ns,
                                 nullptr, ux, uy, uz, nullptr);
    LPM_CHECK(tree_build_device(c, x, y, z, ns, cfg.root_box_min, cfg.root_box_max, LPM_TREE_MAX_PER_NODE,
                                cfg.leaf_max));
    LPM_CHECK(tree_moments_device(c, gamma, cfg.order, 1));
    if (cfg.method == LPM_METHOD_TREE_BATCH)
        return tree_eval_batch_device(c, cfg.prec, cfg.order, cfg.theta, cfg.delta, x, y, z, n, ux, uy, uz);
    if (cfg.method == LPM_METHOD_TREE_LEARNED)
        return tree_eval_learned_device(c, cfg.prec, cfg.theta, cfg.margin, cfg.order, cfg.u_ref, cfg.delta, x, y, z,
                                        n, ux, uy, uz);
    return tree_eval_device(c, cfg.prec, cfg.series, cfg.theta, cfg.delta, x, y, z, n, ux, uy, uz, nullptr, 1.0);
}

// One RK4 step of n = ns + nt particles: the first ns carry circulation
// (zeta_abs[ns], area[ns]), the last nt are passive tracers moved by the same
// velocity field.
int rk4_step_device(lpm_ctx* c, const lpm_solver_config& cfg, double dt, double omega, int64_t ns, int64_t nt,
                    double* x, double* y, double* z, const double* zeta_abs, const double* area) {
    const int64_t n = ns + nt;
    if (n == 0) return LPM_OK;
    hipStream_t st = c->stream;
    LPM_CHECK(c->rk_work.reserve(size_t(n) * 13 * sizeof(double)));
    double* w = c->rk_work.as<double>();
    double *sx = w, *sy = w + n, *sz = w + 2 * n;          // stage positions
    double *kx = w + 3 * n, *ky = w + 4 * n, *kz = w + 5 * n;  // stage velocity
    double *ax = w + 6 * n, *ay = w + 7 * n, *az = w + 8 * n;  // weighted sum
    double* g = w + 9 * n;                                  // circulation
    const double h[4] = {0.5 * dt, 0.5 * dt, dt, 0.0};
    const double wt[4] = {1.0, 2.0, 2.0, 1.0};
    const double* px = x;
    const double* py = y;
    const double* pz = z;
    for (int s = 0; s < 4; ++s) {
        if (ns > 0)
            hipLaunchKernelGGL(rk4_gamma, dim3(blocks(ns)), dim3(256), 0, st, ns, pz, zeta_abs, area, 2.0 * omega, g);
        LPM_CHECK(velocity(c, cfg, n, ns, px, py, pz, g, kx, ky, kz));
        hipLaunchKernelGGL(rk4_stage, dim3(blocks(n)), dim3(256), 0, st, n, x, y, z, kx, ky, kz, h[s], wt[s],
                           int(s == 0), ax, ay, az, s < 3 ? sx : nullptr, sy, sz);
        px = sx; py = sy; pz = sz;
    }
    hipLaunchKernelGGL(rk4_final, dim3(blocks(n)), dim3(256), 0, st, n, x, y, z, ax, ay, az, dt / 6.0);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

extern "C" {

static int check_rk4(lpm_ctx* c, const lpm_solver_config* cfg, double dt, int64_t ns, int64_t nt, const double* x,
                     const double* y, const double* z, const double* zeta_abs, const double* area) {
    if (!c || !cfg) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (ns < 0 || nt < 0 || ns + nt > (int64_t(1) << 30)) return fail(c, LPM_ERR_INVALID_ARG, "invalid particle count");
    if (ns + nt > 0 && (!x || !y || !z)) return fail(c, LPM_ERR_NULL_PTR, "NULL particle array");
    if (ns > 0 && (!zeta_abs || !area)) return fail(c, LPM_ERR_NULL_PTR, "NULL vorticity or area array");
    if (!std::isfinite(dt)) return fail(c, LPM_ERR_INVALID_ARG, "invalid dt");
    if (cfg->method < LPM_METHOD_DIRECT || cfg->method > LPM_METHOD_TREE_LEARNED)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid velocity method");
    if (cfg->prec != LPM_F64 && cfg->prec != LPM_F32C && cfg->prec != LPM_F64F)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid precision");
    if (cfg->prec == LPM_F64F && cfg->method != LPM_METHOD_DIRECT && cfg->method != LPM_METHOD_TREE)
        return fail(c, LPM_ERR_INVALID_ARG, "LPM_F64F serves the direct sum and the per-target treecode only");
    if (cfg->method == LPM_METHOD_TREE_LEARNED) {
        if (!model_weights(c)) return fail(c, LPM_ERR_STATE, "no order model (lpm_model_set_params)");
        if (!(cfg->u_ref > 0.0) || cfg->margin < 0)
            return fail(c, LPM_ERR_INVALID_ARG, "learned treecode needs u_ref > 0 and margin >= 0");
    }
    if (cfg->method != LPM_METHOD_DIRECT) {
        if (cfg->order < 1 || cfg->order > 12 || cfg->leaf_max < 1)
            return fail(c, LPM_ERR_INVALID_ARG, "tree order must be in [1, 12] and leaf_max >= 1");
        if (cfg->series != LPM_SERIES_CARTESIAN && cfg->series != LPM_SERIES_REFERENCE)
            return fail(c, LPM_ERR_INVALID_ARG, "invalid series");
        for (int d = 0; d < 3; ++d)
            if (!(cfg->root_box_min[d] < cfg->root_box_max[d]))
                return fail(c, LPM_ERR_INVALID_ARG, "tree root box must have min < max");
    }
    return LPM_OK;
}

// the solver forms that need every particle on the unit sphere (every form
// but the LPM_F64 direct sum and the LPM_F64 reference-series treecode)
static bool needs_sphere(const lpm_solver_config* cfg) {
    if (cfg->prec == LPM_F32C) return true;
    if (cfg->method == LPM_METHOD_DIRECT) return false;
    return cfg->method != LPM_METHOD_TREE || cfg->series == LPM_SERIES_CARTESIAN;
}

static int steps_host(lpm_ctx* c, const lpm_solver_config* cfg, double dt, double omega, int nsteps, int64_t ns,
                      int64_t nt, double* x, double* y, double* z, const double* zeta_abs, const double* area) {
    if (nsteps < 0) return fail(c, LPM_ERR_INVALID_ARG, "nsteps must be >= 0");
    LPM_HIP(c, hipSetDevice(c->device));
    const int64_t n = ns + nt;
    if (n == 0) return LPM_OK;
    LPM_CHECK(c->h_in.reserve(size_t(3 * n + 2 * std::max<int64_t>(ns, 1)) * sizeof(double)));
    double* d = c->h_in.as<double>();
    double *dz = d + 3 * n, *da = dz + std::max<int64_t>(ns, 1);
    hipStream_t st = c->stream;
    const size_t b = size_t(n) * sizeof(double), bs = size_t(ns) * sizeof(double);
    LPM_HIP(c, hipMemcpyAsync(d, x, b, hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + n, y, b, hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemcpyAsync(d + 2 * n, z, b, hipMemcpyHostToDevice, st));
    if (ns > 0) {
        LPM_HIP(c, hipMemcpyAsync(dz, zeta_abs, bs, hipMemcpyHostToDevice, st));
        LPM_HIP(c, hipMemcpyAsync(da, area, bs, hipMemcpyHostToDevice, st));
    }
    if (needs_sphere(cfg)) LPM_CHECK(check_on_sphere(c, d, d + n, d + 2 * n, n, 1.0, "RK4 particles"));
    for (int s = 0; s < nsteps; ++s) LPM_CHECK(rk4_step_device(c, *cfg, dt, omega, ns, nt, d, d + n, d + 2 * n, dz, da));
    LPM_HIP(c, hipMemcpyAsync(x, d, b, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(y, d + n, b, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipMemcpyAsync(z, d + 2 * n, b, hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

void lpm_solver_config_default(lpm_solver_config* cfg) {
    if (!cfg) return;
    cfg->method = LPM_METHOD_DIRECT;
    cfg->prec = LPM_F64;
    cfg->series = LPM_SERIES_CARTESIAN;
    cfg->order = 8;
    cfg->leaf_max = 128;
    cfg->theta = 0.5;
    cfg->delta = 0.0;
    cfg->margin = 0;
    cfg->u_ref = 0.0;
    for (int d = 0; d < 3; ++d) {
        cfg->root_box_min[d] = -1.0 - 1e-5 - 1e-3;
        cfg->root_box_max[d] = 1.0 + 1e-5 + 1e-3;
    }
}

int lpm_rk4_step_dev(lpm_ctx* c, const lpm_solver_config* cfg, double dt, double omega, int64_t n, double* x,
                     double* y, double* z, const double* zeta_abs, const double* area) {
    LPM_CHECK(check_rk4(c, cfg, dt, n, 0, x, y, z, zeta_abs, area));
    LPM_HIP(c, hipSetDevice(c->device));
    if (needs_sphere(cfg)) LPM_CHECK(check_on_sphere(c, x, y, z, n, 1.0, "lpm_rk4_step_dev particles"));
    return rk4_step_device(c, *cfg, dt, omega, n, 0, x, y, z, zeta_abs, area);
}

int lpm_rk4_steps(lpm_ctx* c, const lpm_solver_config* cfg, double dt, double omega, int nsteps, int64_t n,
                  double* x, double* y, double* z, const double* zeta_abs, const double* area) {
    LPM_CHECK(check_rk4(c, cfg, dt, n, 0, x, y, z, zeta_abs, area));
    return steps_host(c, cfg, dt, omega, nsteps, n, 0, x, y, z, zeta_abs, area);
}

int lpm_bve_step_dev(lpm_ctx* c, const lpm_solver_config* cfg, double dt, double omega, int64_t n_src,
                     int64_t n_tracer, double* x, double* y, double* z, const double* zeta_abs, const double* area) {
    LPM_CHECK(check_rk4(c, cfg, dt, n_src, n_tracer, x, y, z, zeta_abs, area));
    LPM_HIP(c, hipSetDevice(c->device));
    if (needs_sphere(cfg))
        LPM_CHECK(check_on_sphere(c, x, y, z, n_src + n_tracer, 1.0, "lpm_bve_step_dev particles"));
    return rk4_step_device(c, *cfg, dt, omega, n_src, n_tracer, x, y, z, zeta_abs, area);
}

int lpm_bve_steps(lpm_ctx* c, const lpm_solver_config* cfg, double dt, double omega, int nsteps, int64_t n_src,
                  int64_t n_tracer, double* x, double* y, double* z, const double* zeta_abs, const double* area) {
    LPM_CHECK(check_rk4(c, cfg, dt, n_src, n_tracer, x, y, z, zeta_abs, area));
    return steps_host(c, cfg, dt, omega, nsteps, n_src, n_tracer, x, y, z, zeta_abs, area);
}

}  // extern "C"
