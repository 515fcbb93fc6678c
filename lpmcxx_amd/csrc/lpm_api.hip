// C ABI of liblpm_gpu: contexts, status codes, host-pointer wrappers.
#include "lpm_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>

namespace lpm {

int DevBuf::reserve(size_t n) {
    if (n <= bytes) return LPM_OK;
    if (ptr) { (void)hipFree(ptr); ptr = nullptr; bytes = 0; }
    if (hipMalloc(&ptr, n) != hipSuccess) { ptr = nullptr; (void)hipGetLastError(); return LPM_ERR_ALLOC; }
    bytes = n;
    return LPM_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int fail(lpm_ctx* c, int status, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return status;
}

int hip_fail(lpm_ctx* c, hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory) return fail(c, LPM_ERR_ALLOC, "%s: %s", what, hipGetErrorString(e));
    return fail(c, LPM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

namespace {
__global__ void clear_words(unsigned long long* w) { w[threadIdx.x] = 0ull; }

// grid-stride: each thread folds its points' range first, so only the few
// hundred blocks of the grid reach the two words
__global__ __launch_bounds__(256) void r2_range_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, int64_t n,
                                                       unsigned long long* __restrict__ out) {
    double lo = 0.0, hi = 0.0;
    bool any = false;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double r2 = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
        // NaN must win both: compare on the bit patterns the words use
        const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(r2));
        if (!any || b < static_cast<unsigned long long>(__double_as_longlong(lo))) lo = r2;
        if (!any || b > static_cast<unsigned long long>(__double_as_longlong(hi))) hi = r2;
        any = true;
    }
    // the low word keeps ~bits(min), the high word bits(max): fold both ends
    r2_block_range(lo, any, out);
    __syncthreads();  // the second call reuses the block's reduction words
    r2_block_range(hi, any, out);
}
}  // namespace

int r2_range(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, unsigned long long* w,
             double* r2min, double* r2max) {
    hipLaunchKernelGGL(clear_words, dim3(1), dim3(2), 0, c->stream, w);
    if (n > 0)
        hipLaunchKernelGGL(r2_range_kernel, dim3(unsigned(std::min<int64_t>((n + 255) / 256, 512))), dim3(256), 0,
                           c->stream, x, y, z, n, w);
    LPM_HIP(c, hipGetLastError());
    unsigned long long* h = reinterpret_cast<unsigned long long*>(c->pin);
    LPM_HIP(c, hipMemcpyAsync(h, w, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    *r2min = r2_from_words(h, 0);
    *r2max = r2_from_words(h, 1);
    return LPM_OK;
}

int check_on_sphere(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double R,
                    const char* what) {
    if (n <= 0) return LPM_OK;
    LPM_CHECK(c->r2w.reserve(2 * sizeof(unsigned long long)));
    return check_on_sphere_words(c, x, y, z, n, R, c->r2w.as<unsigned long long>(), what);
}

int check_on_sphere_words(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double R,
                          unsigned long long* w, const char* what) {
    if (n <= 0) return LPM_OK;
    double lo = 0.0, hi = 0.0;
    LPM_CHECK(r2_range(c, x, y, z, n, w, &lo, &hi));
    if (!r2_on_sphere(lo, hi, R))
        return fail(c, LPM_ERR_INVALID_ARG,
                    "%s: this form needs |x| = R = %.17g (|x|^2 in [%.17g, %.17g]); use LPM_F64 with the reference "
                    "series for points off the sphere", what, R, lo, hi);
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

extern "C" {

int lpm_create(lpm_ctx** out, int device) {
    if (!out) return LPM_ERR_NULL_PTR;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { (void)hipGetLastError(); return LPM_ERR_HIP; }
    if (device < 0 || device >= n) return LPM_ERR_INVALID_ARG;
    lpm_ctx* c = new (std::nothrow) lpm_ctx();
    if (!c) return LPM_ERR_ALLOC;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_beg) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->pin), kPinWords * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        if (c->ev_beg) (void)hipEventDestroy(c->ev_beg);
        if (c->ev_end) (void)hipEventDestroy(c->ev_end);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return LPM_ERR_HIP;
    }
    if (scratch_init(c) != LPM_OK) {
        lpm_destroy(c);
        return LPM_ERR_HIP;
    }
    *out = c;
    return LPM_OK;
}

void lpm_destroy(lpm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    tree_free(c->tree);
    model_free(c->model);
    c->tree = nullptr;
    for (DevBuf* b : {&c->ds_src, &c->ds_tgt, &c->ds_part, &c->ds_misc, &c->ds_ord, &c->h_in, &c->h_out, &c->rk_work,
                      &c->tb_work, &c->tl_work, &c->mesh_work, &c->lab_work, &c->r2w})
        b->release();
    if (c->ev_beg) (void)hipEventDestroy(c->ev_beg);
    if (c->ev_end) (void)hipEventDestroy(c->ev_end);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->aux) {
        (void)hipStreamSynchronize(c->aux);
        (void)hipStreamDestroy(c->aux);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_build) (void)hipEventDestroy(c->ev_build);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lpm_set_option(lpm_ctx* c, int option, double value) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    switch (option) {
        case LPM_OPT_BATCH_CELL_TARGETS:
            if (!(value >= 1.0 && value <= 1e9)) return fail(c, LPM_ERR_INVALID_ARG, "batch cell targets out of range");
            c->batch_cell_targets = int(value);
            return LPM_OK;
        case LPM_OPT_STEREO:
            if (!(value == 0.0 || value == 1.0 || value == 2.0))
                return fail(c, LPM_ERR_INVALID_ARG, "LPM_OPT_STEREO is 0, 1 or 2");
            c->stereo = int(value);
            return LPM_OK;
        case LPM_OPT_STEREO_EXTRA_ORDER:
            if (!(value >= 0.0 && value <= 14.0 && value == std::floor(value)))
                return fail(c, LPM_ERR_INVALID_ARG, "stereographic extra order must be an integer in 0..14");
            c->stereo_extra = int(value);
            return LPM_OK;
        case LPM_OPT_STEREO_RATIO:
            if (!(value > 0.0 && value <= 4.0)) return fail(c, LPM_ERR_INVALID_ARG, "stereographic ratio out of (0, 4]");
            c->stereo_ratio = value;
            return LPM_OK;
        default: return fail(c, LPM_ERR_INVALID_ARG, "unknown option %d", option);
    }
}

int lpm_last_kernel_ms(lpm_ctx* c, float* ms) {
    if (!c || !ms) return LPM_ERR_NULL_PTR;
    if (!c->ev_valid) return fail(c, LPM_ERR_STATE, "no timed kernel yet");
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_HIP(c, hipEventSynchronize(c->ev_end));
    LPM_HIP(c, hipEventElapsedTime(ms, c->ev_beg, c->ev_end));
    return LPM_OK;
}

const char* lpm_status_string(int s) {
    switch (s) {
        case LPM_OK: return "ok";
        case LPM_ERR_INVALID_ARG: return "invalid argument";
        case LPM_ERR_NULL_PTR: return "null pointer";
        case LPM_ERR_HIP: return "HIP runtime error";
        case LPM_ERR_ALLOC: return "device allocation failed";
        case LPM_ERR_STATE: return "invalid call order";
        case LPM_ERR_TREE_DEPTH: return "octree depth limit exceeded";
        case LPM_ERR_KERNEL: return "invalid kernel";
        default: return "unknown status";
    }
}

const char* lpm_last_error(const lpm_ctx* c) { return c ? c->err.c_str() : ""; }

int lpm_synchronize(lpm_ctx* c) {
    if (!c) return LPM_ERR_NULL_PTR;
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    if (c->aux) LPM_HIP(c, hipStreamSynchronize(c->aux));
    return LPM_OK;
}

void* lpm_stream(lpm_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

static int check_direct_args(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                             const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                             const double* sz, const double* gamma, int64_t ns, double* ux, double* uy, double* uz,
                             double* psi) {
    if (!c) return LPM_ERR_NULL_PTR;
    c->err.clear();
    if (kernels <= 0 || (kernels & ~(LPM_K_BIOT_SAVART_SPHERE | LPM_K_GREENS_SPHERE)))
        return fail(c, LPM_ERR_KERNEL, "invalid kernel mask %d", kernels);
    if (prec != LPM_F64 && prec != LPM_F32C && prec != LPM_F64F)
        return fail(c, LPM_ERR_INVALID_ARG, "invalid precision %d", prec);
    if (nt < 0 || ns < 0 || ns > (int64_t(1) << 31) - 4096 || nt > (int64_t(1) << 40))
        return fail(c, LPM_ERR_INVALID_ARG, "invalid sizes nt=%lld ns=%lld", (long long)nt, (long long)ns);
    if (!(radius > 0.0) || !(delta >= 0.0) || !std::isfinite(radius) || !std::isfinite(delta))
        return fail(c, LPM_ERR_INVALID_ARG, "radius must be > 0 and delta >= 0");
    if (nt > 0 && (!tx || !ty || !tz)) return fail(c, LPM_ERR_NULL_PTR, "target coordinates are NULL");
    if (ns > 0 && (!sx || !sy || !sz || !gamma)) return fail(c, LPM_ERR_NULL_PTR, "source arrays are NULL");
    if (nt > 0 && (kernels & LPM_K_BIOT_SAVART_SPHERE) && (!ux || !uy || !uz))
        return fail(c, LPM_ERR_NULL_PTR, "velocity outputs are NULL");
    if (nt > 0 && (kernels & LPM_K_GREENS_SPHERE) && !psi) return fail(c, LPM_ERR_NULL_PTR, "potential output is NULL");
    return LPM_OK;
}

// LPM_F32C: targets and sources on |x| = radius
static int check_direct_sphere(lpm_ctx* c, int prec, double radius, const double* tx, const double* ty,
                               const double* tz, int64_t nt, const double* sx, const double* sy, const double* sz,
                               int64_t ns) {
    if (prec != LPM_F32C || nt == 0) return LPM_OK;
    LPM_CHECK(check_on_sphere(c, tx, ty, tz, nt, radius, "lpm_direct_sum (LPM_F32C) targets"));
    return check_on_sphere(c, sx, sy, sz, ns, radius, "lpm_direct_sum (LPM_F32C) sources");
}

int lpm_direct_sum_dev(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                       const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                       const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux, double* uy,
                       double* uz, double* psi) {
    LPM_CHECK(check_direct_args(c, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, ux, uy, uz, psi));
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_CHECK(check_direct_sphere(c, prec, radius, tx, ty, tz, nt, sx, sy, sz, ns));
    return direct_sum_device(c, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, skip, ux, uy, uz,
                             psi);
}

int lpm_direct_sum_dev_noalloc(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                               const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                               const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux,
                               double* uy, double* uz, double* psi) {
    LPM_CHECK(check_direct_args(c, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, ux, uy, uz, psi));
    LPM_HIP(c, hipSetDevice(c->device));
    return direct_sum_device_noalloc(c, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, skip, ux,
                                     uy, uz, psi);
}

int lpm_direct_sum(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx, const double* ty,
                   const double* tz, int64_t nt, const double* sx, const double* sy, const double* sz,
                   const double* gamma, int64_t ns, const int64_t* skip, double* ux, double* uy, double* uz,
                   double* psi) {
    LPM_CHECK(check_direct_args(c, kernels, prec, radius, delta, tx, ty, tz, nt, sx, sy, sz, gamma, ns, ux, uy, uz, psi));
    LPM_HIP(c, hipSetDevice(c->device));
    if (nt == 0) return LPM_OK;
    const bool vel = kernels & LPM_K_BIOT_SAVART_SPHERE, pot = kernels & LPM_K_GREENS_SPHERE;
    const size_t in_bytes = sizeof(double) * (3 * size_t(nt) + 4 * size_t(ns)) + (skip ? sizeof(int64_t) * nt : 0);
    LPM_CHECK(c->h_in.reserve(in_bytes));
    LPM_CHECK(c->h_out.reserve(sizeof(double) * 4 * size_t(nt)));
    double* d = c->h_in.as<double>();
    double *dtx = d, *dty = d + nt, *dtz = d + 2 * nt;
    double *dsx = d + 3 * nt, *dsy = dsx + ns, *dsz = dsy + ns, *dg = dsz + ns;
    int64_t* dskip = skip ? reinterpret_cast<int64_t*>(dg + ns) : nullptr;
    double* o = c->h_out.as<double>();
    hipStream_t st = c->stream;
    auto h2d = [&](double* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    LPM_HIP(c, h2d(dtx, tx, nt * sizeof(double)));
    LPM_HIP(c, h2d(dty, ty, nt * sizeof(double)));
    LPM_HIP(c, h2d(dtz, tz, nt * sizeof(double)));
    if (ns > 0) {
        LPM_HIP(c, h2d(dsx, sx, ns * sizeof(double)));
        LPM_HIP(c, h2d(dsy, sy, ns * sizeof(double)));
        LPM_HIP(c, h2d(dsz, sz, ns * sizeof(double)));
        LPM_HIP(c, h2d(dg, gamma, ns * sizeof(double)));
    }
    if (skip) LPM_HIP(c, hipMemcpyAsync(dskip, skip, nt * sizeof(int64_t), hipMemcpyHostToDevice, st));
    LPM_CHECK(check_direct_sphere(c, prec, radius, dtx, dty, dtz, nt, dsx, dsy, dsz, ns));
    LPM_CHECK(direct_sum_device(c, kernels, prec, radius, delta, dtx, dty, dtz, nt, dsx, dsy, dsz, dg, ns, dskip,
                                vel ? o : nullptr, vel ? o + nt : nullptr, vel ? o + 2 * nt : nullptr,
                                pot ? o + 3 * nt : nullptr));
    if (vel) {
        LPM_HIP(c, hipMemcpyAsync(ux, o, nt * sizeof(double), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipMemcpyAsync(uy, o + nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
        LPM_HIP(c, hipMemcpyAsync(uz, o + 2 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (pot) LPM_HIP(c, hipMemcpyAsync(psi, o + 3 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    LPM_HIP(c, hipStreamSynchronize(st));
    return LPM_OK;
}

}  // extern "C"
