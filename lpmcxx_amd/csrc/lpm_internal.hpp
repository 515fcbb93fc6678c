// Internal declarations shared by the HIP translation units of liblpm_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lpm_gpu.h"

namespace lpm {

constexpr int kWave = 64;

// Growable device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t n);   // grows (never shrinks); returns lpm_status
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct Tree;   // octree state (octree.hip)
struct Model;  // order model state (mlp.hip)

// counters of the last tree evaluation: far (target, node) pairs, near
// (target, source) pairs, far-field series terms, order-model calls, the far
// pairs and terms of the order 8..12 list kernel, the stereographic far pairs
// and their complex terms
constexpr int kStats = 8;
// The counters are kept in kStatSlots copies, one 128-byte line each (block b
// adds to copy b % kStatSlots); readers sum the copies.  Tens of thousands of
// waves flush their counts at the end of a launch, which on one line queue
// behind each other.
constexpr int kStatSlots = 32;
constexpr int kStatStride = 16;  // 64-bit words per copy
constexpr size_t kStatBytes = size_t(kStatSlots) * kStatStride * sizeof(unsigned long long);

// Words that follow the evaluation counters in device memory and are cleared
// together with them (one fill per evaluation): the target batching's
// histogram of adjacent key prefixes and its batch count, the interaction
// lists' overflow flag and needed capacities, and the batches per cost bin of
// the learned evaluation.
struct BatchHeader {
    unsigned hist[32];
    int nsel;
    int overflow[4];
    unsigned bins[32];
};
constexpr size_t kHeaderBytes = 512;
static_assert(sizeof(BatchHeader) <= kHeaderBytes, "batch header");
inline BatchHeader* batch_header(unsigned long long* stats) {
    return reinterpret_cast<BatchHeader*>(reinterpret_cast<char*>(stats) + kStatBytes);
}

}  // namespace lpm

struct lpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // direct-sum workspace
    lpm::DevBuf ds_src, ds_tgt, ds_part, ds_misc;
    // sorted copies of the FP32C direct sum's targets and sources
    lpm::DevBuf ds_ord;
    // staging for the host-pointer entry points
    lpm::DevBuf h_in, h_out;
    // RK4 stage storage
    lpm::DevBuf rk_work;
    // target batches of the batch treecode, interaction lists of the learned one
    lpm::DevBuf tb_work, tl_work;
    // level tables of the GPU mesh generator
    lpm::DevBuf mesh_work;
    // ordering of the label rows (lpm_tree_labels)
    lpm::DevBuf lab_work;
    // two words of the sphere check (r2_range)
    lpm::DevBuf r2w;
    // this device's address of the static scratch of lpm_direct_sum_dev_noalloc
    void* scratch = nullptr;
    int far_cap = 1024, near_cap = 512, st_cap = 1024;
    int batch_cell_targets = 256;  // LPM_OPT_BATCH_CELL_TARGETS
    int stereo = 0;                // LPM_OPT_STEREO
    int stereo_extra = 1;          // LPM_OPT_STEREO_EXTRA_ORDER
    double stereo_ratio = 1.0;     // LPM_OPT_STEREO_RATIO
    lpm::Tree* tree = nullptr;
    lpm::Model* model = nullptr;
    unsigned long long* stats_ptr = nullptr;  // kStats counters of the last tree evaluation (device)
    // events around the last pair / evaluation kernel (lpm_last_kernel_ms)
    hipEvent_t ev_beg = nullptr, ev_end = nullptr;
    bool ev_valid = false;
    // pinned host words for small device-to-host reads (build table, flags)
    int* pin = nullptr;
    // second stream (default priority) and its fork / join events: the
    // stereographic moments run beside the Cartesian ones (made on first use)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the second stream's last work (ev_join) is not yet ordered before the
    // first stream's: the learned evaluation joins it just before it reads
    // the stereographic node data, anything that rewrites the tree at its start
    bool aux_pending = false;
    // recorded after the build table's copy to the host (the build waits for it alone)
    hipEvent_t ev_build = nullptr;
};

namespace lpm {
constexpr int kPinWords = 512;
}

namespace lpm {

int fail(lpm_ctx* c, int status, const char* fmt, ...);

// the kStats evaluation counters summed over their copies (waits for the stream)
int read_stats(lpm_ctx* c, unsigned long long* out);

// bracket the main kernel of an entry point with the context's timing events
inline void mark_begin(lpm_ctx* c) {
    if (c->ev_beg) (void)hipEventRecord(c->ev_beg, c->stream);
}
inline void mark_end(lpm_ctx* c) {
    if (c->ev_end) {
        (void)hipEventRecord(c->ev_end, c->stream);
        c->ev_valid = true;
    }
}
int hip_fail(lpm_ctx* c, hipError_t e, const char* what);

// Orders the second stream's pending work before the context stream's next work.
inline hipError_t join_aux(lpm_ctx* c) {
    if (!c->aux_pending) return hipSuccess;
    c->aux_pending = false;
    return hipStreamWaitEvent(c->stream, c->ev_join, 0);
}

#define LPM_HIP(ctx, call)                                  \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return ::lpm::hip_fail(ctx, _e, #call); \
    } while (0)

#define LPM_CHECK(x)                 \
    do {                             \
        int _s = (x);                \
        if (_s != LPM_OK) return _s; \
    } while (0)

// Static device workspace of lpm_direct_sum_dev_noalloc (direct_sum.hip):
// one array of kScratchBytes per GPU, its address looked up once per device
// by scratch_init (called from lpm_create with the device current).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
int scratch_init(lpm_ctx* c);
// the direct sum on the static scratch: no allocation (LPM_ERR_INVALID_ARG
// when the workspace of nt x ns does not fit); LPM_F32C waits once for the
// sphere check
int direct_sum_device_noalloc(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                              const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                              const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux,
                              double* uy, double* uz, double* psi);

// direct sum (direct_sum.hip); device pointers, enqueued on ctx->stream
int direct_sum_device(lpm_ctx* c, int kernels, int prec, double radius, double delta, const double* tx,
                      const double* ty, const double* tz, int64_t nt, const double* sx, const double* sy,
                      const double* sz, const double* gamma, int64_t ns, const int64_t* skip, double* ux,
                      double* uy, double* uz, double* psi);

void tree_free(Tree* t);
int tree_build_device(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
                      const double* box_min, const double* box_max, int control, int param);
int tree_moments_device(lpm_ctx* c, const double* gamma, int p, int sign);
// streamfunction moments mu_k = sum_j gamma_j (y_j - c)^k of every node (one component)
int tree_moments_psi_device(lpm_ctx* c, const double* gamma, int p);
// stereographic node centres, radii and moments of both projections for
// Cartesian orders up to p (stereo.hip), launched on stream st
int tree_stereo_moments(lpm_ctx* c, int p, hipStream_t st);
int tree_eval_device(lpm_ctx* c, int prec, int series, double theta, double delta, const double* tx,
                     const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz,
                     int64_t* mac_out, double radius);
// Morton order of the targets in the tree's root box (tperm: sorted position
// -> target index, device; *keys, nullable: the sorted keys); also clears the
// evaluation counters and the batch header.  *levels (nullable) receives the
// number of top key levels by which the sorted keys are ordered: below them,
// neighbouring keys keep the input order.  Targets that are the tree's
// sources use the build's sorted keys and order in place (read only).
int sort_targets_for_tree(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                          const int** tperm, int* levels = nullptr, const uint64_t** keys = nullptr);
// Cell-aligned target batches: the Morton-sorted targets split at the
// boundaries of the octree cells of level L (the shallowest level with at
// least nt / batch_cell_targets occupied cells, and at most the number of
// key levels the sort ordered) and cut into runs of at most 64.  Enqueued without waiting: bstart[b] = first sorted position of batch
// b, bstart[nb] = nt; the count nb stays in device memory (*nb_dev) and
// *nb_bound >= nb bounds it for launch sizes.
int tree_target_batches(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                        const int** tperm, int** bstart, int64_t* nb_bound, const int** nb_dev);
// waits for the batches and reads their count
int read_batch_count(lpm_ctx* c, const int* nb_dev, int64_t* nb);
// batch-MAC treecode (treecode_batch.hip): prec LPM_F64, or LPM_F32C with
// FP32 far field and FP32C near field; order <= tree order
int tree_eval_batch_device(lpm_ctx* c, int prec, int order, double theta, double delta, const double* tx,
                           const double* ty, const double* tz, int64_t nt, double* ux, double* uy, double* uz);
int tree_eval_learned_device(lpm_ctx* c, int prec, double theta_max, int margin, int max_order, double u_ref,
                             double delta, const double* tx, const double* ty, const double* tz, int64_t nt,
                             double* ux, double* uy, double* uz);

// ---------------------------------------------------------------- fused FP64 weight
// gamma / den for the LPM_F64F pair: r0 = v_rcp_f64(den) (relative error
// below 2^-24), e = 1 - den r0 by one FMA, and 1 / den = r0 / (1 - e) =
// r0 (1 + e + e^2 + O(e^3)), so gamma / den = w0 + w0 (e + e^2) with
// w0 = gamma r0: within about an ulp of the IEEE quotient (e^3 < 1e-22).
// den == 0 gives NaN; the callers' zero rule (|den| <= 1e-13) discards it.
__device__ __forceinline__ double weight_f64f(double gamma, double den) {
    const double r0 = __builtin_amdgcn_rcp(den);
    const double w0 = gamma * r0;
    const double e = __builtin_fma(-den, r0, 1.0);
    return __builtin_fma(w0, __builtin_fma(e, e, e), w0);
}

// ---------------------------------------------------------------- sphere precondition
// The |x - y|^2 / 2 forms of R^2 - x.y (LPM_F32C, and the Cartesian,
// stereographic and logarithmic series) hold only for points on the sphere
// |x| = R.  Public entry points using them check every point: |x|^2 within
// kSphereTol R^2 of R^2 (positions normalized in FP64 are within a few 1e-16).
constexpr double kSphereTol = 1e-12;
inline bool r2_on_sphere(double r2min, double r2max, double R) {
    const double R2 = R * R;
    return std::fabs(r2min - R2) <= kSphereTol * R2 && std::fabs(r2max - R2) <= kSphereTol * R2;
}
// Block-wide range of r2 (every thread of the block calls it; `valid` false
// for threads without a point) folded into out[0] = max of ~bits(min r2) and
// out[1] = max of bits(max r2): r2 >= 0, so its bits order like its value
// (NaN orders above +inf and fails the check).  Both words start at 0.  The
// caller may spread the blocks over several word pairs (out + 2 (block %
// slots)) so that thousands of blocks do not queue on one address.
__device__ inline void r2_block_range(double r2, bool valid, unsigned long long* out) {
    __shared__ unsigned long long s_lo[16], s_hi[16];
    unsigned long long lo = valid ? ~static_cast<unsigned long long>(__double_as_longlong(r2)) : 0ull;
    unsigned long long hi = valid ? static_cast<unsigned long long>(__double_as_longlong(r2)) : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a > lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_lo[w] = lo; s_hi[w] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < int(blockDim.x >> 6); ++k) {
            lo = s_lo[k] > lo ? s_lo[k] : lo;
            hi = s_hi[k] > hi ? s_hi[k] : hi;
        }
        // most blocks change neither word: skip their atomics
        if (lo > out[0]) atomicMax(&out[0], lo);
        if (hi > out[1]) atomicMax(&out[1], hi);
    }
}
inline double r2_from_words(const unsigned long long* w, int which, int slots = 1) {
    unsigned long long lo = 0, hi = 0;
    for (int k = 0; k < slots; ++k) {
        lo = w[2 * k] > lo ? w[2 * k] : lo;
        hi = w[2 * k + 1] > hi ? w[2 * k + 1] : hi;
    }
    const unsigned long long b = which == 0 ? ~lo : hi;
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}
// min and max of x^2 + y^2 + z^2 over n device points, through the two
// device words w (cleared by a kernel, read back through the pinned words;
// waits for the stream)
int r2_range(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n,
             unsigned long long* w, double* r2min, double* r2max);
// LPM_ERR_INVALID_ARG unless every one of the n device points lies on |x| = R
// (the check's words in the context's buffer, or in w)
int check_on_sphere(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double R,
                    const char* what);
int check_on_sphere_words(lpm_ctx* c, const double* x, const double* y, const double* z, int64_t n, double R,
                          unsigned long long* w, const char* what);

// The sphere precondition of the treecode forms that need it (every form but
// LPM_F64 with LPM_SERIES_REFERENCE): the tree's sources (range recorded by
// the build) and the targets on the unit sphere.  Targets that are the tree's
// sources are not read again; others are (one wait for the stream).
int check_tree_sphere(lpm_ctx* c, const double* tx, const double* ty, const double* tz, int64_t nt,
                      const char* what);

void model_free(Model* m);
const unsigned short* model_weights(lpm_ctx* c);  // device BF16 weights + FP32 biases, or nullptr
const float* model_norm(lpm_ctx* c, int which);   // host feature mean (0), inverse std (1), range lo (2), hi (3)

}  // namespace lpm
