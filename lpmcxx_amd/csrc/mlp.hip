// Training and batched inference of the order model on the MFMA units.
//
// Forward (per 16-row tile, v_mfma_f32_16x16x16_bf16, FP32 accumulate):
//   H1 = bf16(relu(bf16(X) W1 + b1)), H2 = bf16(relu(H1 W2 + b2)), Z = H2 W3 + b3
// Loss: mean softmax cross entropy over the 13 used classes.
// Backward (same units, K = rows for the weight gradients):
//   dZ = bf16((softmax - onehot) / B);  dW3 = H2^T dZ, db3 = sum dZ
//   dZ2 = bf16((dZ W3^T) * [H2 > 0]);   dW2 = H1^T dZ2, db2 = sum dZ2
//   dZ1 = bf16((dZ2 W2^T) * [H1 > 0]);  dW1 = X^T dZ1,  db1 = sum dZ1
// One wave per 64 rows, activations staged in LDS (36 KB per wave); the
// per-wave gradients go to a partial buffer that a second kernel sums in
// wave order (deterministic) before an Adam step on FP32 master weights,
// which then refreshes the BF16 copy the kernels read.
#include "lpm_mlp.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

namespace lpm {
namespace {

constexpr int kRows = 64;  // rows per wave / workgroup
constexpr size_t kLds = size_t(kRows) * (kIn + kHid + kHid + kOut + kHid + kHid) * sizeof(unsigned short);

inline unsigned blocks(int64_t n, int b) { return unsigned((n + b - 1) / b); }

struct Frag {
    const unsigned short *W1, *W2, *W3;
    const float *b1, *b2, *b3;
};
__device__ inline Frag frag(const unsigned short* w) {
    Frag f;
    f.W1 = w;
    f.W2 = w + kIn * kHid;
    f.W3 = f.W2 + kHid * kHid;
    f.b1 = reinterpret_cast<const float*>(w + kBfWeights);
    f.b2 = f.b1 + kHid;
    f.b3 = f.b2 + kHid;
    return f;
}

// Forward of the 64 rows staged in Xs; fills H1s, H2s and returns the logits
// of row tile rt in D layout (lane: rows 4g+i, class c).
__device__ inline void forward_tiles(const Frag& f, const unsigned short* Xs, unsigned short* H1s,
                                     unsigned short* H2s, f32x4 logits[4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        bf16x4 a;
        for (int i = 0; i < 4; ++i) a[i] = short(Xs[(r0 + c) * kIn + 4 * g + i]);
        for (int nt = 0; nt < 4; ++nt) {
            bf16x4 b;
            for (int i = 0; i < 4; ++i) b[i] = short(f.W1[(4 * g + i) * kHid + nt * 16 + c]);
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(a, b, acc);
            for (int i = 0; i < 4; ++i)
                H1s[(r0 + 4 * g + i) * kHid + nt * 16 + c] = f2bf(fmaxf(acc[i] + f.b1[nt * 16 + c], 0.f));
        }
    }
    wave_sync();
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < 4; ++kt) {
                bf16x4 a, b;
                for (int i = 0; i < 4; ++i) {
                    a[i] = short(H1s[(r0 + c) * kHid + kt * 16 + 4 * g + i]);
                    b[i] = short(f.W2[(kt * 16 + 4 * g + i) * kHid + nt * 16 + c]);
                }
                acc = mfma16(a, b, acc);
            }
            for (int i = 0; i < 4; ++i)
                H2s[(r0 + 4 * g + i) * kHid + nt * 16 + c] = f2bf(fmaxf(acc[i] + f.b2[nt * 16 + c], 0.f));
        }
    }
    wave_sync();
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 4; ++kt) {
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                a[i] = short(H2s[(r0 + c) * kHid + kt * 16 + 4 * g + i]);
                b[i] = short(f.W3[(kt * 16 + 4 * g + i) * kOut + c]);
            }
            acc = mfma16(a, b, acc);
        }
        for (int i = 0; i < 4; ++i) logits[rt][i] = acc[i] + f.b3[c];
    }
}

__device__ inline void load_rows(const float* __restrict__ X, int64_t row0, int64_t nrows, unsigned short* Xs) {
    const int lane = threadIdx.x & 63;
    for (int k = 0; k < kIn; ++k) {
        const int64_t r = row0 + lane;
        Xs[lane * kIn + k] = (lane < nrows) ? f2bf(X[r * kIn + k]) : (unsigned short)0;
    }
    wave_sync();
}

// partial[wave][kNParams + 1]: gradients of the mean loss over `batch` rows and the loss sum.
__global__ __launch_bounds__(64) void mlp_train_kernel(const float* __restrict__ X, const int* __restrict__ Y,
                                                       int64_t row0, int64_t batch,
                                                       const unsigned short* __restrict__ w,
                                                       float* __restrict__ partial) {
    extern __shared__ unsigned short lds[];
    unsigned short* Xs = lds;
    unsigned short* H1s = Xs + kRows * kIn;
    unsigned short* H2s = H1s + kRows * kHid;
    unsigned short* Ps = H2s + kRows * kHid;
    unsigned short* D2s = Ps + kRows * kOut;
    unsigned short* D1s = D2s + kRows * kHid;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int64_t first = int64_t(blockIdx.x) * kRows;
    const int64_t nrows = std::min<int64_t>(kRows, batch - first);
    const Frag f = frag(w);
    load_rows(X, row0 + first, nrows, Xs);
    f32x4 logits[4];
    forward_tiles(f, Xs, H1s, H2s, logits);
    const float inv_b = 1.0f / float(batch);
    float loss = 0.f;
    for (int rt = 0; rt < 4; ++rt) {
        for (int i = 0; i < 4; ++i) {
            const int r = rt * 16 + 4 * g + i;
            const bool valid = r < nrows;
            const int label = valid ? Y[row0 + first + r] : -1;
            float v = (c < kClasses) ? logits[rt][i] : -3.0e38f;
            float mx = v;
            for (int o = 8; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            const float e = (c < kClasses) ? expf(v - mx) : 0.f;
            float s = e;
            for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            const float pr = e / s;
            if (valid && c == label) loss += -(v - mx - logf(s));
            const float dz = valid ? (pr - (c == label ? 1.f : 0.f)) * inv_b : 0.f;
            Ps[r * kOut + c] = f2bf(dz);
        }
    }
    wave_sync();
    f32x4 dW1[4], dW2[16], dW3[4];
    for (int k = 0; k < 4; ++k) dW1[k] = dW3[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 16; ++k) dW2[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    // layer 3 gradients and dZ2
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        for (int mt = 0; mt < 4; ++mt) {
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                a[i] = short(H2s[(r0 + 4 * g + i) * kHid + mt * 16 + c]);
                b[i] = short(Ps[(r0 + 4 * g + i) * kOut + c]);
            }
            dW3[mt] = mfma16(a, b, dW3[mt]);
        }
        for (int nt = 0; nt < 4; ++nt) {
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                a[i] = short(Ps[(r0 + c) * kOut + 4 * g + i]);
                b[i] = short(f.W3[(nt * 16 + c) * kOut + 4 * g + i]);
            }
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(a, b, acc);
            for (int i = 0; i < 4; ++i) {
                const int idx = (r0 + 4 * g + i) * kHid + nt * 16 + c;
                D2s[idx] = f2bf(H2s[idx] != 0 ? acc[i] : 0.f);
            }
        }
    }
    wave_sync();
    // layer 2 gradients and dZ1
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        for (int mt = 0; mt < 4; ++mt)
            for (int nt = 0; nt < 4; ++nt) {
                bf16x4 a, b;
                for (int i = 0; i < 4; ++i) {
                    a[i] = short(H1s[(r0 + 4 * g + i) * kHid + mt * 16 + c]);
                    b[i] = short(D2s[(r0 + 4 * g + i) * kHid + nt * 16 + c]);
                }
                dW2[mt * 4 + nt] = mfma16(a, b, dW2[mt * 4 + nt]);
            }
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < 4; ++kt) {
                bf16x4 a, b;
                for (int i = 0; i < 4; ++i) {
                    a[i] = short(D2s[(r0 + c) * kHid + kt * 16 + 4 * g + i]);
                    b[i] = short(f.W2[(nt * 16 + c) * kHid + kt * 16 + 4 * g + i]);
                }
                acc = mfma16(a, b, acc);
            }
            for (int i = 0; i < 4; ++i) {
                const int idx = (r0 + 4 * g + i) * kHid + nt * 16 + c;
                D1s[idx] = f2bf(H1s[idx] != 0 ? acc[i] : 0.f);
            }
        }
    }
    wave_sync();
    // layer 1 gradients
    for (int rt = 0; rt < 4; ++rt) {
        const int r0 = rt * 16;
        for (int nt = 0; nt < 4; ++nt) {
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                a[i] = short(Xs[(r0 + 4 * g + i) * kIn + c]);
                b[i] = short(D1s[(r0 + 4 * g + i) * kHid + nt * 16 + c]);
            }
            dW1[nt] = mfma16(a, b, dW1[nt]);
        }
    }
    // bias gradients: column sums over the wave's rows
    float db1 = 0.f, db2 = 0.f, db3 = 0.f;
    for (int r = 0; r < kRows; ++r) {
        db1 += bf2f(D1s[r * kHid + lane]);
        db2 += bf2f(D2s[r * kHid + lane]);
        if (lane < kOut) db3 += bf2f(Ps[r * kOut + lane]);
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    float* out = partial + int64_t(blockIdx.x) * (kNParams + 1);
    for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) out[kOffW1 + (4 * g + i) * kHid + nt * 16 + c] = dW1[nt][i];
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt)
            for (int i = 0; i < 4; ++i)
                out[kOffW2 + (mt * 16 + 4 * g + i) * kHid + nt * 16 + c] = dW2[mt * 4 + nt][i];
    for (int mt = 0; mt < 4; ++mt)
        for (int i = 0; i < 4; ++i) out[kOffW3 + (mt * 16 + 4 * g + i) * kOut + c] = dW3[mt][i];
    out[kOffB1 + lane] = db1;
    out[kOffB2 + lane] = db2;
    if (lane < kOut) out[kOffB3 + lane] = db3;
    if (lane == 0) out[kNParams] = loss;
}

// grad[p] = sum over waves (in order); with adam != 0 also an Adam step.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ partial, int nparts,
                                                         float* __restrict__ grad, float* __restrict__ master,
                                                         float* __restrict__ m, float* __restrict__ v,
                                                         unsigned short* __restrict__ wbf, int adam, float lr,
                                                         float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kNParams) return;
    float s = 0.f;
    for (int k = 0; k < nparts; ++k) s += partial[int64_t(k) * (kNParams + 1) + p];
    grad[p] = s;
    if (!adam || p == kNParams) return;
    const float b1 = 0.9f, b2 = 0.999f;
    m[p] = b1 * m[p] + (1.f - b1) * s;
    v[p] = b2 * v[p] + (1.f - b2) * s * s;
    const float w = master[p] - lr * (m[p] / bc1) / (sqrtf(v[p] / bc2) + 1e-8f);
    master[p] = w;
    float* fb = reinterpret_cast<float*>(wbf + kBfWeights);
    if (p < kOffB1) wbf[p] = f2bf(w);
    else if (p < kOffW2) fb[p - kOffB1] = w;
    else if (p < kOffB2) wbf[kIn * kHid + (p - kOffW2)] = f2bf(w);
    else if (p < kOffW3) fb[kHid + (p - kOffB2)] = w;
    else if (p < kOffB3) wbf[kIn * kHid + kHid * kHid + (p - kOffW3)] = f2bf(w);
    else fb[2 * kHid + (p - kOffB3)] = w;
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ master,
                                                       unsigned short* __restrict__ wbf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParams) return;
    const float w = master[p];
    float* fb = reinterpret_cast<float*>(wbf + kBfWeights);
    if (p < kOffB1) wbf[p] = f2bf(w);
    else if (p < kOffW2) fb[p - kOffB1] = w;
    else if (p < kOffB2) wbf[kIn * kHid + (p - kOffW2)] = f2bf(w);
    else if (p < kOffW3) fb[kHid + (p - kOffB2)] = w;
    else if (p < kOffB3) wbf[kIn * kHid + kHid * kHid + (p - kOffW3)] = f2bf(w);
    else fb[2 * kHid + (p - kOffB3)] = w;
}

// Classes of n (normalized) rows by mlp_classify_regs, the inference of the
// planning pass: 16 rows per wave, lane (g, r) holding inputs 4g..4g+3 of
// row r.
__global__ __launch_bounds__(64) void mlp_classify_kernel(const float* __restrict__ X, int64_t n,
                                                          const unsigned short* __restrict__ w,
                                                          int* __restrict__ cls) {
    __shared__ __align__(16) unsigned char s_model[kModelLds];
    mlp_stage_weights(w, s_model);
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int64_t row = int64_t(blockIdx.x) * 16 + r;
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (row < n)
        for (int i = 0; i < 4; ++i) x[i] = X[row * kIn + 4 * g + i];
    const int q = mlp_classify_regs(s_model, cvt_bf16x4(x[0], x[1], x[2], x[3]));
    if (lane < 16 && row < n) cls[row] = q;
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ X, int64_t n,
                                                         const unsigned short* __restrict__ w,
                                                         float* __restrict__ logits, int* __restrict__ cls) {
    extern __shared__ unsigned short lds[];
    unsigned short* Xs = lds;
    unsigned short* H1s = Xs + kRows * kIn;
    unsigned short* H2s = H1s + kRows * kHid;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int64_t first = int64_t(blockIdx.x) * kRows;
    const int64_t nrows = std::min<int64_t>(kRows, n - first);
    const Frag f = frag(w);
    load_rows(X, first, nrows, Xs);
    f32x4 lg[4];
    forward_tiles(f, Xs, H1s, H2s, lg);
    for (int rt = 0; rt < 4; ++rt)
        for (int i = 0; i < 4; ++i) {
            const int r = rt * 16 + 4 * g + i;
            float v = (c < kClasses) ? lg[rt][i] : -3.0e38f;
            int best = c;
            for (int o = 8; o > 0; o >>= 1) {
                const float ov = __shfl_xor(v, o, 64);
                const int ob = __shfl_xor(best, o, 64);
                if (ov > v || (ov == v && ob < best)) { v = ov; best = ob; }
            }
            if (r < nrows) {
                if (logits) logits[(first + r) * kOut + c] = lg[rt][i];
                if (cls && c == 0) cls[first + r] = best;
            }
        }
}

}  // namespace

static Model& model(lpm_ctx* c) {
    if (!c->model) c->model = new Model();
    return *c->model;
}

void model_free(Model* m) {
    if (!m) return;
    for (DevBuf* b : {&m->master, &m->adam_m, &m->adam_v, &m->wbf, &m->partial, &m->data}) b->release();
    delete m;
}

const unsigned short* model_weights(lpm_ctx* c) {
    return (c->model && c->model->ready) ? c->model->wbf.as<unsigned short>() : nullptr;
}
const float* model_norm(lpm_ctx* c, int which) {
    const Model& M = *c->model;
    return which == 0 ? M.mean : which == 1 ? M.inv_std : which == 2 ? M.lo : M.hi;
}

static void reset_range(Model& M) {
    for (int i = 0; i < kIn; ++i) { M.lo[i] = -kRangeMax; M.hi[i] = kRangeMax; }
}

static int upload_params(lpm_ctx* c, Model& M, const float* params) {
    LPM_CHECK(M.master.reserve(kNParams * sizeof(float)));
    LPM_CHECK(M.adam_m.reserve(kNParams * sizeof(float)));
    LPM_CHECK(M.adam_v.reserve(kNParams * sizeof(float)));
    LPM_CHECK(M.wbf.reserve(kBfWeights * sizeof(unsigned short) + (2 * kHid + kOut) * sizeof(float)));
    hipStream_t st = c->stream;
    LPM_HIP(c, hipMemcpyAsync(M.master.ptr, params, kNParams * sizeof(float), hipMemcpyHostToDevice, st));
    LPM_HIP(c, hipMemsetAsync(M.adam_m.ptr, 0, kNParams * sizeof(float), st));
    LPM_HIP(c, hipMemsetAsync(M.adam_v.ptr, 0, kNParams * sizeof(float), st));
    hipLaunchKernelGGL(mlp_pack_kernel, dim3(blocks(kNParams, 256)), dim3(256), 0, st, M.master.as<float>(),
                       M.wbf.as<unsigned short>());
    LPM_HIP(c, hipStreamSynchronize(st));
    M.step = 0;
    M.ready = true;
    return LPM_OK;
}

}  // namespace lpm

using namespace lpm;

extern "C" {

int lpm_model_init(lpm_ctx* c, uint64_t seed) {
    if (!c) return LPM_ERR_NULL_PTR;
    LPM_HIP(c, hipSetDevice(c->device));
    Model& M = model(c);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> p(kNParams, 0.f);
    auto he = [&](int off, int fan_in, int n) {
        const float s = std::sqrt(2.0f / fan_in);
        for (int i = 0; i < n; ++i) p[off + i] = s * nd(rng);
    };
    he(kOffW1, kFeatures, kIn * kHid);
    for (int i = kFeatures; i < kIn; ++i)
        for (int j = 0; j < kHid; ++j) p[kOffW1 + i * kHid + j] = 0.f;  // padded inputs
    he(kOffW2, kHid, kHid * kHid);
    he(kOffW3, kHid, kHid * kOut);
    for (int i = 0; i < kIn; ++i) { M.mean[i] = 0.f; M.inv_std[i] = 1.f; }
    reset_range(M);
    return upload_params(c, M, p.data());
}

int lpm_model_set_params(lpm_ctx* c, const float* params, const float* mean, const float* inv_std) {
    if (!c || !params) return LPM_ERR_NULL_PTR;
    LPM_HIP(c, hipSetDevice(c->device));
    Model& M = model(c);
    for (int i = 0; i < kIn; ++i) {
        M.mean[i] = mean ? mean[i] : 0.f;
        M.inv_std[i] = inv_std ? inv_std[i] : 1.f;
    }
    reset_range(M);
    return upload_params(c, M, params);
}

int lpm_model_set_range(lpm_ctx* c, const float* lo, const float* hi) {
    if (!c) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    if (!lo != !hi) return fail(c, LPM_ERR_NULL_PTR, "lo and hi are both given or both NULL");
    Model& M = *c->model;
    if (!lo) {
        reset_range(M);
        return LPM_OK;
    }
    for (int i = 0; i < kIn; ++i)
        if (!(lo[i] <= hi[i])) return fail(c, LPM_ERR_INVALID_ARG, "input range with lo > hi or NaN");
    for (int i = 0; i < kIn; ++i) { M.lo[i] = lo[i]; M.hi[i] = hi[i]; }
    return LPM_OK;
}

int lpm_model_get_range(lpm_ctx* c, float* lo, float* hi) {
    if (!c || !lo || !hi) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    for (int i = 0; i < kIn; ++i) { lo[i] = c->model->lo[i]; hi[i] = c->model->hi[i]; }
    return LPM_OK;
}

int lpm_model_get_params(lpm_ctx* c, float* params, float* mean, float* inv_std) {
    if (!c || !params) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    LPM_HIP(c, hipSetDevice(c->device));
    LPM_HIP(c, hipMemcpyAsync(params, c->model->master.ptr, kNParams * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < kIn; ++i) {
        if (mean) mean[i] = c->model->mean[i];
        if (inv_std) inv_std[i] = c->model->inv_std[i];
    }
    return LPM_OK;
}

// Rows x[n][16] are raw features: the model's clamp to its input range and
// normalization (x - mean) * inv_std are applied on upload, as in the
// treecode's planning pass.
static int upload_rows(lpm_ctx* c, Model& M, const float* x, const int32_t* y, int64_t n) {
    std::vector<float> xn(size_t(n) * kIn);
    for (int64_t r = 0; r < n; ++r)
        for (int k = 0; k < kIn; ++k)
            xn[r * kIn + k] = (std::min(std::max(x[r * kIn + k], M.lo[k]), M.hi[k]) - M.mean[k]) * M.inv_std[k];
    LPM_CHECK(M.data.reserve(size_t(n) * (kIn * sizeof(float) + sizeof(int32_t)) + 64));
    float* dx = M.data.as<float>();
    int* dy = reinterpret_cast<int*>(dx + size_t(n) * kIn);
    LPM_HIP(c, hipMemcpyAsync(dx, xn.data(), xn.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (y) LPM_HIP(c, hipMemcpyAsync(dy, y, size_t(n) * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_model_gradient(lpm_ctx* c, const float* x, const int32_t* y, int64_t n, float* grad, float* loss) {
    if (!c || !x || !y || !grad) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    if (n < 1) return fail(c, LPM_ERR_INVALID_ARG, "empty batch");
    for (int64_t r = 0; r < n; ++r)
        if (y[r] < 0 || y[r] >= kClasses) return fail(c, LPM_ERR_INVALID_ARG, "label out of range");
    LPM_HIP(c, hipSetDevice(c->device));
    Model& M = *c->model;
    LPM_CHECK(upload_rows(c, M, x, y, n));
    const int nw = int((n + kRows - 1) / kRows);
    LPM_CHECK(M.partial.reserve(size_t(nw + 1) * (kNParams + 1) * sizeof(float)));
    float* part = M.partial.as<float>();
    float* g = part + size_t(nw) * (kNParams + 1);
    const float* dx = M.data.as<float>();
    const int* dy = reinterpret_cast<const int*>(dx + size_t(n) * kIn);
    hipLaunchKernelGGL(mlp_train_kernel, dim3(nw), dim3(64), kLds, c->stream, dx, dy, int64_t(0), n,
                       M.wbf.as<unsigned short>(), part);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3(blocks(kNParams + 1, 256)), dim3(256), 0, c->stream, part, nw, g,
                       nullptr, nullptr, nullptr, nullptr, 0, 0.f, 1.f, 1.f);
    std::vector<float> h(kNParams + 1);
    LPM_HIP(c, hipMemcpyAsync(h.data(), g, (kNParams + 1) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    std::copy(h.begin(), h.begin() + kNParams, grad);
    if (loss) *loss = h[kNParams] / float(n);
    return LPM_OK;
}

int lpm_model_train(lpm_ctx* c, const float* x, const int32_t* y, int64_t n, int epochs, int batch, float lr,
                    uint64_t seed, float* final_loss) {
    if (!c || !x || !y) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    if (n < 1 || epochs < 0 || batch < 1 || !(lr > 0.f)) return fail(c, LPM_ERR_INVALID_ARG, "bad training args");
    for (int64_t r = 0; r < n; ++r)
        if (y[r] < 0 || y[r] >= kClasses) return fail(c, LPM_ERR_INVALID_ARG, "label out of range");
    for (int64_t i = 0; i < n * kIn; ++i)
        if (!std::isfinite(x[i])) return fail(c, LPM_ERR_INVALID_ARG, "non-finite input");
    LPM_HIP(c, hipSetDevice(c->device));
    Model& M = *c->model;
    // the input range becomes that of the training rows
    for (int k = 0; k < kIn; ++k) { M.lo[k] = x[k]; M.hi[k] = x[k]; }
    for (int64_t r = 1; r < n; ++r)
        for (int k = 0; k < kIn; ++k) {
            M.lo[k] = std::min(M.lo[k], x[r * kIn + k]);
            M.hi[k] = std::max(M.hi[k], x[r * kIn + k]);
        }
    std::vector<int64_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    std::vector<float> xs(size_t(n) * kIn);
    std::vector<int32_t> ys(n);
    const int nw_max = (batch + kRows - 1) / kRows;
    LPM_CHECK(M.partial.reserve(size_t(nw_max + 1) * (kNParams + 1) * sizeof(float)));
    float* part = M.partial.as<float>();
    float* g = part + size_t(nw_max) * (kNParams + 1);
    double last = 0.0;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        for (int64_t r = 0; r < n; ++r) {
            std::copy(x + perm[r] * kIn, x + (perm[r] + 1) * kIn, xs.begin() + r * kIn);
            ys[r] = y[perm[r]];
        }
        LPM_CHECK(upload_rows(c, M, xs.data(), ys.data(), n));
        const float* dx = M.data.as<float>();
        const int* dy = reinterpret_cast<const int*>(dx + size_t(n) * kIn);
        double esum = 0.0;
        for (int64_t r0 = 0; r0 < n; r0 += batch) {
            const int64_t b = std::min<int64_t>(batch, n - r0);
            const int nw = int((b + kRows - 1) / kRows);
            hipLaunchKernelGGL(mlp_train_kernel, dim3(nw), dim3(64), kLds, c->stream, dx, dy, r0, b,
                               M.wbf.as<unsigned short>(), part);
            ++M.step;
            const float bc1 = 1.f - std::pow(0.9f, float(M.step)), bc2 = 1.f - std::pow(0.999f, float(M.step));
            hipLaunchKernelGGL(mlp_reduce_kernel, dim3(blocks(kNParams + 1, 256)), dim3(256), 0, c->stream, part,
                               nw, g, M.master.as<float>(), M.adam_m.as<float>(), M.adam_v.as<float>(),
                               M.wbf.as<unsigned short>(), 1, lr, bc1, bc2);
            float hl = 0.f;
            LPM_HIP(c, hipMemcpyAsync(&hl, g + kNParams, sizeof(float), hipMemcpyDeviceToHost, c->stream));
            LPM_HIP(c, hipStreamSynchronize(c->stream));
            esum += hl;
        }
        last = esum / double(n);
    }
    LPM_HIP(c, hipGetLastError());
    if (final_loss) *final_loss = float(last);
    return LPM_OK;
}

int lpm_model_forward(lpm_ctx* c, const float* x, int64_t n, float* logits, int32_t* cls) {
    if (!c || !x) return LPM_ERR_NULL_PTR;
    if (!c->model || !c->model->ready) return fail(c, LPM_ERR_STATE, "no model");
    if (n < 0) return fail(c, LPM_ERR_INVALID_ARG, "n < 0");
    if (n == 0) return LPM_OK;
    LPM_HIP(c, hipSetDevice(c->device));
    Model& M = *c->model;
    LPM_CHECK(upload_rows(c, M, x, nullptr, n));
    LPM_CHECK(c->h_out.reserve(size_t(n) * (kOut * sizeof(float) + sizeof(int))));
    float* dl = c->h_out.as<float>();
    int* dc = reinterpret_cast<int*>(dl + size_t(n) * kOut);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(unsigned((n + kRows - 1) / kRows)), dim3(64), kLds, c->stream,
                       M.data.as<float>(), n, M.wbf.as<unsigned short>(), dl, nullptr);
    // the classes come from the planning pass's own inference
    hipLaunchKernelGGL(mlp_classify_kernel, dim3(unsigned((n + 15) / 16)), dim3(64), 0, c->stream,
                       M.data.as<float>(), n, M.wbf.as<unsigned short>(), dc);
    if (logits)
        LPM_HIP(c, hipMemcpyAsync(logits, dl, size_t(n) * kOut * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (cls) LPM_HIP(c, hipMemcpyAsync(cls, dc, size_t(n) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

}  // extern "C"
