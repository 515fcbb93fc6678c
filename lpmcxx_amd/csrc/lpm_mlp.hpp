// Order model: a 16 -> 64 -> 64 -> 16 ReLU perceptron (10 features, 13
// classes) that picks the Taylor order p (class p, 1..12) or rejects the
// multipole acceptance (class 0) for a (wave of targets, cluster) pair.
// BF16 weights and activations on the MFMA units (v_mfma_f32_16x16x16_bf16),
// FP32 accumulation, FP32 biases and master weights.
#pragma once

#include "lpm_internal.hpp"

namespace lpm {

constexpr int kIn = 16, kHid = 64, kOut = 16, kClasses = 13, kFeatures = 10;
// flat FP32 parameter layout: W1[16][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut, kNParams = kOffB3 + kOut;
// device copy used by the kernels: BF16 weights then FP32 biases
constexpr int kBfWeights = kIn * kHid + kHid * kHid + kHid * kOut;  // 6144 bf16

struct Model {
    bool ready = false;
    int step = 0;
    float mean[kIn] = {}, inv_std[kIn] = {};
    DevBuf master;   // float [kNParams]
    DevBuf adam_m, adam_v;
    DevBuf wbf;      // ushort [kBfWeights] + float biases [64 + 64 + 16] (+ pad)
    DevBuf partial;  // float [nwg][kNParams + 1]
    DevBuf data;     // training rows / labels
};

// feature normalization passed to the kernels by value: x' = (x - mean) * inv_std
struct MlpNorm {
    float mean[kIn];
    float inv_std[kIn];
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Orders LDS traffic between lanes of one wave (no cross-wave barrier).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// MFMA 16x16x16 operand layout (lane l, group g = l / 16, column c = l % 16):
// A[m][k]: lane holds A[c][4g + i]; B[k][n]: lane holds B[4g + i][c];
// D[m][n]: lane holds D[4g + i][c], i = 0..3.

// One row (wave-uniform, normalized features in h[0..15]) through the network
// on the MFMA units, the row placed in tile row 0; returns the argmax class.
// w: BF16 weights + FP32 biases (Model::wbf layout); h: 144 floats of
// wave-private LDS (inputs 0..15, hidden layers 16..79 and 80..143).
__device__ inline int mlp_classify_row(const unsigned short* __restrict__ w, float* h) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const unsigned short* W1 = w;
    const unsigned short* W2 = w + kIn * kHid;
    const unsigned short* W3 = W2 + kHid * kHid;
    const float* fb = reinterpret_cast<const float*>(w + kBfWeights);
    wave_sync();
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = (c == 0) ? short(f2bf(h[4 * g + i])) : short(0);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = short(W1[(4 * g + i) * kHid + nt * 16 + c]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(a, b, acc);
        if (g == 0) h[16 + nt * 16 + c] = fmaxf(acc[0] + fb[nt * 16 + c], 0.f);
    }
    wave_sync();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            bf16x4 b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = (c == 0) ? short(f2bf(h[16 + kt * 16 + 4 * g + i])) : short(0);
                b[i] = short(W2[(kt * 16 + 4 * g + i) * kHid + nt * 16 + c]);
            }
            acc = mfma16(a, b, acc);
        }
        if (g == 0) h[80 + nt * 16 + c] = fmaxf(acc[0] + fb[kHid + nt * 16 + c], 0.f);
    }
    wave_sync();
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
        bf16x4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = (c == 0) ? short(f2bf(h[80 + kt * 16 + 4 * g + i])) : short(0);
            b[i] = short(W3[(kt * 16 + 4 * g + i) * kOut + c]);
        }
        acc = mfma16(a, b, acc);
    }
    float v = (g == 0 && c < kClasses) ? acc[0] + fb[2 * kHid + c] : -3.0e38f;
    int best = c;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        const float ov = __shfl_xor(v, o, 64);
        const int ob = __shfl_xor(best, o, 64);
        if (ov > v || (ov == v && ob < best)) { v = ov; best = ob; }
    }
    wave_sync();
    return __builtin_amdgcn_readfirstlane(best);
}

}  // namespace lpm
