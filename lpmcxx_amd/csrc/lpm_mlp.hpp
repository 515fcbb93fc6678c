// Order model: a 16 -> 64 -> 64 -> 16 ReLU perceptron (10 features, 13
// classes) that picks the Taylor order p (class p, 1..12) or rejects the
// multipole acceptance (class 0) for a (wave of targets, cluster) pair.
// BF16 weights and activations on the MFMA units (v_mfma_f32_16x16x16_bf16
// for training and the first layer, v_mfma_f32_16x16x32_bf16 for the K = 64
// layers of the inference in the planning pass), FP32 accumulation, FP32
// biases and master weights.
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
    // input range: raw inputs are clamped to [lo, hi] before normalization
    float lo[kIn], hi[kIn];
    DevBuf master;   // float [kNParams]
    DevBuf adam_m, adam_v;
    DevBuf wbf;      // ushort [kBfWeights] + float biases [64 + 64 + 16] (+ pad)
    DevBuf partial;  // float [nwg][kNParams + 1]
    DevBuf data;     // training rows / labels
};

// Unbounded input range (a model without a recorded training range).
constexpr float kRangeMax = 3.0e38f;

// feature normalization passed to the kernels by value:
// x' = (min(max(x, lo), hi) - mean) * inv_std
struct MlpNorm {
    float mean[kIn];
    float inv_std[kIn];
    float lo[kIn];
    float hi[kIn];
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

// Weights of the model as MFMA A fragments in LDS (the layers are computed
// transposed, see mlp_classify_regs): W1 as four 16x16x16 tiles (64 lanes x
// 4 BF16 each), W2 as eight 16x16x32 tiles (tile 2 nt + kb: output tile nt,
// K block kb; 64 lanes x 8 BF16), W3 as two 16x16x32 tiles (K block kb),
// then the FP32 biases b1, b2, b3.
// A 16x16x32 K block kb holds the 32 features 32 kb .. 32 kb + 31 in the
// order in which the previous layer's 16x16 results hand them over: lane
// group g's eight values are features 32 kb + 4 g + i and 32 kb + 16 + 4 g
// + i (i = 0..3), so its B operand is two result fragments side by side.
constexpr int kFragW1 = 4 * 64 * 8, kFragW2 = 8 * 64 * 16, kFragW3 = 2 * 64 * 16;
constexpr int kFragBytes = kFragW1 + kFragW2 + kFragW3;
constexpr int kModelLds = kFragBytes + (2 * kHid + kOut) * 4;

// feature of K block kb held by lane group g as its element j (16x16x32 operands)
__host__ __device__ constexpr int kblock_feature(int kb, int g, int j) {
    return 32 * kb + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// Fill the LDS model image from the device copy (BF16 weights + FP32
// biases); all threads of the block take part, the caller synchronizes.
__device__ inline void mlp_stage_weights(const unsigned short* __restrict__ w, unsigned char* lds) {
    uint2* f1 = reinterpret_cast<uint2*>(lds);
    uint4* f2 = reinterpret_cast<uint4*>(lds + kFragW1);
    uint4* f3 = reinterpret_cast<uint4*>(lds + kFragW1 + kFragW2);
    const unsigned short* W1 = w;
    const unsigned short* W2 = w + kIn * kHid;
    const unsigned short* W3 = W2 + kHid * kHid;
    for (int f = threadIdx.x; f < 14 * 64; f += blockDim.x) {
        const int l = f & 63, g = l >> 4, c = l & 15;
        unsigned short e[8];
        if (f < 4 * 64) {
            const int nt = f >> 6;
            for (int i = 0; i < 4; ++i) e[i] = W1[(4 * g + i) * kHid + nt * 16 + c];
            f1[f] = make_uint2(unsigned(e[0]) | (unsigned(e[1]) << 16), unsigned(e[2]) | (unsigned(e[3]) << 16));
            continue;
        }
        const bool l2 = f < 12 * 64;
        const int tile = l2 ? (f >> 6) - 4 : (f >> 6) - 12;
        const int nt = l2 ? tile >> 1 : 0, kb = l2 ? tile & 1 : tile;
        for (int j = 0; j < 8; ++j) {
            const int k = kblock_feature(kb, g, j);
            e[j] = l2 ? W2[k * kHid + nt * 16 + c] : W3[k * kOut + c];
        }
        const uint4 v = make_uint4(unsigned(e[0]) | (unsigned(e[1]) << 16), unsigned(e[2]) | (unsigned(e[3]) << 16),
                                   unsigned(e[4]) | (unsigned(e[5]) << 16), unsigned(e[6]) | (unsigned(e[7]) << 16));
        if (l2) f2[f - 4 * 64] = v;
        else f3[f - 12 * 64] = v;
    }
    const float* fb = reinterpret_cast<const float*>(w + kBfWeights);
    float* lb = reinterpret_cast<float*>(lds + kFragBytes);
    for (int i = threadIdx.x; i < 2 * kHid + kOut; i += blockDim.x) lb[i] = fb[i];
}

__device__ inline bf16x4 as_bf16x4(uint2 v) {
    bf16x4 r;
    r[0] = short(v.x & 0xFFFFu);
    r[1] = short(v.x >> 16);
    r[2] = short(v.y & 0xFFFFu);
    r[3] = short(v.y >> 16);
    return r;
}

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));

// four floats to BF16 (round to nearest even, the hardware conversion)
__device__ inline bf16x4 cvt_bf16x4(float a, float b, float c, float d) {
    const bf16x2v lo = __builtin_convertvector((f32x2v){a, b}, bf16x2v);
    const bf16x2v hi = __builtin_convertvector((f32x2v){c, d}, bf16x2v);
    const unsigned ulo = __builtin_bit_cast(unsigned, lo), uhi = __builtin_bit_cast(unsigned, hi);
    bf16x4 r;
    r[0] = short(ulo & 0xFFFFu);
    r[1] = short(ulo >> 16);
    r[2] = short(uhi & 0xFFFFu);
    r[3] = short(uhi >> 16);
    return r;
}

// two floats to BF16 (round to nearest even), a in the low half
__device__ inline unsigned cvt_bf16x2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2v){a, b}, bf16x2v));
}

typedef short s16x2 __attribute__((ext_vector_type(2)));

// ReLU of four FP32 values as BF16: rounded first, then the negative ones
// (sign bit set: negative as 16-bit integers, -0 included) raised to +0 by a
// packed 16-bit integer max, two values per instruction.  For finite values
// this is max(x, 0) rounded; an FP32 max would first quiet its operand (the
// MFMA results are not known to be canonical), two instructions per value.
__device__ inline bf16x4 relu_bf16x4(float a, float b, float c, float d) {
    const bf16x2v lo = __builtin_convertvector((f32x2v){a, b}, bf16x2v);
    const bf16x2v hi = __builtin_convertvector((f32x2v){c, d}, bf16x2v);
    const s16x2 z = {0, 0};
    const s16x2 l = __builtin_elementwise_max(__builtin_bit_cast(s16x2, lo), z);
    const s16x2 h = __builtin_elementwise_max(__builtin_bit_cast(s16x2, hi), z);
    bf16x4 r;
    r[0] = l[0];
    r[1] = l[1];
    r[2] = h[0];
    r[3] = h[1];
    return r;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ inline f32x4 mfma32(uint4 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), b, c, 0, 0, 0);
}

// the B operand of K block kb: result fragments 2 kb and 2 kb + 1 side by side
__device__ inline bf16x8 join_bf16(bf16x4 lo, bf16x4 hi) {
    const s16x8 v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_bit_cast(bf16x8, v);
}

// Register-only inference for up to 16 rows.  Every layer is computed
// transposed, H_out^T = W^T H_in^T, so that the MFMA result layout (lane
// (g, c) holds row c, columns 4g..4g+3) is directly the operand layout of
// the next layer and no LDS transposes are needed.  Layer 1 (K = 16) runs on
// v_mfma_f32_16x16x16_bf16 (4 instructions), layers 2 and 3 (K = 64) on
// gfx950's v_mfma_f32_16x16x32_bf16 (8 + 2 instructions).  x: this lane's
// four normalized inputs, row (lane % 16), columns 4g..4g+3 (g = lane / 16),
// as BF16.  Returns the argmax class of row (lane % 16) in every lane.
__device__ inline int mlp_classify_regs(const unsigned char* model, bf16x4 x) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint2* f1 = reinterpret_cast<const uint2*>(model);
    const uint4* f2 = reinterpret_cast<const uint4*>(model + kFragW1);
    const uint4* f3 = reinterpret_cast<const uint4*>(model + kFragW1 + kFragW2);
    const float* b1 = reinterpret_cast<const float*>(model + kFragBytes);
    const float* b2 = b1 + kHid;
    const float* b3 = b2 + kHid;
    bf16x4 h1[4], h2[4];
    // the biases enter as the accumulators' initial values (the MFMA result
    // layout is the bias layout: lane (g, c) holds outputs 4g..4g+3)
    auto bias = [&](const float* bv) {
        const float4 b = *reinterpret_cast<const float4*>(bv + 4 * g);
        return f32x4{b.x, b.y, b.z, b.w};
    };
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const f32x4 acc = mfma16(as_bf16x4(f1[nt * 64 + lane]), x, bias(b1 + nt * 16));
        h1[nt] = relu_bf16x4(acc[0], acc[1], acc[2], acc[3]);
    }
    const bf16x8 k1[2] = {join_bf16(h1[0], h1[1]), join_bf16(h1[2], h1[3])};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = bias(b2 + nt * 16);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) acc = mfma32(f2[(2 * nt + kb) * 64 + lane], k1[kb], acc);
        h2[nt] = relu_bf16x4(acc[0], acc[1], acc[2], acc[3]);
    }
    f32x4 acc = bias(b3);
    acc = mfma32(f3[lane], join_bf16(h2[0], h2[1]), acc);
    acc = mfma32(f3[64 + lane], join_bf16(h2[2], h2[3]), acc);
    const float z[4] = {acc[0], acc[1], acc[2], acc[3]};
    float v = -3.0e38f;
    int best = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (4 * g + i < kClasses && z[i] > v) { v = z[i]; best = 4 * g + i; }
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
        const float ov = __shfl_xor(v, o, 64);
        const int ob = __shfl_xor(best, o, 64);
        if (ov > v || (ov == v && ob < best)) { v = ov; best = ob; }
    }
    return best;
}

// Inference for up to 16 RT rows in RT row tiles (tile t: rows 16 t ..
// 16 t + 15), with the layout and operations of mlp_classify_regs per tile:
// x[t] holds lane (g, c)'s inputs 4g..4g+3 of row 16 t + c.  The tiles go
// through the layers in groups of GRP: each weight fragment is read from LDS
// once per group and applied to the group's tiles.  Returns in lane l the
// argmax class of row l (for RT = 1, of row l % 16 in every lane).
template <int GRP>
__device__ inline void mlp_tile_group(const unsigned char* model, const bf16x4* x, int t0, int& mine) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint2* f1 = reinterpret_cast<const uint2*>(model);
    const uint4* f2 = reinterpret_cast<const uint4*>(model + kFragW1);
    const uint4* f3 = reinterpret_cast<const uint4*>(model + kFragW1 + kFragW2);
    const float* b1 = reinterpret_cast<const float*>(model + kFragBytes);
    const float* b2 = b1 + kHid;
    const float* b3 = b2 + kHid;
    auto bias = [&](const float* bv) {
        const float4 b = *reinterpret_cast<const float4*>(bv + 4 * g);
        return f32x4{b.x, b.y, b.z, b.w};
    };
    bf16x4 h1[GRP][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const bf16x4 w = as_bf16x4(f1[nt * 64 + lane]);
        const f32x4 b = bias(b1 + nt * 16);
#pragma unroll
        for (int t = 0; t < GRP; ++t) {
            const f32x4 acc = mfma16(w, x[t], b);
            h1[t][nt] = relu_bf16x4(acc[0], acc[1], acc[2], acc[3]);
        }
    }
    f32x4 out[GRP];
#pragma unroll
    for (int t = 0; t < GRP; ++t) out[t] = bias(b3);
    // layer 2 in two halves of two output tiles each; each half is the K
    // block of layer 3 that consumes it
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        bf16x4 h2[GRP][2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int nt = 2 * hf + j;
            f32x4 acc[GRP];
#pragma unroll
            for (int t = 0; t < GRP; ++t) acc[t] = bias(b2 + nt * 16);
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const uint4 w = f2[(2 * nt + kb) * 64 + lane];
#pragma unroll
                for (int t = 0; t < GRP; ++t) acc[t] = mfma32(w, join_bf16(h1[t][2 * kb], h1[t][2 * kb + 1]), acc[t]);
            }
#pragma unroll
            for (int t = 0; t < GRP; ++t) h2[t][j] = relu_bf16x4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
        }
        const uint4 w = f3[hf * 64 + lane];
#pragma unroll
        for (int t = 0; t < GRP; ++t) out[t] = mfma32(w, join_bf16(h2[t][0], h2[t][1]), out[t]);
    }
#pragma unroll
    for (int t = 0; t < GRP; ++t) {
        float v = -3.0e38f;
        int best = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < kClasses && out[t][i] > v) { v = out[t][i]; best = 4 * g + i; }
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            const float ov = __shfl_xor(v, o, 64);
            const int ob = __shfl_xor(best, o, 64);
            if (ov > v || (ov == v && ob < best)) { v = ov; best = ob; }
        }
        if (t0 + t == g) mine = best;
    }
}

template <int RT, int GRP = RT>
__device__ inline int mlp_classify_tiles(const unsigned char* model, const bf16x4 (&x)[RT]) {
    static_assert(RT % GRP == 0, "whole groups of tiles");
    if (RT == 1) return mlp_classify_regs(model, x[0]);
    int mine = 0;
#pragma unroll
    for (int t0 = 0; t0 < RT; t0 += GRP) mlp_tile_group<GRP>(model, x + t0, t0, mine);
    return mine;
}

}  // namespace lpm
