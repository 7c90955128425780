// Device pieces of the learned proposal network shared by the MFMA kernels
// (ud2_mlp.hip) and the fused learned MC kernel (ud2_learned.hip).
//
// Network: y = w3 . relu(W2^T a1 + b2) + b3, a1 = bf16(relu(W1^T xn + b1)),
// xn = bf16((x - mu) * inv_sigma); 80 inputs, two hidden layers of 256.
//
// Tiling: one wave scores a group of 16 candidates. Every layer is computed
// transposed, out^T[hidden][cand] = W^T[hidden][k] . in^T[k][cand], with
// v_mfma_f32_16x16x16_bf16: the weights are the A operand (lane l holds row
// l % 16, k = 4 (l / 16) .. +3), the activations the B operand (lane l holds
// candidate l % 16, the same k). The accumulator of a 16 x 16 output tile puts
// rows 4 (l / 16) .. +3 of candidate l % 16 on lane l, which is exactly the B
// operand of the next layer for the K block of that tile, so hidden
// activations never leave the registers. Weights stream from L2 (166 KB in
// bf16) as 8-byte loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ud2_internal.h"

namespace ud2 {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ float bf2f_s(short b) { return bf2f(uint16_t(b)); }

// Round to nearest even; NaN stays NaN (quietened), overflow rounds to inf.
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) {
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        return uint16_t(u >> 16);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Sum over the 16 lanes of a lane row (lanes 16 q .. 16 q + 15).
__device__ __forceinline__ float row16_sum(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// Normalised inputs of candidate `row` as B operands of the 5 K blocks of layer 1.
__device__ __forceinline__ void load_inputs(const MlpDev& W, const uint16_t* xr, int q, bf16x4 xb[5]) {
#pragma unroll
    for (int ks = 0; ks < 5; ks++) {
        const bf16x4 raw = ld4(xr + 16 * ks + 4 * q);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int f = 16 * ks + 4 * q + i;
            xb[ks][i] = short(f2bf((bf2f_s(raw[i]) - W.mu[f]) * W.inv_sigma[f]));
        }
    }
}

// Layer 1: a1 tiles (hidden 16 mt + 4 q + i of candidate `row` on this lane).
__device__ __forceinline__ void layer1(const MlpDev& W, const bf16x4 xb[5], int row, int q, bf16x4 a1[16]) {
#pragma unroll
    for (int mt = 0; mt < 16; mt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* wr = W.w1t + (16 * mt + row) * kFeat + 4 * q;
#pragma unroll
        for (int ks = 0; ks < 5; ks++) acc = mfma16(ld4(wr + 16 * ks), xb[ks], acc);
#pragma unroll
        for (int i = 0; i < 4; i++) a1[mt][i] = short(f2bf(fmaxf(acc[i] + W.b1[16 * mt + 4 * q + i], 0.0f)));
    }
}

// z2 tile ot (outputs 16 ot + 4 q + i of candidate `row`), before bias.
__device__ __forceinline__ f32x4 layer2_tile(const MlpDev& W, const bf16x4 a1[16], int ot, int row, int q) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* wr = W.w2t + (16 * ot + row) * kHidden + 4 * q;
#pragma unroll
    for (int kt = 0; kt < 16; kt++) acc = mfma16(ld4(wr + 16 * kt), a1[kt], acc);
    return acc;
}

// Scores of the 16 candidates whose bf16 inputs are rows [16][80] at x (global
// or LDS memory, 8-byte aligned): lanes 16 q + r all return the score of candidate r.
__device__ __forceinline__ float mlp_group_scores(const MlpDev& W, const uint16_t* x) {
    const int l = threadIdx.x % kWave, row = l & 15, q = l >> 4;
    bf16x4 xb[5];
    load_inputs(W, x + row * kFeat, q, xb);
    bf16x4 a1[16];
    layer1(W, xb, row, q, a1);
    float yp = 0.0f;
#pragma unroll 2
    for (int ot = 0; ot < 16; ot++) {
        const f32x4 acc = layer2_tile(W, a1, ot, row, q);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = 16 * ot + 4 * q + i;
            yp += W.w3[o] * fmaxf(acc[i] + W.b2[o], 0.0f);
        }
    }
    yp += __shfl_xor(yp, 16);
    yp += __shfl_xor(yp, 32);
    return yp + W.b3[0];
}

// The same scores with the hidden activations staged through LDS instead of
// registers (for kernels whose other phases need the registers): x holds the
// 16 bf16 input rows, the hidden layer is [16][kA1Stride] bf16 with rows below
// `split` at a1 and the others at a1b (all LDS, 8-byte aligned, private to the
// wave). The rows at a1 may overlay x: x is read into registers first.

__device__ __forceinline__ float mlp_group_scores_lds(const MlpDev& W, const uint16_t* x, uint16_t* a1,
                                                      uint16_t* a1b, int split) {
    const int l = threadIdx.x % kWave, row = l & 15, q = l >> 4;
    bf16x4 xb[5];
    load_inputs(W, x + row * kFeat, q, xb);
    uint16_t* a1row = row < split ? a1 + row * kA1Stride : a1b + (row - split) * kA1Stride;
    for (int mt = 0; mt < 16; mt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* wr = W.w1t + (16 * mt + row) * kFeat + 4 * q;
#pragma unroll
        for (int ks = 0; ks < 5; ks++) acc = mfma16(ld4(wr + 16 * ks), xb[ks], acc);
        bf16x4 h;
#pragma unroll
        for (int i = 0; i < 4; i++) h[i] = short(f2bf(fmaxf(acc[i] + W.b1[16 * mt + 4 * q + i], 0.0f)));
        *reinterpret_cast<bf16x4*>(a1row + 16 * mt + 4 * q) = h;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float yp = 0.0f;
    for (int ot = 0; ot < 16; ot++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* wr = W.w2t + (16 * ot + row) * kHidden + 4 * q;
        const uint16_t* ar = a1row + 4 * q;
#pragma unroll 4
        for (int kt = 0; kt < 16; kt++) acc = mfma16(ld4(wr + 16 * kt), ld4(ar + 16 * kt), acc);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = 16 * ot + 4 * q + i;
            yp += W.w3[o] * fmaxf(acc[i] + W.b2[o], 0.0f);
        }
    }
    yp += __shfl_xor(yp, 16);
    yp += __shfl_xor(yp, 32);
    return yp + W.b3[0];
}

}  // namespace ud2
