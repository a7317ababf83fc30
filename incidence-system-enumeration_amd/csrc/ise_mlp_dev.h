// ise_mlp_dev.h -- device pieces of the row-order network shared by the
// scoring kernels (ise_mlp.hip) and the learned-order canonicity kernel
// (ise_kernels.hip): bf16 conversion, the weights as MFMA A fragments and the
// forward pass of one tile of 16 feature rows.
//
// v_mfma_f32_16x16x16_bf16 operand fragments, per lane l of the wave:
//   A (16 x 16, row m, column k):  m = l % 16, k = 4*(l/16) + 0..3
//   B (16 x 16, row k, column n):  n = l % 16, k = 4*(l/16) + 0..3
//   C (16 x 16):                   n = l % 16, m = 4*(l/16) + 0..3
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise_mlp.h"

namespace ise {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
    // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// The products are computed transposed, H^T = W^T X^T, with the weights as
// the A operand: the C fragment of one layer (n = feature row = l % 16, m =
// unit 4*(l/16) + t) is then exactly the B fragment of the next (k = unit,
// n = row), so activations stay in registers from the features to the score.
// Weights are pre-arranged as A fragments (frag_pack): tile (mt, kt) of a
// layer holds, for lane l, W[kt*16 + 4*(l/16) + t][mt*16 + l % 16], t = 0..3.
constexpr int FRAG_L1 = 0, FRAG_L2 = 8, FRAG_L3 = 8 + 64, FRAG_TILES = 8 + 64 + 8;

__device__ __forceinline__ bf16x4 relu_bias_bf16(f32x4 acc, const float *bias) {
    bf16x4 h;
    for (int t = 0; t < 4; t++) {
        const float v = acc[t] + bias[t];
        h[t] = (short)f2bf(v > 0.f ? v : 0.f);
    }
    return h;
}

// Scores of one tile of 16 pair rows given as a B fragment; lane l < 16
// returns the score of row l of the tile.
__device__ __forceinline__ float mlp_tile(bf16x4 xb, const bf16x4 *__restrict__ frag, const float *__restrict__ w,
                                          int l) {
    const int q = l >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4 h1[8], h2[8];
    for (int mt = 0; mt < 8; mt++) {
        const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L1 + mt) * 64 + l], xb, zero, 0, 0, 0);
        h1[mt] = relu_bias_bf16(acc, w + MLP_B1 + mt * 16 + 4 * q);
    }
    for (int mt = 0; mt < 8; mt++) {
        f32x4 acc = zero;
        for (int kt = 0; kt < 8; kt++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L2 + mt * 8 + kt) * 64 + l], h1[kt], acc, 0, 0,
                                                            0);
        h2[mt] = relu_bias_bf16(acc, w + MLP_B2 + mt * 16 + 4 * q);
    }
    f32x4 acc = zero;
    for (int kt = 0; kt < 8; kt++)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L3 + kt) * 64 + l], h2[kt], acc, 0, 0, 0);
    return acc[0] + w[MLP_B3];
}

// The same forward pass with the fragments read from global memory, eight
// per step: each step addresses its fragments through an opaque zero, so the
// compiler cannot load them ahead of the step and at most eight are in
// registers at once (the kernels that call it keep their LDS for the search
// and their registers for its state).
__device__ __forceinline__ const bf16x4 *opaque_step(const bf16x4 *p) {
    int zero;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
    return p + zero;
}

__device__ __forceinline__ float mlp_tile_streamed(bf16x4 xb, const bf16x4 *frag_g, const float *w, int l) {
    const int q = l >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4 h1[8];
    const bf16x4 *frag = opaque_step(frag_g);
    for (int mt = 0; mt < 8; mt++) {
        const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L1 + mt) * 64 + l], xb, zero, 0, 0, 0);
        h1[mt] = relu_bias_bf16(acc, w + MLP_B1 + mt * 16 + 4 * q);
    }
    // each tile of the second layer's output goes into the third layer at once,
    // so only one of them is live
    f32x4 out = zero;
    for (int mt = 0; mt < 8; mt++) {
        frag = opaque_step(frag_g);
        f32x4 acc = zero;
        for (int kt = 0; kt < 8; kt++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L2 + mt * 8 + kt) * 64 + l], h1[kt], acc, 0, 0,
                                                            0);
        const bf16x4 h2 = relu_bias_bf16(acc, w + MLP_B2 + mt * 16 + 4 * q);
        out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag[(FRAG_L3 + mt) * 64 + l], h2, out, 0, 0, 0);
    }
    return out[0] + w[MLP_B3];
}

} // namespace ise
