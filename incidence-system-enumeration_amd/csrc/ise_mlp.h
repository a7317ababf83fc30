// ise_mlp.h -- kernels of the learned row-order scorer (ise_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ise {

constexpr int MLP_IN = 16, MLP_HID = 128, MLP_OUT_PAD = 16;

// Packed parameter layout (floats): W1[16][128], b1[128], W2[128][128], b2[128],
// W3[128][16] (column 0 used), b3[16] (entry 0 used).
constexpr long MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_IN * MLP_HID, MLP_W2 = MLP_B1 + MLP_HID,
               MLP_B2 = MLP_W2 + MLP_HID * MLP_HID, MLP_W3 = MLP_B2 + MLP_HID,
               MLP_B3 = MLP_W3 + MLP_HID * MLP_OUT_PAD, MLP_PARAMS = MLP_B3 + MLP_OUT_PAD;

hipError_t gemm_bf16(int M, int N, int K, const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk,
                     long sbn, const float *bias, int relu, float *Cf, uint16_t *Cb, long ldc, int accumulate,
                     hipStream_t s);
hipError_t bce(int M, const float *z, long ldz, const float *y, const float *w, float inv_wsum, uint16_t *dz,
               float *loss_terms, hipStream_t s);
hipError_t relu_back(long count, const float *dh, const float *z, uint16_t *dz, hipStream_t s);
hipError_t colsum(int M, int N, const uint16_t *x, float *out, hipStream_t s);
hipError_t sum(long count, const float *x, float *out, hipStream_t s);
hipError_t adam(long count, float *w, const float *g, float *m1, float *m2, uint16_t *wb, float lr, float b1,
                float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t to_bf16(long count, const float *x, uint16_t *y, hipStream_t s);
hipError_t features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, uint32_t od,
                    uint16_t *X, hipStream_t s);
// Features of every row as the row of position 0: X [n][16][16] bf16.
hipError_t lead_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, uint16_t *X,
                         hipStream_t s);
// Weights as MFMA A fragments (FRAG words of uint16) for learned_order.
constexpr long MLP_FRAG_WORDS = 80L * 64 * 4;
hipError_t frag_pack(const uint16_t *wb, uint16_t *frag, hipStream_t s);
// Fused features -> MLP -> per-position order, one wave per candidate.
// LDS of one workgroup: the weight fragments (40 KiB) plus the orbit
// geometry of each wave's candidate (8 KiB per wave).
size_t learned_order_lds(uint32_t waves_per_block);
// With `list`, only the candidates list[0 .. *list_count) are scored.
hipError_t learned_order(uint32_t blocks, uint32_t waves_per_block, uint32_t n, uint32_t stride,
                         const uint64_t *rows, const uint32_t *meta, uint32_t od, const uint16_t *frag,
                         const float *w, uint8_t *order, float *score, hipStream_t s,
                         const uint32_t *list = nullptr, const uint32_t *list_count = nullptr);
hipError_t learned_order_occupancy(uint32_t waves_per_block, int *blocks_per_cu);
hipError_t order_from_scores(uint32_t n, const uint32_t *meta, uint32_t od, const float *score, long score_ld,
                             uint8_t *order, hipStream_t s);

} // namespace ise
