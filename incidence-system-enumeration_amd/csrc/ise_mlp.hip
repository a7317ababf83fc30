// ise_mlp.hip -- the learned row-order scorer: pair features, a 3-layer MLP
// (16 -> 128 -> 128 -> 1, ReLU) with bf16 weights and FP32 accumulation on
// MFMA, its backward pass, Adam, and the conversion of scores to per-
// candidate row orders for the witness hunt.
//
// Matrix products use v_mfma_f32_16x16x16_bf16 (one wave per 16x16 output
// tile, K stepped by 16).  Operand fragments, per lane l of the wave:
//   A (16 x 16, row m, column k):  m = l % 16, k = 4*(l/16) + 0..3
//   B (16 x 16, row k, column n):  n = l % 16, k = 4*(l/16) + 0..3
//   C (16 x 16):                   n = l % 16, m = 4*(l/16) + 0..3
// Operands are addressed through row/column strides so that one kernel
// serves X*W, H^T*dZ (weight gradients) and dZ*W^T (input gradients).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise_mlp.h"
#include "ise_wave.h"
#include "ise_mlp_dev.h"
#include "ise_lead.h"

namespace ise {

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]) (relu), for a 16-aligned M, N, K.
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; outputs optional.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const uint16_t *__restrict__ A,
                                                        long sam, long sak, const uint16_t *__restrict__ B,
                                                        long sbk, long sbn, const float *__restrict__ bias,
                                                        int relu, float *__restrict__ Cf, uint16_t *__restrict__ Cb,
                                                        long ldc, int accumulate) {
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int tiles_n = N / 16;
    const int tm = wave / tiles_n, tn = wave % tiles_n;
    if (tm * 16 >= M) return;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, q = l >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
        for (int t = 0; t < 4; t++) {
            const int k = k0 + 4 * q + t;
            a[t] = (short)A[(long)(tm * 16 + r16) * sam + (long)k * sak];
            b[t] = (short)B[(long)k * sbk + (long)(tn * 16 + r16) * sbn];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = tn * 16 + r16;
    for (int t = 0; t < 4; t++) {
        const int m = tm * 16 + 4 * q + t;
        float v = acc[t];
        if (accumulate && Cf) v += Cf[(long)m * ldc + n];
        if (bias) v += bias[n];
        if (relu) v = v > 0.f ? v : 0.f;
        if (Cf) Cf[(long)m * ldc + n] = v;
        if (Cb) Cb[(long)m * ldc + n] = f2bf(v);
    }
}

hipError_t gemm_bf16(int M, int N, int K, const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk,
                     long sbn, const float *bias, int relu, float *Cf, uint16_t *Cb, long ldc, int accumulate,
                     hipStream_t s) {
    if (M % 16 || N % 16 || K % 16 || M <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
    const long waves = (long)(M / 16) * (N / 16);
    const long blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                       bias, relu, Cf, Cb, ldc, accumulate);
    return hipGetLastError();
}

// ---- loss, masks, reductions, optimiser ------------------------------------

// Binary cross-entropy on logits z[m] (column 0 of a [M x 16] block) against
// labels y[m] with weights w[m]; writes dz (bf16, [M x 16], only column 0
// non-zero) = w (sigmoid(z) - y) / sum(w), and the weighted loss terms.
__global__ void bce_kernel(int M, const float *__restrict__ z, long ldz, const float *__restrict__ y,
                           const float *__restrict__ w, float inv_wsum, uint16_t *__restrict__ dz,
                           float *__restrict__ loss_terms) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float x = z[(long)m * ldz];
    const float p = 1.f / (1.f + __expf(-x));
    const float g = w[m] * (p - y[m]) * inv_wsum;
    // log(1 + e^-|x|) + max(x, 0) - x*y
    const float l = __logf(1.f + __expf(-fabsf(x))) + fmaxf(x, 0.f) - x * y[m];
    loss_terms[m] = w[m] * l * inv_wsum;
    uint16_t *row = dz + (long)m * 16;
    row[0] = f2bf(g);
    for (int c = 1; c < 16; c++) row[c] = 0;
}

// dz = dh * (z > 0) as bf16 (z: pre-activation, fp32)
__global__ void relu_back_kernel(long count, const float *__restrict__ dh, const float *__restrict__ z,
                                 uint16_t *__restrict__ dz) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dz[i] = f2bf(z[i] > 0.f ? dh[i] : 0.f);
}

// out[n] = sum_m bf16 x[m][n] (column sums, one block per column)
__global__ void colsum_kernel(int M, int N, const uint16_t *__restrict__ x, float *__restrict__ out) {
    __shared__ float part[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += bf2f(x[(long)m * N + n]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = part[0];
}

// sum of `count` floats into out[0] (single block)
__global__ void sum_kernel(long count, const float *__restrict__ x, float *__restrict__ out) {
    __shared__ float part[256];
    float s = 0.f;
    for (long i = threadIdx.x; i < count; i += blockDim.x) s += x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

// Adam on fp32 master weights; the bf16 copy used by the products follows.
__global__ void adam_kernel(long count, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, uint16_t *__restrict__ wb, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[i] = v;
    if (wb) wb[i] = f2bf(v);
}

__global__ void to_bf16_kernel(long count, const float *__restrict__ x, uint16_t *__restrict__ y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) y[i] = f2bf(x[i]);
}

static unsigned blocks_for(long count, int threads) { return (unsigned)((count + threads - 1) / threads); }

hipError_t bce(int M, const float *z, long ldz, const float *y, const float *w, float inv_wsum, uint16_t *dz,
               float *loss_terms, hipStream_t s) {
    hipLaunchKernelGGL(bce_kernel, dim3(blocks_for(M, 256)), dim3(256), 0, s, M, z, ldz, y, w, inv_wsum, dz,
                       loss_terms);
    return hipGetLastError();
}
hipError_t relu_back(long count, const float *dh, const float *z, uint16_t *dz, hipStream_t s) {
    hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, count, dh, z, dz);
    return hipGetLastError();
}
hipError_t colsum(int M, int N, const uint16_t *x, float *out, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3(N), dim3(256), 0, s, M, N, x, out);
    return hipGetLastError();
}
hipError_t sum(long count, const float *x, float *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, count, x, out);
    return hipGetLastError();
}
hipError_t adam(long count, float *w, const float *g, float *m1, float *m2, uint16_t *wb, float lr, float b1,
                float b2, float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, count, w, g, m1, m2, wb, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t to_bf16(long count, const float *x, uint16_t *y, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, count, x, y);
    return hipGetLastError();
}

// ---- pair features -------------------------------------------------------
//
// For candidate c, position i and row j (i, j < d), 16 features describing
// row j placed at position i under the identity prefix: the orbit-by-orbit
// comparison of row j with row i (the same rule as the search), row
// intersections, column-orbit sizes, positions and depth.  All values are
// small multiples of powers of two, exact in bf16.

// Wave-uniform statistics of one (i, j) pair.
struct PairStats {
    int out;        // compare_row outcome: -1 row j larger at position i, 1 smaller, 0 equal
    int dec;        // deciding orbit's first column (b when equal)
    int dsize;      // deciding orbit's size (0 when equal)
    int ddiff;      // ones of row j minus ones of row i inside it
    int inter_i;    // |R_j & R_i|
    int inter_n;    // |R_j & R_{d-1}|
    int first;      // ones of row j in the first column orbit of position i
    int norb;       // column orbits at position i
};

ISE_DEV PairStats pair_stats(const Grp<64> &g, const Staged &c, const uint32_t *geo, int i, int j) {
    PairStats s;
    const int b = c.b, d = c.d;
    int pc = g.ln, dec = b;
    uint32_t dummy0 = 0, dummy1 = 0;
    s.out = compare_row<64, false>(g, c, geo, i, j, pc, &dec, dummy0, dummy1);
    const uint64_t Ri = g.read64(c.rowv, i), Ci = g.read64(c.cellv, i);
    const uint64_t Rj = g.read64(c.rowv, j), Rn = g.read64(c.rowv, d - 1);
    s.dec = s.out ? dec : b;
    const uint64_t after = Ci & ~lowmask(s.dec + 1);
    const int e = after ? ctz64(after) : b;
    const uint64_t om = s.out ? (lowmask(e) & ~lowmask(s.dec)) : 0ull;
    s.dsize = popc64(om);
    s.ddiff = popc64(Rj & om) - popc64(Ri & om);
    s.inter_i = popc64(Rj & Ri);
    s.inter_n = popc64(Rj & Rn);
    const uint64_t heads = Ci & ~1ull;
    s.first = popc64(Rj & lowmask(heads ? ctz64(heads) : b));
    s.norb = popc64(Ci);
    return s;
}

ISE_DEV float pair_feature(int id, const PairStats &s, int i, int j, int d) {
    switch (id) {
    case 0: return s.out < 0 ? 1.f : (s.out > 0 ? -1.f : 0.f);
    case 1: return (float)s.dec * (1.f / 64.f);
    case 2: return (float)s.ddiff * 0.125f;
    case 3: return (float)i * (1.f / 32.f);
    case 4: return (float)j * (1.f / 32.f);
    case 5: return (float)(j - i) * (1.f / 32.f);
    case 6: return (float)d * (1.f / 32.f);
    case 7: return j == i ? 1.f : 0.f;
    case 8: return j == d - 1 ? 1.f : 0.f;
    case 9: return (float)s.inter_i * 0.0625f;
    case 10: return (float)s.norb * (1.f / 64.f);
    case 11: return (float)s.first * 0.0625f;
    case 12: return j < i ? 1.f : 0.f;
    case 13: return (float)s.inter_n * 0.0625f;
    case 14: return (float)s.dsize * (1.f / 64.f);
    default: return 1.f;
    }
}

// Features of every pair to X[(c*od + i)*od + j][16] (pairs with i or j >= d
// are zero); one wave per candidate.
__global__ __launch_bounds__(256) void features_kernel(uint32_t n, uint32_t stride, const uint64_t *__restrict__ rows,
                                                       const uint32_t *__restrict__ meta, uint32_t od,
                                                       uint16_t *__restrict__ X) {
    extern __shared__ uint32_t lds_geo[];
    Grp<64> g;
    uint32_t *geo = lds_geo + (threadIdx.x >> 6) * ISE_MAX_ROWS * 64;
    const uint32_t ci = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (ci >= n) return;
    const uint32_t m = meta[ci];
    const int d = m & 0xff, b = (m >> 8) & 0xff;
    const int l = g.ln;
    uint16_t *xc = X + (uint64_t)ci * od * od * 16;
    const bool ok = d >= 1 && d <= (int)od && d <= ISE_MAX_ROWS && (uint32_t)d <= stride && b >= 1 && b <= 64;
    if (!ok) {
        for (uint32_t t = l; t < od * od * 16; t += 64) xc[t] = 0;
        return;
    }
    const Staged c = stage(g, ci, m, rows, nullptr, stride, geo);
    for (int i = 0; i < (int)od; i++) {
        for (int j = 0; j < (int)od; j++) {
            float f = 0.f;
            if (i < d && j < d) {
                const PairStats s = pair_stats(g, c, geo, i, j);
                f = pair_feature(l & 15, s, i, j, d);
            }
            if (l < 16) xc[((uint64_t)i * od + j) * 16 + l] = f2bf(f);
        }
    }
}

hipError_t features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, uint32_t od,
                    uint16_t *X, hipStream_t s) {
    const uint32_t wpb = 4;
    const size_t lds = (size_t)wpb * ISE_MAX_ROWS * 64 * sizeof(uint32_t);
    hipLaunchKernelGGL(features_kernel, dim3((n + wpb - 1) / wpb), dim3(64 * wpb), lds, s, n, stride, rows, meta, od,
                       X);
    return hipGetLastError();
}

// Features of every row as the row of position 0 (the lead order's network
// input): X[(c*16 + j)*16 + f], zero past the depth and for candidates the
// lead order does not take (depth outside 2..16, more than 64 columns).
__global__ __launch_bounds__(256) void lead_features_kernel(uint32_t n, uint32_t stride,
                                                            const uint64_t *__restrict__ rows,
                                                            const uint32_t *__restrict__ meta,
                                                            uint16_t *__restrict__ X) {
    Grp<64> g;
    const uint32_t ci = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (ci >= n) return;
    const uint32_t m = meta[ci];
    const int d = m & 0xff, b = (m >> 8) & 0xff;
    const int l = g.ln;
    uint16_t *xc = X + (uint64_t)ci * LEAD_MAX_ROWS * MLP_IN;
    if (d < 2 || d > LEAD_MAX_ROWS || (uint32_t)d > stride || b < 1 || b > 64) {
        for (int t = l; t < LEAD_MAX_ROWS * MLP_IN; t += 64) xc[t] = 0;
        return;
    }
    const Staged c = load_candidate(g, ci, m, rows, nullptr, stride);
    const LeadStats s = lead_stats(g, c);
    if (l < LEAD_MAX_ROWS)
        for (int f = 0; f < MLP_IN; f++) xc[l * MLP_IN + f] = f2bf(lead_feature(f, s, l, d));
}

hipError_t lead_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, uint16_t *X,
                         hipStream_t s) {
    const uint32_t wpb = 4;
    hipLaunchKernelGGL(lead_features_kernel, dim3((n + wpb - 1) / wpb), dim3(64 * wpb), 0, s, n, stride, rows, meta,
                       X);
    return hipGetLastError();
}

// ---- fused scoring: features -> MLP -> per-position row order -------------
//
// The products are computed transposed, H^T = W^T X^T, with the weights as
// the A operand: the C fragment of one layer (n = pair row = l % 16, m = unit
// 4*(l/16) + t) is then exactly the B fragment of the next (k = unit, n =
// row), so activations stay in registers from the features to the score.
// The weights are pre-arranged as A fragments (frag_pack_kernel) and staged
// in LDS once per workgroup: tile (mt, kt) of a layer holds, for lane l,
// W[kt*16 + 4*(l/16) + t][mt*16 + l % 16], t = 0..3.

__global__ void frag_pack_kernel(const uint16_t *__restrict__ wb, uint16_t *__restrict__ frag) {
    const int tile = blockIdx.x, l = threadIdx.x;   // 80 blocks of 64 lanes
    const int q = l >> 4, r = l & 15;
    for (int t = 0; t < 4; t++) {
        uint16_t v;
        if (tile < FRAG_L2) {                         // W1 [16][128]: mt = tile, kt = 0
            v = wb[MLP_W1 + (4 * q + t) * MLP_HID + tile * 16 + r];
        } else if (tile < FRAG_L3) {                  // W2 [128][128]
            const int mt = (tile - FRAG_L2) / 8, kt = (tile - FRAG_L2) % 8;
            v = wb[MLP_W2 + (kt * 16 + 4 * q + t) * MLP_HID + mt * 16 + r];
        } else {                                      // W3 [128][16]: mt = 0
            const int kt = tile - FRAG_L3;
            v = wb[MLP_W3 + (kt * 16 + 4 * q + t) * MLP_OUT_PAD + r];
        }
        frag[((size_t)tile * 64 + l) * 4 + t] = v;
    }
}

hipError_t frag_pack(const uint16_t *wb, uint16_t *frag, hipStream_t s) {
    hipLaunchKernelGGL(frag_pack_kernel, dim3(FRAG_TILES), dim3(64), 0, s, wb, frag);
    return hipGetLastError();
}

// One wave per candidate (grid-stride).  For every position i < d: features
// of the pairs (i, j), j < d, in tiles of 16, the MLP, then the rows j < d
// sorted by decreasing score (ties by increasing j) into order[c][i][0..d-1];
// entries j >= d, and positions i >= d, hold the identity.  Optionally the
// scores (score[c][i][j], 0 outside d x d).
__global__ __launch_bounds__(256) void learned_order_kernel(uint32_t n, uint32_t stride,
                                                            const uint64_t *__restrict__ rows,
                                                            const uint32_t *__restrict__ meta, uint32_t od,
                                                            const uint16_t *__restrict__ frag_g,
                                                            const float *__restrict__ w, uint8_t *__restrict__ order,
                                                            float *__restrict__ score,
                                                            const uint32_t *__restrict__ list,
                                                            const uint32_t *__restrict__ list_count) {
    extern __shared__ uint32_t lds[];
    bf16x4 *frag = (bf16x4 *)lds;                                       // FRAG_TILES * 64 * 8 bytes
    uint32_t *geo = lds + FRAG_TILES * 64 * 2 + (threadIdx.x >> 6) * ISE_MAX_ROWS * 64;
    {
        const uint2 *src = (const uint2 *)frag_g;
        uint2 *dst = (uint2 *)frag;
        for (int t = threadIdx.x; t < FRAG_TILES * 64; t += blockDim.x) dst[t] = src[t];
    }
    __syncthreads();
    Grp<64> g;
    const int l = g.ln, q = l >> 4, r16 = l & 15;
    const uint32_t wpb = blockDim.x >> 6;
    // every candidate, or only those of `list` (list_count of them)
    const uint32_t total = list ? *list_count : n;
    for (uint32_t k = blockIdx.x * wpb + (threadIdx.x >> 6); k < total; k += gridDim.x * wpb) {
        const uint32_t ci = list ? list[k] : k;
        if (ci >= n) continue;
        const uint32_t m = meta[ci];
        const int d = m & 0xff, b = (m >> 8) & 0xff;
        uint8_t *oc = order + (uint64_t)ci * od * od;
        float *sc = score ? score + (uint64_t)ci * od * od : nullptr;
        const bool ok = d >= 1 && d <= (int)od && d <= ISE_MAX_ROWS && (uint32_t)d <= stride && b >= 1 && b <= 64;
        // identity outside the d x d block (written once; the block is written below)
        const uint32_t dd = ok ? (uint32_t)d : 0u;
        for (uint32_t t = l; t < od * od; t += 64) {
            if (t / od < dd && t % od < dd) continue;
            oc[t] = (uint8_t)(t % od);
            if (sc) sc[t] = 0.f;
        }
        if (!ok) continue;
        const Staged c = stage(g, ci, m, rows, nullptr, stride, geo);
        for (int i = 0; i < d; i++) {
            float mine = 0.f;   // lane j < d: score of row j at position i
            for (int jt = 0; jt * 16 < d; jt++) {
                bf16x4 xb = {0, 0, 0, 0};
                for (int jj = 0; jj < 16; jj++) {
                    const int j = jt * 16 + jj;
                    if (j >= d) break;
                    const PairStats s = pair_stats(g, c, geo, i, j);
                    if (r16 == jj)
                        for (int t = 0; t < 4; t++) xb[t] = (short)f2bf(pair_feature(4 * q + t, s, i, j, d));
                }
                // an opaque zero offset keeps the weight fragments in LDS
                // (loaded per tile) instead of hoisted into registers
                int zero;
                asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
                const float s16 = mlp_tile(xb, frag + zero, w, l);
                // move the tile's scores (lanes 0..15) to lanes jt*16 .. jt*16+15
                const float moved = __shfl(s16, (l - jt * 16) & 63);
                if (l >= jt * 16 && l < jt * 16 + 16) mine = moved;
            }
            // stable descending rank among rows 0..d-1
            int rank = 0;
            for (int k = 0; k < d; k++) {
                const float sk = __shfl(mine, k);
                rank += (sk > mine) || (sk == mine && k < l);
            }
            if (l < d) {
                oc[i * od + rank] = (uint8_t)l;
                if (sc) sc[i * od + l] = mine;
            }
        }
    }
}

size_t learned_order_lds(uint32_t wpb) {
    return (size_t)FRAG_TILES * 64 * 8 + (size_t)wpb * ISE_MAX_ROWS * 64 * sizeof(uint32_t);
}

hipError_t learned_order(uint32_t blocks, uint32_t wpb, uint32_t n, uint32_t stride, const uint64_t *rows,
                         const uint32_t *meta, uint32_t od, const uint16_t *frag, const float *w, uint8_t *order,
                         float *score, hipStream_t s, const uint32_t *list, const uint32_t *list_count) {
    if (wpb < 1 || wpb > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(learned_order_kernel, dim3(blocks), dim3(64 * wpb), learned_order_lds(wpb), s, n, stride, rows,
                       meta, od, frag, w, order, score, list, list_count);
    return hipGetLastError();
}

hipError_t learned_order_occupancy(uint32_t wpb, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)&learned_order_kernel,
                                                        (int)(64 * wpb), learned_order_lds(wpb));
}

// Scores [n][od][od] -> order [n][od][od]: row i of a candidate lists the rows
// j < d by decreasing score (ties by increasing j); entries past d are j.
__global__ void order_kernel(uint32_t n, const uint32_t *__restrict__ meta, uint32_t od, const float *__restrict__ score,
                             long score_ld, uint8_t *__restrict__ order) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long c = t / od, i = t % od;
    if (c >= n) return;
    const int d = meta[c] & 0xff;
    const float *sc = score + ((c * od + i) * od) * score_ld;
    uint8_t *out = order + (c * od + i) * od;
    for (int j = 0; j < (int)od; j++) {
        if (j >= d) { out[j] = (uint8_t)j; continue; }
        const float sj = sc[(long)j * score_ld];
        int rank = 0;
        for (int k = 0; k < d; k++) {
            const float sk = sc[(long)k * score_ld];
            rank += (sk > sj) || (sk == sj && k < j);
        }
        out[rank] = (uint8_t)j;
    }
}

hipError_t order_from_scores(uint32_t n, const uint32_t *meta, uint32_t od, const float *score, long score_ld,
                             uint8_t *order, hipStream_t s) {
    const long threads = (long)n * od;
    hipLaunchKernelGGL(order_kernel, dim3(blocks_for(threads, 256)), dim3(256), 0, s, n, meta, od, score, score_ld,
                       order);
    return hipGetLastError();
}

} // namespace ise
