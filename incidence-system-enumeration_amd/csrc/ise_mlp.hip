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
// comparison of row j with row i (the same rule as the search), intersections
// and positions.  Written as bf16 to X[(c*od + i)*od + j][0..15]; pairs with
// i or j >= d are zero.
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
    uint32_t dummy0 = 0, dummy1 = 0;
    const float inv32 = 1.f / 32.f;
    for (int i = 0; i < (int)od; i++) {
        const uint64_t Ri = i < d ? g.read64(c.rowv, i) : 0ull;
        const uint64_t Ci = i < d ? g.read64(c.cellv, i) : 0ull;
        for (int j = 0; j < (int)od; j++) {
            float f = 0.f;
            if (i < d && j < d) {
                int pc = l, dec = b;
                const int out = compare_row<64, false>(g, c, geo, i, j, pc, &dec, dummy0, dummy1);
                const uint64_t Rj = g.read64(c.rowv, j);
                const uint64_t Rn = g.read64(c.rowv, d - 1);
                // ones of row j and row i inside the deciding orbit
                const uint64_t after = Ci & ~lowmask(dec + 1);
                const int e = after ? ctz64(after) : b;
                const uint64_t om = (dec < b) ? (lowmask(e) & ~lowmask(dec)) : 0ull;
                const uint64_t first = Ci & ~1ull;
                const int e0 = first ? ctz64(first) : b;
                switch (l) {
                case 0: f = out < 0 ? 1.f : (out > 0 ? -1.f : 0.f); break;
                case 1: f = (float)(out ? dec : b) / 64.f; break;
                case 2: f = (float)(popc64(Rj & om) - popc64(Ri & om)) * 0.125f; break;
                case 3: f = i * inv32; break;
                case 4: f = j * inv32; break;
                case 5: f = (j - i) * inv32; break;
                case 6: f = d * inv32; break;
                case 7: f = j == i ? 1.f : 0.f; break;
                case 8: f = j == d - 1 ? 1.f : 0.f; break;
                case 9: f = popc64(Rj & Ri) * 0.0625f; break;
                case 10: f = popc64(Ci) / 64.f; break;
                case 11: f = popc64(Rj & lowmask(e0)) * 0.0625f; break;
                case 12: f = j < i ? 1.f : 0.f; break;
                case 13: f = popc64(Rj & Rn) * 0.0625f; break;
                case 14: f = b / 64.f; break;
                case 15: f = 1.f; break;
                default: break;
                }
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
