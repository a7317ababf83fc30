/*
 * ise_oracle_mlp.c -- CPU oracle of the learned row-order scorer.
 *
 * TEST INFRASTRUCTURE ONLY (see ise_oracle.h).  The reference has no learned
 * component; what is restated here is the project's own definition of it, so
 * that the GPU kernels can be checked:
 *   - the 16 pair features of (position i, row j) under the identity prefix,
 *     built on the reference's orbit-by-orbit row comparison
 *     (CCanonicityChecker::checkColOrbit, Source/CanonicityChecker.cpp:191-223,
 *     with the identity column permutation);
 *   - the 16 -> 128 -> 128 -> 1 ReLU network, its weighted binary
 *     cross-entropy and the gradient of that loss, either exactly (double, no
 *     rounding; pinned against central finite differences in tests/) or with
 *     the bf16 rounding points of the GPU path (weights, inputs, hidden
 *     activations and the back-propagated pre-activation gradients rounded to
 *     bf16; sums in double).
 * Parameter packing: W1[16][128], b1[128], W2[128][128], b2[128], W3[128][16]
 * (column 0 used), b3[16] (entry 0 used).
 */
#include "ise_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

#define IN 16
#define HID 128
#define OPAD 16
#define P_W1 0
#define P_B1 (P_W1 + IN * HID)
#define P_W2 (P_B1 + HID)
#define P_B2 (P_W2 + HID * HID)
#define P_W3 (P_B2 + HID)
#define P_B3 (P_W3 + HID * OPAD)
#define P_ALL (P_B3 + OPAD)

static int popc(uint64_t x) { return __builtin_popcountll(x); }
static uint64_t low(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

/* round to nearest even bf16, returned as float */
static float bf(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { u = (u & 0xffff0000u) | ((u & 0xffffu) ? 0x400000u : 0u); }
    else { u += 0x7fffu + ((u >> 16) & 1u); u &= 0xffff0000u; }
    float y;
    memcpy(&y, &u, 4);
    return y;
}

int orc_mlp_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                     uint32_t od, float *X)
{
    memset(X, 0, sizeof(float) * (size_t)n * od * od * IN);
    for (uint32_t c = 0; c < n; c++) {
        const int d = meta[c] & 0xff, b = (meta[c] >> 8) & 0xff;
        if (d < 1 || d > (int)od || d > ORC_MAX_ROWS || d > (int)stride || b < 1 || b > 64) continue;
        const uint64_t *R = rows + (size_t)c * stride;
        uint64_t cells[ORC_MAX_ROWS];
        orc_cell_masks(R, d, b, cells);
        for (int i = 0; i < d; i++) {
            const uint64_t Ci = (cells[i] | 1ull) & low(b), Ri = R[i] & low(b);
            for (int j = 0; j < d; j++) {
                const uint64_t Rj = R[j] & low(b), Rn = R[d - 1] & low(b);
                int out = 0, dec = b, e = b;
                /* orbits of position i in column order; first differing count decides */
                for (int s = 0; s < b;) {
                    const uint64_t rest = Ci & ~low(s + 1);
                    const int t = rest ? __builtin_ctzll(rest) : b;
                    const uint64_t om = low(t) & ~low(s);
                    const int cR = popc(Ri & om), cP = popc(Rj & om);
                    if (cR != cP) { out = cR > cP ? 1 : -1; dec = s; e = t; break; }
                    s = t;
                }
                const uint64_t om = out ? (low(e) & ~low(dec)) : 0ull;
                const uint64_t heads = Ci & ~1ull;
                const int first_end = heads ? __builtin_ctzll(heads) : b;
                float *f = X + (((size_t)c * od + i) * od + j) * IN;
                f[0] = out < 0 ? 1.f : (out > 0 ? -1.f : 0.f);
                f[1] = (float)dec / 64.f;
                f[2] = (float)(popc(Rj & om) - popc(Ri & om)) * 0.125f;
                f[3] = (float)i / 32.f;
                f[4] = (float)j / 32.f;
                f[5] = (float)(j - i) / 32.f;
                f[6] = (float)d / 32.f;
                f[7] = j == i ? 1.f : 0.f;
                f[8] = j == d - 1 ? 1.f : 0.f;
                f[9] = (float)popc(Rj & Ri) * 0.0625f;
                f[10] = (float)popc(Ci) / 64.f;
                f[11] = (float)popc(Rj & low(first_end)) * 0.0625f;
                f[12] = j < i ? 1.f : 0.f;
                f[13] = (float)popc(Rj & Rn) * 0.0625f;
                f[14] = (float)popc(om) / 64.f;
                f[15] = 1.f;
            }
        }
    }
    return 0;
}

typedef struct {
    double *h1, *h2, *z;     /* post-activation values as the next layer sees them */
} acts_t;

static double rnd(double x, int q) { return q ? (double)bf((float)x) : x; }

/* forward of M rows; q = emulate the bf16 rounding points */
static void fwd(const float *p, const float *X, uint32_t M, int q, acts_t *a)
{
    for (uint32_t m = 0; m < M; m++) {
        const float *x = X + (size_t)m * IN;
        double *h1 = a->h1 + (size_t)m * HID, *h2 = a->h2 + (size_t)m * HID;
        for (int u = 0; u < HID; u++) {
            double s = 0;
            for (int k = 0; k < IN; k++) s += rnd(x[k], q) * rnd(p[P_W1 + k * HID + u], q);
            s += p[P_B1 + u];
            h1[u] = rnd(s > 0 ? s : 0, q);
        }
        for (int u = 0; u < HID; u++) {
            double s = 0;
            for (int k = 0; k < HID; k++) s += h1[k] * rnd(p[P_W2 + k * HID + u], q);
            s += p[P_B2 + u];
            h2[u] = rnd(s > 0 ? s : 0, q);
        }
        double s = 0;
        for (int k = 0; k < HID; k++) s += h2[k] * rnd(p[P_W3 + k * OPAD], q);
        a->z[m] = s + p[P_B3];
    }
}

int orc_mlp_forward(const float *params, const float *X, uint32_t M, int bf16_rounding, double *z)
{
    acts_t a;
    a.h1 = (double *)malloc(sizeof(double) * M * HID);
    a.h2 = (double *)malloc(sizeof(double) * M * HID);
    a.z = z;
    fwd(params, X, M, bf16_rounding, &a);
    free(a.h1); free(a.h2);
    return 0;
}

/* weighted mean BCE on logits and its gradient (packed like the parameters) */
int orc_mlp_gradient(const float *p, const float *X, const float *y, const float *w, uint32_t M,
                     int q, double *grad, double *loss)
{
    acts_t a;
    a.h1 = (double *)malloc(sizeof(double) * M * HID);
    a.h2 = (double *)malloc(sizeof(double) * M * HID);
    a.z = (double *)malloc(sizeof(double) * M);
    double *dz2 = (double *)malloc(sizeof(double) * HID);
    double *dz1 = (double *)malloc(sizeof(double) * HID);
    fwd(p, X, M, q, &a);
    double ws = 0, L = 0;
    for (uint32_t m = 0; m < M; m++) ws += w[m];
    if (!(ws > 0)) { free(a.h1); free(a.h2); free(a.z); free(dz2); free(dz1); return -1; }
    memset(grad, 0, sizeof(double) * P_ALL);
    for (uint32_t m = 0; m < M; m++) {
        const double z = a.z[m];
        const double sg = 1.0 / (1.0 + exp(-z));
        L += w[m] * (log1p(exp(-fabs(z))) + (z > 0 ? z : 0) - z * y[m]) / ws;
        const double g3 = rnd(w[m] * (sg - y[m]) / ws, q);
        const double *h1 = a.h1 + (size_t)m * HID, *h2 = a.h2 + (size_t)m * HID;
        const float *x = X + (size_t)m * IN;
        grad[P_B3] += g3;
        for (int k = 0; k < HID; k++) {
            grad[P_W3 + k * OPAD] += h2[k] * g3;
            const double dh = g3 * rnd(p[P_W3 + k * OPAD], q);
            dz2[k] = h2[k] > 0 ? rnd(dh, q) : 0.0;
        }
        for (int u = 0; u < HID; u++) {
            grad[P_B2 + u] += dz2[u];
            for (int k = 0; k < HID; k++) grad[P_W2 + k * HID + u] += h1[k] * dz2[u];
        }
        for (int k = 0; k < HID; k++) {
            double dh = 0;
            for (int u = 0; u < HID; u++) dh += dz2[u] * rnd(p[P_W2 + k * HID + u], q);
            dz1[k] = h1[k] > 0 ? rnd(dh, q) : 0.0;
        }
        for (int u = 0; u < HID; u++) {
            grad[P_B1 + u] += dz1[u];
            for (int k = 0; k < IN; k++) grad[P_W1 + k * HID + u] += rnd(x[k], q) * dz1[u];
        }
    }
    if (loss) *loss = L;
    free(a.h1); free(a.h2); free(a.z); free(dz2); free(dz1);
    return 0;
}
