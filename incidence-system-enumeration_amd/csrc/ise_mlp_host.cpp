// ise_mlp_host.cpp -- the learned row-order model: parameters, Adam state,
// activation workspace and the forward / backward / training / ordering
// sequences built from the kernels of ise_mlp.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>
#include "ise_internal.h"
#include "ise_mlp.h"

using namespace ise;

struct ise_mlp {
    ise_ctx *ctx = nullptr;
    hipStream_t s = nullptr;
    uint32_t max_rows = 0;
    int step = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // [MLP_PARAMS]
    uint16_t *wb = nullptr;                                           // bf16 copy [MLP_PARAMS]
    // workspace
    float *h1f = nullptr, *h2f = nullptr, *z3 = nullptr, *dh2 = nullptr, *dh1 = nullptr, *lossv = nullptr;
    uint16_t *h1b = nullptr, *h2b = nullptr, *dz3 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *scalar = nullptr;
    // ordering workspace
    uint16_t *X = nullptr;
    float *score = nullptr;
};

namespace {

int chk(hipError_t e) { return e == hipSuccess ? ISE_OK : ISE_ERR_HIP; }
#define CK(x)                          \
    do {                               \
        int rc_ = chk(x);              \
        if (rc_) return rc_;           \
    } while (0)

template <typename T> hipError_t dalloc(T **p, size_t count) { return hipMalloc((void **)p, count * sizeof(T) + 64); }

// He-uniform initial weights from a fixed generator (splitmix64).
void init_params(std::vector<float> &p, uint64_t seed) {
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0;
    };
    p.assign(MLP_PARAMS, 0.f);
    const double a1 = sqrt(6.0 / MLP_IN), a2 = sqrt(6.0 / MLP_HID);
    for (long i = 0; i < MLP_IN * MLP_HID; i++) p[MLP_W1 + i] = (float)((2 * next() - 1) * a1);
    for (long i = 0; i < MLP_HID * MLP_HID; i++) p[MLP_W2 + i] = (float)((2 * next() - 1) * a2);
    for (long k = 0; k < MLP_HID; k++) p[MLP_W3 + k * MLP_OUT_PAD] = (float)((2 * next() - 1) * a2);
}

// logits of M rows into m->z3 ([M][16], column 0)
int forward(ise_mlp *m, const uint16_t *X, uint32_t M) {
    const uint16_t *W1 = m->wb + MLP_W1, *W2 = m->wb + MLP_W2, *W3 = m->wb + MLP_W3;
    CK(gemm_bf16(M, MLP_HID, MLP_IN, X, MLP_IN, 1, W1, MLP_HID, 1, m->w + MLP_B1, 1, m->h1f, m->h1b, MLP_HID, 0, m->s));
    CK(gemm_bf16(M, MLP_HID, MLP_HID, m->h1b, MLP_HID, 1, W2, MLP_HID, 1, m->w + MLP_B2, 1, m->h2f, m->h2b, MLP_HID, 0,
                 m->s));
    CK(gemm_bf16(M, MLP_OUT_PAD, MLP_HID, m->h2b, MLP_HID, 1, W3, MLP_OUT_PAD, 1, m->w + MLP_B3, 0, m->z3, nullptr,
                 MLP_OUT_PAD, 0, m->s));
    return ISE_OK;
}

// gradient of the loss into m->g; loss into m->scalar[0]
int backward(ise_mlp *m, const uint16_t *X, const float *y, const float *wt, uint32_t M, float inv_wsum) {
    const uint16_t *W2 = m->wb + MLP_W2, *W3 = m->wb + MLP_W3;
    float *g = m->g;
    CK(bce((int)M, m->z3, MLP_OUT_PAD, y, wt, inv_wsum, m->dz3, m->lossv, m->s));
    CK(sum(M, m->lossv, m->scalar, m->s));
    // layer 3: dW3 = H2^T dZ3, db3 = colsum dZ3, dH2 = dZ3 W3^T
    CK(gemm_bf16(MLP_HID, MLP_OUT_PAD, M, m->h2b, 1, MLP_HID, m->dz3, MLP_OUT_PAD, 1, nullptr, 0, g + MLP_W3, nullptr,
                 MLP_OUT_PAD, 0, m->s));
    CK(colsum(M, MLP_OUT_PAD, m->dz3, g + MLP_B3, m->s));
    CK(gemm_bf16(M, MLP_HID, MLP_OUT_PAD, m->dz3, MLP_OUT_PAD, 1, W3, 1, MLP_OUT_PAD, nullptr, 0, m->dh2, nullptr,
                 MLP_HID, 0, m->s));
    CK(relu_back((long)M * MLP_HID, m->dh2, m->h2f, m->dz2, m->s));
    // layer 2
    CK(gemm_bf16(MLP_HID, MLP_HID, M, m->h1b, 1, MLP_HID, m->dz2, MLP_HID, 1, nullptr, 0, g + MLP_W2, nullptr, MLP_HID,
                 0, m->s));
    CK(colsum(M, MLP_HID, m->dz2, g + MLP_B2, m->s));
    CK(gemm_bf16(M, MLP_HID, MLP_HID, m->dz2, MLP_HID, 1, W2, 1, MLP_HID, nullptr, 0, m->dh1, nullptr, MLP_HID, 0,
                 m->s));
    CK(relu_back((long)M * MLP_HID, m->dh1, m->h1f, m->dz1, m->s));
    // layer 1
    CK(gemm_bf16(MLP_IN, MLP_HID, M, X, 1, MLP_IN, m->dz1, MLP_HID, 1, nullptr, 0, g + MLP_W1, nullptr, MLP_HID, 0,
                 m->s));
    CK(colsum(M, MLP_HID, m->dz1, g + MLP_B1, m->s));
    return ISE_OK;
}

float weight_sum(ise_mlp *m, const float *wt, uint32_t M) {
    float ws = 0.f;
    if (sum(M, wt, m->scalar + 1, m->s) != hipSuccess) return -1.f;
    if (hipMemcpyAsync(&ws, m->scalar + 1, sizeof(float), hipMemcpyDeviceToHost, m->s) != hipSuccess) return -1.f;
    if (hipStreamSynchronize(m->s) != hipSuccess) return -1.f;
    return ws;
}

} // namespace

extern "C" {

int ise_mlp_create(ise_ctx *ctx, ise_mlp **out, uint32_t max_rows, uint64_t seed) {
    if (!ctx || !out || max_rows == 0) return ISE_ERR_ARG;
    *out = nullptr;
    ise_mlp *m = new ise_mlp();
    m->ctx = ctx;
    m->max_rows = (max_rows + 15) / 16 * 16;
    const size_t R = m->max_rows;
    hipError_t e = hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = dalloc(&m->w, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->g, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->m1, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->m2, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->wb, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->h1f, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->h2f, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->dh1, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->dh2, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->h1b, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->h2b, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->dz2, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->dz1, R * MLP_HID);
    if (e == hipSuccess) e = dalloc(&m->z3, R * MLP_OUT_PAD);
    if (e == hipSuccess) e = dalloc(&m->dz3, R * MLP_OUT_PAD);
    if (e == hipSuccess) e = dalloc(&m->lossv, R);
    if (e == hipSuccess) e = dalloc(&m->scalar, 16);
    if (e == hipSuccess) e = dalloc(&m->X, R * MLP_IN);
    if (e == hipSuccess) e = dalloc(&m->score, R);
    if (e != hipSuccess) {
        ise_mlp_destroy(m);
        return ISE_ERR_HIP;
    }
    std::vector<float> p;
    init_params(p, seed);
    int rc = ise_mlp_set_params(m, p.data());
    if (rc) {
        ise_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return ISE_OK;
}

int ise_mlp_destroy(ise_mlp *m) {
    if (!m) return ISE_ERR_ARG;
    if (m->s) (void)hipStreamSynchronize(m->s);
    void *bufs[] = {m->w, m->g, m->m1, m->m2, m->wb, m->h1f, m->h2f, m->dh1, m->dh2, m->h1b, m->h2b,
                    m->dz2, m->dz1, m->z3, m->dz3, m->lossv, m->scalar, m->X, m->score};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return ISE_OK;
}

int ise_mlp_set_params(ise_mlp *m, const float *p) {
    if (!m || !p) return ISE_ERR_ARG;
    CK(hipMemcpyAsync(m->w, p, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, m->s));
    CK(hipMemsetAsync(m->m1, 0, MLP_PARAMS * sizeof(float), m->s));
    CK(hipMemsetAsync(m->m2, 0, MLP_PARAMS * sizeof(float), m->s));
    CK(to_bf16(MLP_PARAMS, m->w, m->wb, m->s));
    CK(hipStreamSynchronize(m->s));
    m->step = 0;
    return ISE_OK;
}

int ise_mlp_get_params(const ise_mlp *m, float *p) {
    if (!m || !p) return ISE_ERR_ARG;
    CK(hipMemcpyAsync(p, m->w, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, m->s));
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int ise_mlp_features(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint32_t *d_meta,
                     uint32_t od, uint16_t *d_X) {
    if (!ctx || !d_rows || !d_meta || !d_X || od < 1 || od > ISE_MAX_ROWS || stride < 1) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    hipStream_t s = nullptr;
    CK(features(n, stride, d_rows, d_meta, od, d_X, s));
    CK(hipStreamSynchronize(s));
    return ISE_OK;
}

int ise_mlp_forward(ise_mlp *m, const uint16_t *X, uint32_t M, float *d_score) {
    if (!m || !X || !d_score || M == 0 || M % 16 || M > m->max_rows) return ISE_ERR_ARG;
    int rc = forward(m, X, M);
    if (rc) return rc;
    CK(hipMemcpy2DAsync(d_score, sizeof(float), m->z3, MLP_OUT_PAD * sizeof(float), sizeof(float), M,
                        hipMemcpyDeviceToDevice, m->s));
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int ise_mlp_gradient(ise_mlp *m, const uint16_t *X, const float *y, const float *wt, uint32_t M, float *host_grad,
                     float *loss) {
    if (!m || !X || !y || !wt || M == 0 || M % 16 || M > m->max_rows) return ISE_ERR_ARG;
    const float ws = weight_sum(m, wt, M);
    if (!(ws > 0.f)) return ISE_ERR_ARG;
    int rc = forward(m, X, M);
    if (!rc) rc = backward(m, X, y, wt, M, 1.f / ws);
    if (rc) return rc;
    if (host_grad) CK(hipMemcpyAsync(host_grad, m->g, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, m->s));
    float l = 0.f;
    CK(hipMemcpyAsync(&l, m->scalar, sizeof(float), hipMemcpyDeviceToHost, m->s));
    CK(hipStreamSynchronize(m->s));
    if (loss) *loss = l;
    return ISE_OK;
}

int ise_mlp_train_step(ise_mlp *m, const uint16_t *X, const float *y, const float *wt, uint32_t M, float lr,
                       float *loss) {
    float l = 0.f;
    int rc = ise_mlp_gradient(m, X, y, wt, M, nullptr, &l);
    if (rc) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - powf(b1, (float)m->step)), c2 = 1.f / (1.f - powf(b2, (float)m->step));
    CK(adam(MLP_PARAMS, m->w, m->g, m->m1, m->m2, m->wb, lr, b1, b2, 1e-8f, c1, c2, m->s));
    CK(hipStreamSynchronize(m->s));
    if (loss) *loss = l;
    return ISE_OK;
}

int ise_mlp_order(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint32_t *d_meta,
                  uint32_t od, uint8_t *d_order) {
    if (!m || !d_rows || !d_meta || !d_order || od < 1 || od > ISE_MAX_ROWS) return ISE_ERR_ARG;
    const uint32_t pairs = od * od;
    const uint32_t per = m->max_rows / pairs;   // candidates per chunk
    if (per == 0 || (pairs % 16 && per * pairs % 16)) return ISE_ERR_ARG;
    for (uint32_t c0 = 0; c0 < n; c0 += per) {
        uint32_t cnt = n - c0 < per ? n - c0 : per;
        uint32_t M = cnt * pairs;
        const uint32_t Mp = (M + 15) / 16 * 16;
        if (Mp != M) CK(hipMemsetAsync(m->X + (size_t)M * MLP_IN, 0, (size_t)(Mp - M) * MLP_IN * 2, m->s));
        CK(features(cnt, stride, d_rows + (size_t)c0 * stride, d_meta + c0, od, m->X, m->s));
        int rc = forward(m, m->X, Mp);
        if (rc) return rc;
        CK(order_from_scores(cnt, d_meta + c0, od, m->z3, MLP_OUT_PAD, d_order + (size_t)c0 * pairs, m->s));
    }
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

} // extern "C"
