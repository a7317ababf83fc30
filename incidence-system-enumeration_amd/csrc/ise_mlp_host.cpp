// ise_mlp_host.cpp -- the learned row-order model: parameters, Adam state,
// activation workspace and the forward / backward / training / ordering
// sequences built from the kernels of ise_mlp.hip, plus the witness-log
// labels it is trained on.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>
#include "ise_internal.h"
#include "ise_mlp.h"

using namespace ise;

struct ise_mlp {
    ise_ctx *ctx = nullptr;
    hipStream_t s = nullptr;   // the context's stream
    uint32_t max_rows = 0;
    int step = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // [MLP_PARAMS]
    uint16_t *wb = nullptr;                                           // bf16 copy [MLP_PARAMS]
    uint16_t *frag = nullptr;                                         // A fragments [MLP_FRAG_WORDS]
    // workspace of the unfused forward/backward
    float *h1f = nullptr, *h2f = nullptr, *z3 = nullptr, *dh2 = nullptr, *dh1 = nullptr, *lossv = nullptr;
    uint16_t *h1b = nullptr, *h2b = nullptr, *dz3 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *scalar = nullptr;
    int order_blocks = 0;
    uint32_t order_wpb = 4;
    // learned-order batches: candidates decided within this many plain
    // permutations are not scored (ise_mlp_set_gate; 0 scores all)
    uint32_t gate_perms = ISE_DEFAULT_LEARNED_GATE;
};

namespace {

int chk(hipError_t e) { return e == hipSuccess ? ISE_OK : ISE_ERR_HIP; }
#define CK(x)                          \
    do {                               \
        int rc_ = chk(x);              \
        if (rc_) return rc_;           \
    } while (0)

template <typename T> hipError_t dalloc(T **p, size_t count) { return hipMalloc((void **)p, count * sizeof(T) + 64); }

// He-uniform initial weights from splitmix64(seed); biases zero.
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

// gradient of the loss into m->g (same packing as the parameters); loss into m->scalar[0]
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

int weight_sum(ise_mlp *m, const float *wt, uint32_t M, float *ws) {
    CK(sum(M, wt, m->scalar + 1, m->s));
    CK(hipMemcpyAsync(ws, m->scalar + 1, sizeof(float), hipMemcpyDeviceToHost, m->s));
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int gradient(ise_mlp *m, const uint16_t *X, const float *y, const float *wt, uint32_t M) {
    if (!m || !X || !y || !wt || M == 0 || M % 16 || M > m->max_rows) return ISE_ERR_ARG;
    float ws = 0.f;
    int rc = weight_sum(m, wt, M, &ws);
    if (rc) return rc;
    if (!(ws > 0.f)) return ISE_ERR_ARG;
    rc = forward(m, X, M);
    if (!rc) rc = backward(m, X, y, wt, M, 1.f / ws);
    return rc;
}

int adam_step(ise_mlp *m, float lr) {
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - powf(b1, (float)m->step)), c2 = 1.f / (1.f - powf(b2, (float)m->step));
    CK(adam(MLP_PARAMS, m->w, m->g, m->m1, m->m2, m->wb, lr, b1, b2, 1e-8f, c1, c2, m->s));
    CK(frag_pack(m->wb, m->frag, m->s));
    return ISE_OK;
}

int order_launch(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint32_t *d_meta, uint32_t od,
                 uint8_t *d_order, float *d_score, const uint32_t *list = nullptr,
                 const uint32_t *list_count = nullptr) {
    uint32_t blocks = (uint32_t)m->order_blocks;
    const uint32_t need = (n + m->order_wpb - 1) / m->order_wpb;
    if (blocks > need) blocks = need;
    if (blocks == 0) blocks = 1;
    CK(learned_order(blocks, m->order_wpb, n, stride, d_rows, d_meta, od, m->frag, m->w, d_order, d_score, m->s,
                     list, list_count));
    return ISE_OK;
}

} // namespace

extern "C" {

int ise_mlp_create(ise_ctx *ctx, ise_mlp **out, uint32_t max_rows, uint64_t seed) {
    if (!ctx || !out || max_rows == 0 || max_rows > (1u << 26)) return ISE_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ISE_ERR_HIP;
    ise_mlp *m = new ise_mlp();
    m->ctx = ctx;
    m->s = ctx_stream(ctx);
    m->max_rows = (max_rows + 15) / 16 * 16;
    const size_t R = m->max_rows;
    // as many waves per workgroup (up to 4) as the per-workgroup LDS limit allows
    int max_lds = 0;
    hipError_t e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx_device(ctx));
    m->order_wpb = 4;
    while (m->order_wpb > 1 && learned_order_lds(m->order_wpb) > (size_t)max_lds) m->order_wpb--;
    int occ = 0;
    if (e == hipSuccess) e = learned_order_occupancy(m->order_wpb, &occ);
    m->order_blocks = (occ > 0 ? occ : 1) * (ctx_num_cus(ctx) > 0 ? ctx_num_cus(ctx) : 1);
    if (e == hipSuccess) e = dalloc(&m->w, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->g, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->m1, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->m2, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->wb, MLP_PARAMS);
    if (e == hipSuccess) e = dalloc(&m->frag, MLP_FRAG_WORDS);
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
    // a context still ordering by this network goes back to the plain order
    if (m->ctx && m->frag && ctx_lead_frag(m->ctx) == m->frag) (void)ctx_set_lead(m->ctx, nullptr, nullptr, 0);
    void *bufs[] = {m->w, m->g, m->m1, m->m2, m->wb, m->frag, m->h1f, m->h2f, m->dh1, m->dh2,
                    m->h1b, m->h2b, m->dz2, m->dz1, m->z3, m->dz3, m->lossv, m->scalar};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete m;
    return ISE_OK;
}

int ise_mlp_set_params(ise_mlp *m, const float *p) {
    if (!m || !p) return ISE_ERR_ARG;
    CK(hipMemcpyAsync(m->w, p, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, m->s));
    CK(hipMemsetAsync(m->m1, 0, MLP_PARAMS * sizeof(float), m->s));
    CK(hipMemsetAsync(m->m2, 0, MLP_PARAMS * sizeof(float), m->s));
    CK(to_bf16(MLP_PARAMS, m->w, m->wb, m->s));
    CK(frag_pack(m->wb, m->frag, m->s));
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
    CK(hipSetDevice(ctx_device(ctx)));
    CK(features(n, stride, d_rows, d_meta, od, d_X, ctx_stream(ctx)));
    CK(hipStreamSynchronize(ctx_stream(ctx)));
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
    int rc = gradient(m, X, y, wt, M);
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
    if (!(lr > 0.f)) return ISE_ERR_ARG;
    int rc = gradient(m, X, y, wt, M);
    if (!rc) rc = adam_step(m, lr);
    if (rc) return rc;
    float l = 0.f;
    CK(hipMemcpyAsync(&l, m->scalar, sizeof(float), hipMemcpyDeviceToHost, m->s));
    CK(hipStreamSynchronize(m->s));
    if (loss) *loss = l;
    return ISE_OK;
}

int ise_mlp_fit(ise_mlp *m, const uint16_t *X, const float *y, const float *wt, uint64_t M_total, uint32_t batch,
                uint32_t steps, float lr, float *loss_curve) {
    if (!m || !X || !y || !wt || batch == 0 || batch % 16 || batch > m->max_rows || M_total < batch || !(lr > 0.f))
        return ISE_ERR_ARG;
    const uint64_t nb = M_total / batch;
    for (uint32_t s = 0; s < steps; s++) {
        const uint64_t off = (uint64_t)(s % nb) * batch;
        int rc = gradient(m, X + off * MLP_IN, y + off, wt + off, batch);
        if (rc == ISE_ERR_ARG) continue;   // a batch without weight: nothing to learn from it
        if (!rc) rc = adam_step(m, lr);
        if (rc) return rc;
        if (loss_curve) CK(hipMemcpyAsync(&loss_curve[s], m->scalar, sizeof(float), hipMemcpyDeviceToHost, m->s));
    }
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int ise_mlp_order(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint32_t *d_meta,
                  uint32_t od, uint8_t *d_order, float *d_score) {
    if (!m || !d_rows || !d_meta || !d_order || od < 1 || od > ISE_MAX_ROWS || stride < 1) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    int rc = order_launch(m, n, stride, d_rows, d_meta, od, d_order, d_score);
    if (rc) return rc;
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int ise_mlp_check_batch_device_timed(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                     const uint32_t *d_meta, uint32_t od, uint8_t *d_order, uint32_t hunt_budget,
                                     const ise_results *d_out, uint32_t *d_counter, uint32_t repeats, float *ms_out,
                                     float *order_ms_out) {
    if (!m || !d_rows || !d_meta || !d_order || !d_out || !ms_out || repeats == 0 || hunt_budget == 0 || od < 1 ||
        od > ISE_MAX_ROWS)
        return ISE_ERR_ARG;
    ise_ctx *c = m->ctx;
    const OrderState saved = ctx_save_order(c);
    int rc = ise_set_order(c, ISE_ORDER_LEARNED, hunt_budget, d_order, od);
    if (rc) return rc;
    std::vector<hipEvent_t> ev(2 * repeats + 1, nullptr);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ISE_ERR_HIP;
    if (!rc && hipEventRecord(ev[0], m->s) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && !rc; i++) {
        if (m->gate_perms) {
            // plain search with a small budget for all; the network orders only
            // the candidates it leaves open, which then run hunt + plain search
            const uint32_t *list = nullptr;
            rc = launch_gated_pass(c, n, stride, d_rows, d_meta, *d_out, d_counter, m->gate_perms, 1, 0, nullptr,
                                   od, &list);
            if (!rc) rc = order_launch(m, n, stride, d_rows, d_meta, od, d_order, nullptr, list,
                                       defer_count_ptr(d_counter, 1));
            if (!rc && hipEventRecord(ev[2 * i + 1], m->s) != hipSuccess) rc = ISE_ERR_HIP;
            if (!rc) rc = launch_gated_pass(c, n, stride, d_rows, d_meta, *d_out, d_counter, m->gate_perms, 2,
                                            hunt_budget, d_order, od, nullptr);
        } else {
            rc = order_launch(m, n, stride, d_rows, d_meta, od, d_order, nullptr);
            if (!rc && hipEventRecord(ev[2 * i + 1], m->s) != hipSuccess) rc = ISE_ERR_HIP;
            if (!rc) rc = ise_check_batch_device(c, n, stride, d_rows, nullptr, d_meta, d_out, d_counter, nullptr);
        }
        if (!rc && hipEventRecord(ev[2 * i + 2], m->s) != hipSuccess) rc = ISE_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(m->s) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && !rc; i++) {
        if (hipEventElapsedTime(&ms_out[i], ev[2 * i], ev[2 * i + 2]) != hipSuccess) rc = ISE_ERR_HIP;
        if (!rc && order_ms_out && hipEventElapsedTime(&order_ms_out[i], ev[2 * i], ev[2 * i + 1]) != hipSuccess)
            rc = ISE_ERR_HIP;
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    const int rc2 = ctx_restore_order(c, saved);
    return rc ? rc : rc2;
}

int ise_mlp_lead_features(ise_ctx *ctx, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                          const uint32_t *d_meta, uint16_t *d_X) {
    if (!ctx || !d_X || stride < 1 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    CK(hipSetDevice(ctx_device(ctx)));
    CK(lead_features(n, stride, d_rows, d_meta, d_X, ctx_stream(ctx)));
    CK(hipStreamSynchronize(ctx_stream(ctx)));
    return ISE_OK;
}

int ise_mlp_check_lead_device(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                              const uint32_t *d_meta, const ise_results *d_out, uint32_t *d_counter,
                              const uint8_t *d_lead_in, uint8_t *d_lead_out, float *d_score_out) {
    if (!m || !d_out || !d_counter || stride < 1 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    int rc = launch_lead_ctx(m->ctx, n, stride, d_rows, d_meta, *d_out, d_counter, m->gate_perms, m->frag, m->w,
                             d_lead_in, d_lead_out, d_score_out);
    if (rc) return rc;
    CK(hipStreamSynchronize(m->s));
    return ISE_OK;
}

int ise_mlp_check_lead_device_timed(ise_mlp *m, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                    const uint32_t *d_meta, const ise_results *d_out, uint32_t *d_counter,
                                    uint32_t repeats, float *ms_out) {
    if (!m || !d_out || !d_counter || !ms_out || repeats == 0 || stride < 1 || stride > ISE_MAX_ROWS)
        return ISE_ERR_ARG;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    std::vector<hipEvent_t> ev(repeats + 1, nullptr);
    int rc = ISE_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ISE_ERR_HIP;
    if (!rc && hipEventRecord(ev[0], m->s) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && !rc; i++) {
        rc = launch_lead_ctx(m->ctx, n, stride, d_rows, d_meta, *d_out, d_counter, m->gate_perms, m->frag, m->w,
                             nullptr, nullptr, nullptr);
        if (!rc && hipEventRecord(ev[i + 1], m->s) != hipSuccess) rc = ISE_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(m->s) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && !rc; i++)
        if (hipEventElapsedTime(&ms_out[i], ev[i], ev[i + 1]) != hipSuccess) rc = ISE_ERR_HIP;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int ise_mlp_use_lead(ise_mlp *m, int on) {
    if (!m) return ISE_ERR_ARG;
    if (on) return ctx_set_lead(m->ctx, m->frag, m->w, m->gate_perms);
    if (ctx_lead_frag(m->ctx) == m->frag) return ctx_set_lead(m->ctx, nullptr, nullptr, 0);
    return ISE_OK;
}

int ise_mlp_set_gate(ise_mlp *m, uint32_t gate_perms) {
    if (!m) return ISE_ERR_ARG;
    m->gate_perms = gate_perms;
    // a context ordering by this network takes the new gate
    if (m->ctx && ctx_lead_frag(m->ctx) == m->frag) return ctx_set_lead(m->ctx, m->frag, m->w, gate_perms);
    return ISE_OK;
}

int ise_mlp_witness_labels(uint32_t n, uint32_t od, const uint32_t *meta, const uint8_t *verdict,
                           const uint8_t *fail_row, const uint8_t *perm_row, float *y, float *wt) {
    if (!meta || !verdict || !fail_row || !perm_row || !y || !wt || od < 1 || od > ISE_MAX_ROWS) return ISE_ERR_ARG;
    for (uint32_t c = 0; c < n; c++) {
        const uint32_t d = meta[c] & 0xff;
        const uint32_t f = fail_row[c];
        float *yc = y + (size_t)c * od * od, *wc = wt + (size_t)c * od * od;
        memset(yc, 0, sizeof(float) * od * od);
        memset(wc, 0, sizeof(float) * od * od);
        if (verdict[c] != ISE_VERDICT_NONCANONICAL || d > od || f >= d) continue;
        const uint8_t *w = perm_row + (size_t)c * ISE_MAX_ROWS;
        for (uint32_t i = 0; i <= f; i++)
            for (uint32_t j = 0; j < d; j++) {
                yc[i * od + j] = w[i] == j ? 1.f : 0.f;
                wc[i * od + j] = 1.f;
            }
    }
    return ISE_OK;
}

} // extern "C"
