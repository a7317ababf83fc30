// ise_runtime.cpp -- contexts, device buffers and the C ABI of the
// canonicity kernels (declared in include/ise.h).
//
// A context owns one HIP stream and device buffers for `capacity` candidates
// (inputs and every result array), so that a steady stream of batches causes
// no allocation after ise_open.  This replaces the reference's per-matrix
// device-heap checker objects (Source/CanonicityChecker.cu:17-68) and its
// batches of at most 128 matrices (Source/CudaInfo.h:9).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <map>
#include <string>
#include <tuple>
#include <vector>
#include "ise_internal.h"
#include "ise_kfact.h"
#include "ise_noalloc.h"
#include "ise_canon.h"

struct ise_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    uint32_t capacity = 0, stride = 0;
    // device inputs
    uint64_t *d_rows = nullptr, *d_cells = nullptr;
    uint32_t *d_meta = nullptr, *d_head = nullptr;
    // device results
    ise_results d{};
    uint32_t gen_cap_alloc = 0;
    // host queue filled by ise_assign
    std::vector<uint64_t> q_rows;
    std::vector<uint32_t> q_meta;
    // last submit
    uint32_t last_n = 0, last_gen_cap = 0;
    bool pending = false, last_cells = false;
    // ise_submit reports column permutations (ise_report_columns)
    bool report_columns = true, last_columns = true;
    // deferred long searches (ise_set_defer): first-pass budget and the list
    uint32_t defer_perms[ise::DEFER_BUDGETS] = {ISE_DEFAULT_DEFER_PERMS, ISE_DEFAULT_DEFER_PERMS2};
    uint32_t *d_defer = nullptr;
    size_t defer_cap = 0;
    // deepest candidates first (ise_set_schedule): list of n + SCHED_EXTRA_WORDS words
    int schedule = ISE_SCHEDULE_DEEPEST_FIRST;
    uint32_t *d_sched = nullptr;
    size_t sched_cap = 0;
    ise_launch_cfg cfg{0, 4, 0, 1u << 26};
    // search order (ise_set_order; ISE_ORDER_LEAD by ise_mlp_use_lead: the
    // network's fragments and packed parameters, owned by the ise_mlp)
    int order_mode = ISE_ORDER_PLAIN;
    uint32_t hunt_budget = 0, order_dim = 0;
    const uint8_t *d_order = nullptr;
    const uint16_t *lead_frag = nullptr;
    const float *lead_w = nullptr;
    uint32_t lead_gate = 0;
    // 1-factorization path: QuickPerm tables and per-candidate results
    uint64_t *d_kf_tab = nullptr;
    ise::KfTables kf_tab{};
    uint64_t *d_kf_witness = nullptr, *d_kf_auts = nullptr, *d_kf_tried = nullptr;
    // split launch (ise_kf_set_split): two lists of kf_cap entries (candidate,
    // automorphisms before the window), published witnesses, range results
    uint32_t kf_first_chunks = ISE_DEFAULT_KF_FIRST_CHUNKS, kf_max_ranges = ISE_DEFAULT_KF_MAX_RANGES;
    uint32_t *d_kf_list = nullptr, *d_kf_best = nullptr;
    uint64_t *d_kf_part = nullptr, *d_kf_ranges = nullptr;
    size_t kf_cap = 0, kf_range_words = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // batched CanonizeMatrix (ise_canonize_*): two compact work batches of
    // canon_cap matrices (rows, meta, original index), the open count, and
    // result arrays for the host entry point
    uint64_t *d_canon_rows[2] = {nullptr, nullptr};
    uint64_t *d_canon_cells[2] = {nullptr, nullptr};   // orbit-start masks under a column partition
    uint32_t *d_canon_meta[2] = {nullptr, nullptr}, *d_canon_idx[2] = {nullptr, nullptr};
    uint32_t *d_canon_count = nullptr;
    uint64_t *d_canon_out = nullptr, *d_canon_aut = nullptr;
    uint8_t *d_canon_status = nullptr;
    uint32_t *d_canon_iters = nullptr;
    size_t canon_cap = 0;
    // the device's scratch array of ise_check_batch_device_noalloc (not owned)
    uint32_t *noalloc_scratch = nullptr;
    std::string noalloc_err;
    // executable graphs of ise_submit launches, keyed by batch shape without
    // the batch size: a batch of another size updates the graph in place
    // (hipGraphExecUpdate), so every level and tail chunk of an enumeration
    // replays one graph.  The kernels' arguments hold the addresses of context
    // buffers and the deferral/schedule settings, so every cached graph is
    // dropped when one of those buffers is reallocated or a setting changes
    // (drop_graphs).
    bool use_graphs = false;
    using GraphKey = std::tuple<uint32_t, int, bool, uint32_t, uint32_t, uint32_t, uint32_t, int,
                                uint32_t, uintptr_t, uint32_t, bool>;
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        uint32_t n = 0;
    };
    std::map<GraphKey, GraphEntry> graphs;
    uint64_t graph_replays = 0, graph_updates = 0;
    float last_ms = 0.f;
    std::string err;
};

namespace {

int fail(ise_ctx *c, hipError_t e, const char *what) {
    if (c) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
    }
    return ISE_ERR_HIP;
}

#define HIPCHK(ctx, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return fail(ctx, e_, #call);  \
    } while (0)

template <typename T> hipError_t dalloc(T **p, size_t count) {
    return hipMalloc((void **)p, count * sizeof(T) + 16);
}

// Destroy every cached graph (after the work already queued has finished):
// called before a buffer whose address a graph holds is freed, and when a
// launch setting that the graph key leaves out changes.
int drop_graphs(ise_ctx *c) {
    if (c->graphs.empty()) return ISE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return ISE_OK;
}

int alloc_gens(ise_ctx *c, uint32_t gen_cap) {
    if (gen_cap <= c->gen_cap_alloc) return ISE_OK;
    if (int rc = drop_graphs(c)) return rc;
    if (c->d.gens) HIPCHK(c, hipFree(c->d.gens));
    c->d.gens = nullptr;
    HIPCHK(c, dalloc(&c->d.gens, (size_t)c->capacity * gen_cap * ISE_MAX_COLS));
    c->gen_cap_alloc = gen_cap;
    return ISE_OK;
}

// Validate host-side records and pick the lane-group size.
int scan_meta(const uint32_t *meta, uint32_t n, uint32_t stride, uint32_t *max_d, uint32_t *max_b) {
    uint32_t md = 0, mb = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t m = meta[i], d = m & 0xff, b = (m >> 8) & 0xff, f = (m >> 16) & 0xff;
        if (d < 1 || d > ISE_MAX_ROWS || d > stride || b < 1 || b > ISE_MAX_COLS || f >= d)
            return ISE_ERR_ARG;
        if (d > md) md = d;
        if (b > mb) mb = b;
    }
    *max_d = md;
    *max_b = mb;
    return ISE_OK;
}

// Persistent grid size: resident workgroups per CU times CUs, capped by the work.
int grid_blocks(ise_ctx *c, uint32_t n, uint32_t stride, int group_lanes, const ise_results &out,
                uint32_t *blocks) {
    const uint32_t wpb = c->cfg.waves_per_block ? c->cfg.waves_per_block : 4;
    const uint32_t budget = c->order_mode == ISE_ORDER_PLAIN ? 0u : c->hunt_budget;
    uint32_t b = c->cfg.blocks;
    if (!b) {
        int per_cu = 0;
        if (c->order_mode == ISE_ORDER_LEAD) {
            // one candidate per wave
            group_lanes = 64;
            HIPCHK(c, ise::lead_occupancy(wpb, stride, &per_cu));
        } else {
            HIPCHK(c, ise::occupancy_blocks(group_lanes, wpb, stride, ise::uses_sets(out, budget), &per_cu));
        }
        if (per_cu < 1) per_cu = 1;
        b = (uint32_t)per_cu * (uint32_t)c->num_cus;
        const uint64_t groups_needed = ((uint64_t)n * group_lanes + 64ull * wpb - 1) / (64ull * wpb);
        if (b > groups_needed) b = (uint32_t)(groups_needed ? groups_needed : 1);
    }
    *blocks = b;
    return ISE_OK;
}

// Enqueue queue reset + kernel; `blocks` from grid_blocks (no runtime queries
// here, so the sequence can be captured into a graph).
// use_sched = false: candidates in index order, no deepest-first schedule
// (it changes no result; batches of one depth gain nothing from it).
int launch_with(ise_ctx *c, uint32_t blocks, uint32_t n, uint32_t stride, const uint64_t *rows,
                const uint64_t *cells, const uint32_t *meta, const ise_results &out, uint32_t *head,
                hipStream_t s, int group_lanes, bool allow_lead = true, bool use_sched = true) {
    const uint32_t wpb = c->cfg.waves_per_block ? c->cfg.waves_per_block : 4;
    if (c->order_mode == ISE_ORDER_LEAD && allow_lead && !cells && !out.perm_col) {
        // the learned order of the first row: one launch, deepest-first schedule
        uint32_t *sched = (c->schedule == ISE_SCHEDULE_DEEPEST_FIRST && n <= c->sched_cap) ? c->d_sched : nullptr;
        HIPCHK(c, ise::launch_lead(blocks, wpb, n, stride, rows, meta, out, head,
                                   c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), c->lead_gate, c->lead_frag,
                                   c->lead_w, nullptr, nullptr, nullptr, sched, s));
        return ISE_OK;
    }
    const uint32_t budget = c->order_mode == ISE_ORDER_PLAIN ? 0u : c->hunt_budget;
    const uint8_t *order = c->order_mode == ISE_ORDER_LEARNED ? c->d_order : nullptr;
    uint32_t *defer = (c->defer_perms[0] && n <= c->defer_cap) ? c->d_defer : nullptr;
    uint32_t *sched =
        (use_sched && c->schedule == ISE_SCHEDULE_DEEPEST_FIRST && n <= c->sched_cap) ? c->d_sched : nullptr;
    HIPCHK(c, ise::launch_check(group_lanes, blocks, wpb, n, stride, rows, cells, meta, out, head,
                                c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), budget, order,
                                c->order_dim, defer, c->defer_perms, sched, s));
    return ISE_OK;
}

// Schedule list for batches of up to n, grown outside any launch.
int ensure_sched(ise_ctx *c, size_t n) {
    if (c->schedule != ISE_SCHEDULE_DEEPEST_FIRST || n <= c->sched_cap) return ISE_OK;
    if (int rc = drop_graphs(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_sched) HIPCHK(c, hipFree(c->d_sched));
    c->d_sched = nullptr;
    c->sched_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_sched, (n + ise::SCHED_EXTRA_WORDS) * sizeof(uint32_t)));
    c->sched_cap = n;
    return ISE_OK;
}

// Lists of deferred candidates (two halves of n words) for batches of up to
// n, grown outside any launch.
int ensure_defer(ise_ctx *c, size_t n) {
    if (!c->defer_perms[0] || n <= c->defer_cap) return ISE_OK;
    if (int rc = drop_graphs(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_defer) HIPCHK(c, hipFree(c->d_defer));
    c->d_defer = nullptr;
    c->defer_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_defer, 2 * n * sizeof(uint32_t)));
    c->defer_cap = n;
    return ISE_OK;
}

// Range results of the split 1-factorization launch: at least this many
// words whatever the capacity, so that small batches get many ranges each.
constexpr size_t KF_MIN_RANGE_WORDS = size_t(1) << 17;

int ensure_kf_split(ise_ctx *c, size_t n) {
    if (!c->kf_first_chunks || n <= c->kf_cap) return ISE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    void *bufs[] = {c->d_kf_list, c->d_kf_best, c->d_kf_part, c->d_kf_ranges};
    for (void *p : bufs)
        if (p) HIPCHK(c, hipFree(p));
    c->d_kf_list = c->d_kf_best = nullptr;
    c->d_kf_part = c->d_kf_ranges = nullptr;
    c->kf_cap = c->kf_range_words = 0;
    const size_t words = n > KF_MIN_RANGE_WORDS ? n : KF_MIN_RANGE_WORDS;
    HIPCHK(c, hipMalloc(&c->d_kf_list, 2 * n * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&c->d_kf_best, n * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&c->d_kf_part, 2 * n * sizeof(uint64_t)));
    HIPCHK(c, hipMalloc(&c->d_kf_ranges, words * sizeof(uint64_t)));
    c->kf_cap = n;
    c->kf_range_words = words;
    return ISE_OK;
}

int launch(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint64_t *cells,
           const uint32_t *meta, const ise_results &out, uint32_t *head, hipStream_t s,
           int group_lanes, bool allow_lead = true, bool use_sched = true) {
    uint32_t blocks = 0;
    int rc = grid_blocks(c, n, stride, group_lanes, out, &blocks);
    if (rc) return rc;
    return launch_with(c, blocks, n, stride, rows, cells, meta, out, head, s, group_lanes, allow_lead, use_sched);
}

} // namespace

namespace ise {
int launch_gated_pass(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                      const ise_results &out, uint32_t *head, uint32_t gate, int pass, uint32_t hunt_budget,
                      const uint8_t *order, uint32_t order_dim, const uint32_t **list) {
    if (!c || gate == 0 || (pass != 1 && pass != 2) || !head || !out.verdict) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    // the lists live in the deferral buffer even when deferral is off
    const uint32_t saved = c->defer_perms[0];
    c->defer_perms[0] = gate;
    int rc = ensure_defer(c, n);
    c->defer_perms[0] = saved;
    if (rc == ISE_OK) rc = ensure_sched(c, n);
    if (rc) return rc;
    const int group = c->cfg.group_lanes ? (int)c->cfg.group_lanes : 64;
    uint32_t blocks = 0;
    rc = grid_blocks(c, n, stride, group, out, &blocks);
    if (rc) return rc;
    const uint32_t wpb = c->cfg.waves_per_block ? c->cfg.waves_per_block : 4;
    const uint32_t budgets[DEFER_BUDGETS] = {gate, 0u};
    uint32_t *sched = c->schedule == ISE_SCHEDULE_DEEPEST_FIRST ? c->d_sched : nullptr;
    HIPCHK(c, launch_check(group, blocks, wpb, n, stride, rows, nullptr, meta, out, head,
                           c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), pass == 2 ? hunt_budget : 0u,
                           pass == 2 ? order : nullptr, order_dim, c->d_defer, budgets, sched, c->stream, pass,
                           pass));
    if (list) *list = c->d_defer;
    return ISE_OK;
}

int launch_lead_ctx(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                    const ise_results &out, uint32_t *head, uint32_t gate, const uint16_t *frag, const float *w,
                    const uint8_t *lead_in, uint8_t *lead_out, float *score_out) {
    if (!c || !head || !out.verdict || out.perm_col || (!frag && !lead_in)) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_sched(c, n);
    if (rc) return rc;
    // one candidate per wave; the context's waves per workgroup
    const uint32_t wpb = c->cfg.waves_per_block ? c->cfg.waves_per_block : 4;
    uint32_t blocks = c->cfg.blocks;
    if (!blocks) {
        int per_cu = 0;
        HIPCHK(c, ise::lead_occupancy(wpb, stride, &per_cu));
        if (per_cu < 1) per_cu = 1;
        blocks = (uint32_t)per_cu * (uint32_t)c->num_cus;
        const uint64_t need = ((uint64_t)n + wpb - 1) / wpb;
        if (blocks > need) blocks = (uint32_t)(need ? need : 1);
    }
    uint32_t *sched = c->schedule == ISE_SCHEDULE_DEEPEST_FIRST ? c->d_sched : nullptr;
    HIPCHK(c, ise::launch_lead(blocks, wpb, n, stride, rows, meta, out, head,
                               c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), gate, frag, w, lead_in, lead_out,
                               score_out, sched, c->stream));
    return ISE_OK;
}

uint32_t ctx_stride(const ise_ctx *c) { return c->stride; }
uint32_t ctx_capacity(const ise_ctx *c) { return c->capacity; }
hipStream_t ctx_stream(const ise_ctx *c) { return c->stream; }
int ctx_device(const ise_ctx *c) { return c->device; }
int ctx_num_cus(const ise_ctx *c) { return c->num_cus; }
void ctx_set_error(ise_ctx *c, const char *msg) { c->err = msg; }
bool ctx_report_columns(ise_ctx *c, bool on) {
    const bool prev = c->report_columns;
    c->report_columns = on;
    return prev;
}
int ctx_set_lead(ise_ctx *c, const uint16_t *frag, const float *w, uint32_t gate) {
    if (int rc = drop_graphs(c)) return rc;
    c->order_mode = frag ? ISE_ORDER_LEAD : ISE_ORDER_PLAIN;
    c->hunt_budget = 0;
    c->d_order = nullptr;
    c->order_dim = 0;
    c->lead_frag = frag;
    c->lead_w = frag ? w : nullptr;
    c->lead_gate = frag ? gate : 0;
    return ISE_OK;
}
const uint16_t *ctx_lead_frag(const ise_ctx *c) { return c->order_mode == ISE_ORDER_LEAD ? c->lead_frag : nullptr; }
OrderState ctx_save_order(const ise_ctx *c) {
    return OrderState{c->order_mode, c->hunt_budget, c->d_order, c->order_dim, c->lead_frag, c->lead_w,
                      c->lead_gate};
}
int ctx_restore_order(ise_ctx *c, const OrderState &st) {
    if (st.mode == ISE_ORDER_LEAD) return ctx_set_lead(c, st.lead_frag, st.lead_w, st.lead_gate);
    c->lead_frag = nullptr;
    c->lead_w = nullptr;
    c->lead_gate = 0;
    return ise_set_order(c, st.mode, st.budget, st.order, st.dim);
}
void ctx_get_order(const ise_ctx *c, int *mode, uint32_t *budget, const uint8_t **order, uint32_t *dim) {
    *mode = c->order_mode;
    *budget = c->hunt_budget;
    *order = c->d_order;
    *dim = c->order_dim;
}
} // namespace ise

extern "C" {

int ise_open(ise_ctx **out, int device, uint32_t capacity, uint32_t stride) {
    if (!out || capacity == 0 || stride == 0 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return ISE_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ISE_ERR_NODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ISE_ERR_NODEV;
    ise_ctx *c = new ise_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    c->capacity = capacity;
    c->stride = stride;
    int rc = ISE_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = ISE_ERR_NODEV; break; }
        const size_t n = capacity;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = dalloc(&c->d_rows, n * stride);
        if (e == hipSuccess) e = dalloc(&c->d_cells, n * stride);
        if (e == hipSuccess) e = dalloc(&c->d_meta, n);
        if (e == hipSuccess) e = dalloc(&c->d_head, ISE_QUEUE_WORDS);
        if (e == hipSuccess) e = dalloc(&c->d_defer, 2 * n);
        if (e == hipSuccess) c->defer_cap = n;
        if (e == hipSuccess) e = dalloc(&c->d_sched, n + ise::SCHED_EXTRA_WORDS);
        if (e == hipSuccess) c->sched_cap = n;
        if (e == hipSuccess) e = dalloc(&c->d.verdict, n);
        if (e == hipSuccess) e = dalloc(&c->d.fail_row, n);
        if (e == hipSuccess) e = dalloc(&c->d.fail_col, n);
        if (e == hipSuccess) e = dalloc(&c->d.row_to_change, n);
        if (e == hipSuccess) e = dalloc(&c->d.group_order, n);
        if (e == hipSuccess) e = dalloc(&c->d.group_order_hi, n);
        if (e == hipSuccess) e = dalloc(&c->d.perm_row, n * ISE_MAX_ROWS);
        if (e == hipSuccess) e = dalloc(&c->d.perm_col, n * ISE_MAX_COLS);
        if (e == hipSuccess) e = dalloc(&c->d.orbits, n * ISE_MAX_ROWS);
        if (e == hipSuccess) e = dalloc(&c->d.n_gens, n);
        if (e == hipSuccess) e = dalloc(&c->d.counters, n * 4);
        if (e == hipSuccess) e = dalloc(&c->d_kf_witness, n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_auts, n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_tried, n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_list, 2 * n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_best, n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_part, 2 * n);
        if (e == hipSuccess) e = dalloc(&c->d_kf_ranges, n > KF_MIN_RANGE_WORDS ? n : KF_MIN_RANGE_WORDS);
        if (e == hipSuccess) {
            c->kf_cap = n;
            c->kf_range_words = n > KF_MIN_RANGE_WORDS ? n : KF_MIN_RANGE_WORDS;
        }
        if (e == hipSuccess) {
            std::vector<uint64_t> words(ise::kf_table_words());
            ise::kf_build_tables(words.data(), c->kf_tab.offset, c->kf_tab.count);
            e = dalloc(&c->d_kf_tab, words.size());
            if (e == hipSuccess)
                e = hipMemcpy(c->d_kf_tab, words.data(), words.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
            c->kf_tab.base = c->d_kf_tab;
            c->kf_tab.inv = c->d_kf_tab + words.size() / 2;
        }
        if (e == hipSuccess) {
            // allocation-free entry point: its scratch address on this device
            const hipError_t se = ise::noalloc_scratch_init(device, &c->noalloc_scratch);
            if (se != hipSuccess) {
                c->noalloc_scratch = nullptr;
                c->noalloc_err = std::string("scratch lookup: ") + hipGetErrorString(se);
                (void)hipGetLastError();
            }
        }
        if (e == hipSuccess) e = hipEventCreate(&c->ev0);
        if (e == hipSuccess) e = hipEventCreate(&c->ev1);
        if (e != hipSuccess) { rc = fail(c, e, "ise_open"); break; }
    } while (0);
    if (rc != ISE_OK) {
        ise_close(c);
        return rc;
    }
    *out = c;
    return ISE_OK;
}

int ise_close(ise_ctx *c) {
    if (!c) return ISE_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->d_rows, c->d_cells, c->d_meta, c->d_head, c->d.verdict, c->d.fail_row,
                    c->d.fail_col, c->d.row_to_change, c->d.group_order, c->d.group_order_hi, c->d.perm_row,
                    c->d.perm_col, c->d.orbits, c->d.n_gens, c->d.counters, c->d.gens,
                    c->d_kf_tab, c->d_kf_witness, c->d_kf_auts, c->d_kf_tried, c->d_defer, c->d_sched,
                    c->d_kf_list, c->d_kf_best, c->d_kf_part, c->d_kf_ranges,
                    c->d_canon_rows[0], c->d_canon_rows[1], c->d_canon_cells[0], c->d_canon_cells[1],
                    c->d_canon_meta[0], c->d_canon_meta[1],
                    c->d_canon_idx[0], c->d_canon_idx[1], c->d_canon_count, c->d_canon_out, c->d_canon_aut,
                    c->d_canon_status, c->d_canon_iters};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ISE_OK;
}

int ise_set_launch(ise_ctx *c, const ise_launch_cfg *cfg) {
    if (!c || !cfg) return ISE_ERR_ARG;
    if (cfg->group_lanes && cfg->group_lanes != 16 && cfg->group_lanes != 32 && cfg->group_lanes != 64)
        return ISE_ERR_ARG;
    if (cfg->waves_per_block > 16) return ISE_ERR_ARG;
    c->cfg = *cfg;
    return ISE_OK;
}

int ise_get_launch(const ise_ctx *c, ise_launch_cfg *cfg) {
    if (!c || !cfg) return ISE_ERR_ARG;
    *cfg = c->cfg;
    return ISE_OK;
}

int ise_assign(ise_ctx *c, const uint64_t *rows, uint32_t depth, uint32_t ncols,
               uint32_t len_stab, uint32_t out_info) {
    if (!c || !rows || depth < 1 || depth > c->stride || ncols < 1 || ncols > ISE_MAX_COLS ||
        len_stab >= depth || out_info > 0xff)
        return ISE_ERR_ARG;
    if (c->q_meta.size() >= c->capacity) return ISE_ERR_CAPACITY;
    const size_t at = c->q_rows.size();
    c->q_rows.resize(at + c->stride, 0ull);
    memcpy(&c->q_rows[at], rows, depth * sizeof(uint64_t));
    c->q_meta.push_back(depth | (ncols << 8) | (len_stab << 16) | (out_info << 24));
    return ISE_OK;
}

int ise_num_assigned(const ise_ctx *c, uint32_t *n) {
    if (!c || !n) return ISE_ERR_ARG;
    *n = (uint32_t)c->q_meta.size();
    return ISE_OK;
}

int ise_submit(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows,
               const uint64_t *cells, const uint32_t *meta, uint32_t gen_cap) {
    if (!c) return ISE_ERR_ARG;
    bool from_queue = false;
    if (n == 0 && !rows) {
        from_queue = true;
        n = (uint32_t)c->q_meta.size();
        stride = c->stride;
        rows = c->q_rows.data();
        meta = c->q_meta.data();
        cells = nullptr;
    }
    if (n > c->capacity) return ISE_ERR_CAPACITY;
    if (n && (!rows || !meta || stride != c->stride)) return ISE_ERR_ARG;
    uint32_t md = 1, mb = 1;
    if (n && scan_meta(meta, n, stride, &md, &mb) != ISE_OK) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (gen_cap) {
        int rc = alloc_gens(c, gen_cap);
        if (rc) return rc;
    }
    c->last_n = n;
    c->last_gen_cap = gen_cap;
    c->last_cells = cells != nullptr;
    c->last_columns = c->report_columns;
    if (n) {
        HIPCHK(c, hipMemcpyAsync(c->d_rows, rows, (size_t)n * stride * 8, hipMemcpyHostToDevice, c->stream));
        if (cells)
            HIPCHK(c, hipMemcpyAsync(c->d_cells, cells, (size_t)n * stride * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_meta, meta, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        ise_results o = c->d;
        o.gens = gen_cap ? c->d.gens : nullptr;
        o.gen_cap = gen_cap;
        if (!c->report_columns) o.perm_col = nullptr;
        int group = c->cfg.group_lanes ? (int)c->cfg.group_lanes : ise::fit_group_lanes(md, mb);
        if ((uint32_t)group < md || (uint32_t)group < mb) return ISE_ERR_ARG;
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        if (c->use_graphs) {
            const ise_ctx::GraphKey key{stride, group, cells != nullptr, gen_cap, c->cfg.waves_per_block,
                                        c->cfg.blocks, c->cfg.max_perms, c->order_mode, c->hunt_budget,
                                        (uintptr_t)c->d_order, c->order_dim, c->report_columns};
            auto it = c->graphs.find(key);
            if (it == c->graphs.end() || it->second.n != n) {
                // capture: every pointer is a context buffer, so replays see new data
                hipGraph_t g = nullptr;
                uint32_t blocks = 0;
                int rc = grid_blocks(c, n, stride, group, o, &blocks);
                if (rc) return rc;
                HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                rc = launch_with(c, blocks, n, stride, c->d_rows, cells ? c->d_cells : nullptr, c->d_meta,
                                 o, c->d_head, c->stream, group);
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return fail(c, e, "hipStreamEndCapture");
                bool updated = false;
                if (it != c->graphs.end()) {
                    // same kernels with other sizes: update the executable graph in
                    // place, once its replays already queued have finished
                    HIPCHK(c, hipStreamSynchronize(c->stream));
                    hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
                    hipGraphNode_t err_node = nullptr;
                    updated = hipGraphExecUpdate(it->second.exec, g, &err_node, &ur) == hipSuccess &&
                              ur == hipGraphExecUpdateSuccess;
                    if (updated) {
                        it->second.n = n;
                        c->graph_updates++;
                    } else {
                        (void)hipGetLastError();
                        HIPCHK(c, hipStreamSynchronize(c->stream));
                        (void)hipGraphExecDestroy(it->second.exec);
                        c->graphs.erase(it);
                        it = c->graphs.end();
                    }
                }
                if (!updated) {
                    hipGraphExec_t ge = nullptr;
                    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                    if (e != hipSuccess) {
                        (void)hipGraphDestroy(g);
                        return fail(c, e, "hipGraphInstantiate");
                    }
                    it = c->graphs.emplace(key, ise_ctx::GraphEntry{ge, n}).first;
                }
                (void)hipGraphDestroy(g);
            }
            HIPCHK(c, hipGraphLaunch(it->second.exec, c->stream));
            c->graph_replays++;
        } else {
            int rc = launch(c, n, stride, c->d_rows, cells ? c->d_cells : nullptr, c->d_meta, o,
                            c->d_head, c->stream, group);
            if (rc) return rc;
        }
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    }
    if (from_queue) {
        c->q_rows.clear();
        c->q_meta.clear();
    }
    c->pending = true;
    return ISE_OK;
}

int ise_collect(ise_ctx *c, ise_results *h) {
    if (!c || !h) return ISE_ERR_ARG;
    if (!c->pending) return ISE_ERR_STATE;
    const size_t n = c->last_n;
    hipStream_t s = c->stream;
    HIPCHK(c, hipSetDevice(c->device));
    if (n && h->perm_col && !c->last_columns) {
        c->err = "column permutations were not reported for this batch (ise_report_columns)";
        return ISE_ERR_STATE;
    }
    if (n) {
        if (h->verdict) HIPCHK(c, hipMemcpyAsync(h->verdict, c->d.verdict, n, hipMemcpyDeviceToHost, s));
        if (h->fail_row) HIPCHK(c, hipMemcpyAsync(h->fail_row, c->d.fail_row, n, hipMemcpyDeviceToHost, s));
        if (h->fail_col) HIPCHK(c, hipMemcpyAsync(h->fail_col, c->d.fail_col, n, hipMemcpyDeviceToHost, s));
        if (h->row_to_change)
            HIPCHK(c, hipMemcpyAsync(h->row_to_change, c->d.row_to_change, n, hipMemcpyDeviceToHost, s));
        if (h->group_order)
            HIPCHK(c, hipMemcpyAsync(h->group_order, c->d.group_order, n * 8, hipMemcpyDeviceToHost, s));
        if (h->group_order_hi)
            HIPCHK(c, hipMemcpyAsync(h->group_order_hi, c->d.group_order_hi, n * 8, hipMemcpyDeviceToHost, s));
        if (h->perm_row)
            HIPCHK(c, hipMemcpyAsync(h->perm_row, c->d.perm_row, n * ISE_MAX_ROWS, hipMemcpyDeviceToHost, s));
        if (h->perm_col)
            HIPCHK(c, hipMemcpyAsync(h->perm_col, c->d.perm_col, n * ISE_MAX_COLS, hipMemcpyDeviceToHost, s));
        if (h->orbits)
            HIPCHK(c, hipMemcpyAsync(h->orbits, c->d.orbits, n * ISE_MAX_ROWS, hipMemcpyDeviceToHost, s));
        if (h->n_gens) HIPCHK(c, hipMemcpyAsync(h->n_gens, c->d.n_gens, n * 2, hipMemcpyDeviceToHost, s));
        if (h->counters)
            HIPCHK(c, hipMemcpyAsync(h->counters, c->d.counters, n * 16, hipMemcpyDeviceToHost, s));
        if (h->gens && c->last_gen_cap) {
            if (h->gen_cap != c->last_gen_cap) return ISE_ERR_ARG;
            HIPCHK(c, hipMemcpyAsync(h->gens, c->d.gens, n * c->last_gen_cap * ISE_MAX_COLS,
                                     hipMemcpyDeviceToHost, s));
        }
    }
    HIPCHK(c, hipStreamSynchronize(s));
    if (n) HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    c->pending = false;
    return ISE_OK;
}

int ise_test_canonicity(ise_ctx *c, const uint8_t *mat, uint32_t nrows, uint32_t ncols,
                        uint32_t len_stab, uint32_t out_info, uint32_t gen_cap, ise_results *h) {
    if (!c || !mat || !h || nrows < 1 || nrows > c->stride || ncols < 1 || ncols > ISE_MAX_COLS ||
        len_stab >= nrows)
        return ISE_ERR_ARG;
    std::vector<uint64_t> rows(c->stride, 0ull);
    for (uint32_t i = 0; i < nrows; i++)
        for (uint32_t q = 0; q < ncols; q++)
            if (mat[i * ncols + q]) rows[i] |= 1ull << q;
    const uint32_t meta = nrows | (ncols << 8) | (len_stab << 16) | ((out_info & 0xff) << 24);
    int rc = ise_submit(c, 1, c->stride, rows.data(), nullptr, &meta, gen_cap);
    if (rc) return rc;
    return ise_collect(c, h);
}

int ise_check_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                           const uint64_t *d_cells, const uint32_t *d_meta, const ise_results *d_out,
                           uint32_t *d_counter, void *stream) {
    if (!c || !d_out || !d_out->verdict || !d_counter || stride < 1 || stride > ISE_MAX_ROWS)
        return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const int group = c->cfg.group_lanes ? (int)c->cfg.group_lanes : 64;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_defer(c, n);
    if (rc == ISE_OK) rc = ensure_sched(c, n);
    if (rc) return rc;
    return launch(c, n, stride, d_rows, d_cells, d_meta, *d_out, d_counter, s, group);
}

// Work buffers of the batched CanonizeMatrix for up to the context capacity
// (allocated on the first call).
static int ensure_canon(ise_ctx *c) {
    if (c->canon_cap) return ISE_OK;
    const size_t n = c->capacity, words = n * c->stride;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = dalloc(&c->d_canon_rows[i], words);
        if (e == hipSuccess) e = dalloc(&c->d_canon_cells[i], words);
        if (e == hipSuccess) e = dalloc(&c->d_canon_meta[i], n);
        if (e == hipSuccess) e = dalloc(&c->d_canon_idx[i], n);
    }
    if (e == hipSuccess) e = dalloc(&c->d_canon_count, 64);
    if (e == hipSuccess) e = dalloc(&c->d_canon_out, words);
    if (e == hipSuccess) e = dalloc(&c->d_canon_aut, n);
    if (e == hipSuccess) e = dalloc(&c->d_canon_status, n);
    if (e == hipSuccess) e = dalloc(&c->d_canon_iters, n);
    if (e != hipSuccess) return fail(c, e, "ensure_canon");
    c->canon_cap = n;
    return ISE_OK;
}

int ise_canonize_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                              const uint32_t *d_meta, uint64_t *d_canon, uint8_t *d_status, uint64_t *d_aut,
                              uint32_t *d_iters, uint32_t max_iters) {
    return ise_canonize_batch_device_part(c, n, stride, d_rows, d_meta, 0ull, d_canon, d_status, d_aut, d_iters,
                                          max_iters);
}

int ise_canonize_batch_device_part(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                   const uint32_t *d_meta, uint64_t part, uint64_t *d_canon, uint8_t *d_status,
                                   uint64_t *d_aut, uint32_t *d_iters, uint32_t max_iters) {
    if (!c || !d_canon || !d_status || stride < 1 || stride > c->stride || max_iters < 1) return ISE_ERR_ARG;
    if (n > c->capacity) return ISE_ERR_CAPACITY;
    if (n == 0) return ISE_OK;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_canon(c);
    if (rc == ISE_OK) rc = ensure_defer(c, n);
    if (rc == ISE_OK) rc = ensure_sched(c, n);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const uint32_t step_blocks = (uint32_t)c->num_cus * 8u;
    ise::CanonStep a{};
    a.first = 1;
    a.n_cur = n;
    a.stride = stride;
    a.stride_in = stride;
    a.stride_out = stride;
    a.iter = 0;
    a.max_iters = max_iters;
    a.rows_in = d_rows;
    a.meta_in = d_meta;
    a.canon = d_canon;
    a.status = d_status;
    a.aut = d_aut;
    a.iters = d_iters;
    a.part = part;
    a.n_cur_dev = nullptr;
    // the test of every round: verdict, witness row permutation and |Aut| (row action)
    ise_results out{};
    out.verdict = c->d.verdict;
    out.perm_row = c->d.perm_row;
    out.group_order = c->d.group_order;
    // The host reads the number of open matrices after the first round and
    // every CANON_SYNC_ROUNDS rounds after that; in between each round's step
    // reads the exact count on the device and the test runs over the last
    // count read, an upper bound (open matrices only leave), whose slots past
    // the true count hold an empty meta word.  The two batches' counts sit on
    // their own cache lines.
    constexpr int CANON_SYNC_ROUNDS = 4;
    uint32_t bound = n;
    int since_sync = 0;
    int cur = 0;
    for (;;) {
        a.rows_next = c->d_canon_rows[cur];
        a.meta_next = c->d_canon_meta[cur];
        a.idx_next = c->d_canon_idx[cur];
        a.cells_next = part ? c->d_canon_cells[cur] : nullptr;
        a.count_next = c->d_canon_count + 32 * cur;
        a.n_cur = bound;
        HIPCHK(c, ise::canon_step_launch(a, step_blocks, s));
        if (a.first || ++since_sync >= CANON_SYNC_ROUNDS) {
            uint32_t open = 0;
            HIPCHK(c, hipMemcpyAsync(&open, a.count_next, sizeof open, hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
            since_sync = 0;
            bound = open;
            if (open == 0) break;
        }
        // the reference's witness order (CanonizeMatrix keeps the plain order); index
        // order, no deepest-first schedule: the order changes no result, and the
        // batches of the k-Sys canonizer hold one depth
        rc = launch(c, bound, stride, c->d_canon_rows[cur], a.cells_next, c->d_canon_meta[cur], out, c->d_head, s,
                    64, false, false);
        if (rc) return rc;
        a.first = 0;
        a.n_cur_dev = a.count_next;
        a.rows_cur = c->d_canon_rows[cur];
        a.meta_cur = c->d_canon_meta[cur];
        a.idx_cur = c->d_canon_idx[cur];
        a.verdict = out.verdict;
        a.perm_row = out.perm_row;
        a.group_order = out.group_order;
        a.iter++;
        cur ^= 1;
    }
    return ISE_OK;
}

int ise_canonize_batch(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                       uint64_t *canon_out, uint8_t *status_out, uint64_t *aut_out, uint32_t *iters_out,
                       uint32_t max_iters) {
    if (!c || !canon_out || !status_out || (n && (!rows || !meta)) || stride < 1 || stride > c->stride)
        return ISE_ERR_ARG;
    if (n > c->capacity) return ISE_ERR_CAPACITY;
    if (n == 0) return ISE_OK;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_canon(c);
    if (rc) return rc;
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemcpyAsync(c->d_rows, rows, (size_t)n * stride * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->d_meta, meta, (size_t)n * 4, hipMemcpyHostToDevice, s));
    rc = ise_canonize_batch_device(c, n, stride, c->d_rows, c->d_meta, c->d_canon_out, c->d_canon_status,
                                   c->d_canon_aut, c->d_canon_iters, max_iters);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(canon_out, c->d_canon_out, (size_t)n * stride * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(status_out, c->d_canon_status, n, hipMemcpyDeviceToHost, s));
    if (aut_out) HIPCHK(c, hipMemcpyAsync(aut_out, c->d_canon_aut, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (iters_out) HIPCHK(c, hipMemcpyAsync(iters_out, c->d_canon_iters, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return ISE_OK;
}

int ise_canonize_matrix(ise_ctx *c, const uint8_t *mat, uint32_t nrows, uint32_t ncols, uint32_t len_stab,
                        uint8_t *out, uint64_t *aut, uint32_t *iters) {
    if (!c || !mat || !out || nrows < 1 || nrows > c->stride || ncols < 1 || ncols > ISE_MAX_COLS ||
        len_stab >= nrows)
        return ISE_ERR_ARG;
    std::vector<uint64_t> rows(c->stride, 0ull), canon(c->stride, 0ull);
    for (uint32_t i = 0; i < nrows; i++)
        for (uint32_t q = 0; q < ncols; q++)
            if (mat[i * ncols + q]) rows[i] |= 1ull << q;
    const uint32_t meta = nrows | (ncols << 8) | (len_stab << 16);
    uint8_t status = 0;
    uint64_t a = 0;
    uint32_t it = 0;
    int rc = ise_canonize_batch(c, 1, c->stride, rows.data(), &meta, canon.data(), &status, &a, &it,
                                ISE_CANONIZE_MAX_ITERS);
    if (rc) return rc;
    if (status != ISE_VERDICT_CANONICAL) {
        c->err = "ise_canonize_matrix: the matrix did not reach its canonical form";
        return ISE_ERR_STATE;
    }
    for (uint32_t i = 0; i < nrows; i++)
        for (uint32_t q = 0; q < ncols; q++) out[i * ncols + q] = (uint8_t)((canon[i] >> q) & 1ull);
    if (aut) *aut = a;
    if (iters) *iters = it;
    return ISE_OK;
}

int ise_check_batch_device_noalloc(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                   const uint64_t *d_cells, const uint32_t *d_meta, const ise_results *d_out,
                                   void *stream) {
    if (!c || !d_out || !d_out->verdict || stride < 1 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if (n > ISE_NOALLOC_MAX_CANDIDATES) {
        c->err = "ise_check_batch_device_noalloc: " + std::to_string(n) + " candidates exceed the maximum of " +
                 std::to_string(ISE_NOALLOC_MAX_CANDIDATES) + " (ISE_NOALLOC_MAX_CANDIDATES)";
        return ISE_ERR_CAPACITY;
    }
    if (n == 0) return ISE_OK;
    if (!d_rows || !d_meta) return ISE_ERR_ARG;
    if (!c->noalloc_scratch) {
        c->err = "ise_check_batch_device_noalloc: no scratch on this device (" + c->noalloc_err + ")";
        return ISE_ERR_HIP;
    }
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const int group = c->cfg.group_lanes ? (int)c->cfg.group_lanes : 64;
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t blocks = 0;
    int rc = grid_blocks(c, n, stride, group, *d_out, &blocks);
    if (rc) return rc;
    // one pass (no deferral); queue words and schedule list in the scratch array
    uint32_t *head = c->noalloc_scratch;
    uint32_t *sched = c->schedule == ISE_SCHEDULE_DEEPEST_FIRST ? c->noalloc_scratch + ISE_QUEUE_WORDS : nullptr;
    const uint32_t wpb = c->cfg.waves_per_block ? c->cfg.waves_per_block : 4;
    if (c->order_mode == ISE_ORDER_LEAD && !d_cells && !d_out->perm_col) {
        HIPCHK(c, ise::launch_lead(blocks, wpb, n, stride, d_rows, d_meta, *d_out, head,
                                   c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), c->lead_gate, c->lead_frag,
                                   c->lead_w, nullptr, nullptr, nullptr, sched, s));
        return ISE_OK;
    }
    const uint32_t budget = c->order_mode == ISE_ORDER_PLAIN ? 0u : c->hunt_budget;
    const uint8_t *order = c->order_mode == ISE_ORDER_LEARNED ? c->d_order : nullptr;
    HIPCHK(c, ise::launch_check(group, blocks, wpb, n, stride, d_rows, d_cells, d_meta, *d_out, head,
                                c->cfg.max_perms ? c->cfg.max_perms : (1u << 26), budget, order, c->order_dim,
                                nullptr, nullptr, sched, s));
    return ISE_OK;
}

int ise_check_batch_device_timed(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows,
                                 const uint64_t *d_cells, const uint32_t *d_meta,
                                 const ise_results *d_out, uint32_t *d_counter, uint32_t repeats,
                                 float *ms_out) {
    if (!c || !ms_out || repeats == 0) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<hipEvent_t> ev(repeats + 1, nullptr);
    int rc = ISE_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ISE_ERR_HIP;
    if (rc == ISE_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++) {
        rc = ise_check_batch_device(c, n, stride, d_rows, d_cells, d_meta, d_out, d_counter, nullptr);
        if (rc == ISE_OK && hipEventRecord(ev[i + 1], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    }
    if (rc == ISE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++)
        if (hipEventElapsedTime(&ms_out[i], ev[i], ev[i + 1]) != hipSuccess) rc = ISE_ERR_HIP;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == ISE_ERR_HIP && c->err.empty()) c->err = "timed launch failed";
    return rc;
}

int ise_device_alloc(ise_ctx *c, void **d_ptr, uint64_t bytes) {
    if (!c || !d_ptr || bytes == 0) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMalloc(d_ptr, bytes));
    HIPCHK(c, hipMemsetAsync(*d_ptr, 0, bytes, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

int ise_device_free(ise_ctx *c, void *d_ptr) {
    if (!c) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (d_ptr) HIPCHK(c, hipFree(d_ptr));
    return ISE_OK;
}

int ise_copy_to_device(ise_ctx *c, void *d_dst, const void *src, uint64_t bytes) {
    if (!c || !d_dst || !src) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

int ise_copy_to_host(ise_ctx *c, void *dst, const void *d_src, uint64_t bytes) {
    if (!c || !dst || !d_src) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

int ise_copy_on_device(ise_ctx *c, void *d_dst, const void *d_src, uint64_t bytes) {
    if (!c || !d_dst || !d_src) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

static ise::KfOut kf_out(const ise_kf_results &r) {
    ise::KfOut o;
    o.verdict = r.verdict; o.witness = r.witness; o.n_auts = r.n_auts; o.n_tried = r.n_tried;
    o.player_index = r.player_index;
    return o;
}

static int kf_launch_on(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                        const uint32_t *d_meta, const ise_kf_results &d_out, uint32_t *d_queue,
                        hipStream_t s) {
    // one wave per workgroup; the range pass fills the chip whatever n is
    uint32_t range_blocks = c->cfg.blocks;
    if (!range_blocks) {
        int per_cu = 0;
        HIPCHK(c, ise::kf_occupancy(&per_cu));
        range_blocks = (uint32_t)(per_cu < 1 ? 1 : per_cu) * (uint32_t)c->num_cus;
    }
    const uint32_t blocks = c->cfg.blocks ? c->cfg.blocks : (range_blocks > n ? (n ? n : 1) : range_blocks);
    const uint32_t limit = c->kf_first_chunks * 64u;
    if (limit) {
        const int rc = ensure_kf_split(c, n);
        if (rc) return rc;
    }
    const size_t cap = c->kf_cap;
    const ise::KfSplit sp{{c->d_kf_list, c->d_kf_list ? c->d_kf_list + cap : nullptr},
                          {c->d_kf_part, c->d_kf_part ? c->d_kf_part + cap : nullptr},
                          c->d_kf_best, c->d_kf_ranges, (uint32_t)c->kf_range_words, c->kf_max_ranges};
    HIPCHK(c, ise::kf_launch(blocks, range_blocks, n, stride, d_rounds, d_meta, kf_out(d_out), d_queue,
                             c->kf_tab, limit, sp, s));
    HIPCHK(c, ise::kf_backjump_launch(0, n, stride, d_rounds, d_meta, kf_out(d_out), c->kf_tab, s));
    return ISE_OK;
}

int ise_kf_check_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                              const uint32_t *d_meta, const ise_kf_results *d_out, uint32_t *d_queue,
                              void *stream) {
    if (!c || !d_out || !d_out->verdict || !d_queue || stride < 1 || stride > ISE_MAX_ROWS)
        return ISE_ERR_ARG;
    if (d_out->player_index && !d_out->witness) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rounds || !d_meta) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    return kf_launch_on(c, n, stride, d_rounds, d_meta, *d_out, d_queue,
                        stream ? (hipStream_t)stream : c->stream);
}

int ise_kf_check_batch_device_timed(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                    const uint32_t *d_meta, const ise_kf_results *d_out,
                                    uint32_t *d_queue, uint32_t repeats, float *ms_out) {
    if (!c || !ms_out || repeats == 0) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<hipEvent_t> ev(repeats + 1, nullptr);
    int rc = ISE_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ISE_ERR_HIP;
    if (rc == ISE_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++) {
        rc = ise_kf_check_batch_device(c, n, stride, d_rounds, d_meta, d_out, d_queue, nullptr);
        if (rc == ISE_OK && hipEventRecord(ev[i + 1], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    }
    if (rc == ISE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++)
        if (hipEventElapsedTime(&ms_out[i], ev[i], ev[i + 1]) != hipSuccess) rc = ISE_ERR_HIP;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == ISE_ERR_HIP && c->err.empty()) c->err = "timed launch failed";
    return rc;
}

// The two-row canonizers: perfect (p1f) or uniform 1-factorizations.
enum KfMode { KF_GENERAL = 0, KF_2P1F = 1, KF_2U1F = 2, KF_3U1F = 3 };

// cnvCheck3U1F's repetitions of a pair's relabellings for groups of 2 (its loop
// over MAX_CYCLE_SETS cycle sets, k-Factorization/TripleSys/Include/Global.h:12)
constexpr uint32_t KF_3U1F_REPS = 32;

// The cycle types a 3U1F run allows ([n_types][4] cycle lengths, ascending,
// zero after the last) in the kernel's form; false for a malformed list.
static bool kf_u1f_mode(const uint8_t *u1f_cycles, uint32_t n_types, ise::KfU1fMode *md) {
    *md = ise::KfU1fMode{KF_3U1F_REPS, n_types, {0u, 0u, 0u, 0u}};
    if (n_types > ISE_KF_MAX_U1F_TYPES || (n_types && !u1f_cycles)) return false;
    for (uint32_t k = 0; k < n_types; k++) {
        uint32_t nc = 0, halves = 0, prev = 0;
        for (uint32_t c = 0; c < ISE_KF_MAX_U1F_CYCLES; c++) {
            const uint32_t len = u1f_cycles[k * ISE_KF_MAX_U1F_CYCLES + c];
            if (!len) {
                for (uint32_t e = c; e < ISE_KF_MAX_U1F_CYCLES; e++)
                    if (u1f_cycles[k * ISE_KF_MAX_U1F_CYCLES + e]) return false;
                break;
            }
            if (len < 4 || (len & 1u) || len > 2u * 15u || len < prev) return false;
            halves |= (len / 2u) << (4 * nc++);
            prev = len;
        }
        if (!nc) return false;
        md->types[k] = (nc << 16) | halves;
    }
    return true;
}

static int kf2_launch_on(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                         const uint32_t *d_meta, const ise_kf_results &d_out, hipStream_t s, int mode,
                         const ise::KfU1fMode *md3 = nullptr) {
    uint32_t blocks = c->cfg.blocks;
    if (!blocks) {
        int per_cu = 0;
        HIPCHK(c, ise::kf_occupancy(&per_cu));
        blocks = (uint32_t)(per_cu < 1 ? 1 : per_cu) * (uint32_t)c->num_cus;
        if (blocks > n) blocks = n ? n : 1;
    }
    if (mode == KF_2U1F || mode == KF_3U1F) {
        const ise::KfU1fMode one{1u, 0u, {0u, 0u, 0u, 0u}};
        HIPCHK(c, ise::kf_u1f_launch(blocks, n, stride, d_rounds, d_meta, kf_out(d_out),
                                     mode == KF_3U1F ? *md3 : one, s));
    } else {
        HIPCHK(c, ise::kf2_launch(blocks, n, stride, d_rounds, d_meta, kf_out(d_out), s));
    }
    if (mode != KF_3U1F)
        HIPCHK(c, ise::kf_backjump_launch(mode == KF_2U1F ? 2 : 1, n, stride, d_rounds, d_meta, kf_out(d_out),
                                          c->kf_tab, s));
    return ISE_OK;
}

int ise_kf_check_2p1f_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const ise_kf_results *d_out, void *stream) {
    if (!c || !d_out || !d_out->verdict || stride < 1 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if (d_out->player_index && !d_out->witness) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rounds || !d_meta) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    return kf2_launch_on(c, n, stride, d_rounds, d_meta, *d_out, stream ? (hipStream_t)stream : c->stream, KF_2P1F);
}

int ise_kf_check_2u1f_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const ise_kf_results *d_out, void *stream) {
    if (!c || !d_out || !d_out->verdict || stride < 1 || stride > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if (d_out->player_index && !d_out->witness) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rounds || !d_meta) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    return kf2_launch_on(c, n, stride, d_rounds, d_meta, *d_out, stream ? (hipStream_t)stream : c->stream, KF_2U1F);
}

int ise_kf_check_3u1f_batch_device(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                   const uint32_t *d_meta, const uint8_t *u1f_cycles, uint32_t n_types,
                                   const ise_kf_results *d_out, void *stream) {
    if (!c || !d_out || !d_out->verdict || d_out->player_index || stride < 1 || stride > ISE_MAX_ROWS)
        return ISE_ERR_ARG;
    ise::KfU1fMode md;
    if (!kf_u1f_mode(u1f_cycles, n_types, &md)) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_rounds || !d_meta) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    return kf2_launch_on(c, n, stride, d_rounds, d_meta, *d_out, stream ? (hipStream_t)stream : c->stream, KF_3U1F,
                         &md);
}

// `repeats` launches of `launch_one` on the context stream, each timed with HIP events.
extern "C++" template <typename F> static int kf_timed(ise_ctx *c, uint32_t repeats, float *ms_out, F launch_one) {
    if (!c || !ms_out || repeats == 0) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<hipEvent_t> ev(repeats + 1, nullptr);
    int rc = ISE_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ISE_ERR_HIP;
    if (rc == ISE_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++) {
        rc = launch_one();
        if (rc == ISE_OK && hipEventRecord(ev[i + 1], c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    }
    if (rc == ISE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = ISE_ERR_HIP;
    for (uint32_t i = 0; i < repeats && rc == ISE_OK; i++)
        if (hipEventElapsedTime(&ms_out[i], ev[i], ev[i + 1]) != hipSuccess) rc = ISE_ERR_HIP;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == ISE_ERR_HIP && c->err.empty()) c->err = "timed launch failed";
    return rc;
}

int ise_kf_check_2p1f_batch_device_timed(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                         const uint32_t *d_meta, const ise_kf_results *d_out,
                                         uint32_t repeats, float *ms_out) {
    return kf_timed(c, repeats, ms_out, [&] {
        return ise_kf_check_2p1f_batch_device(c, n, stride, d_rounds, d_meta, d_out, nullptr);
    });
}

int ise_kf_check_2u1f_batch_device_timed(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rounds,
                                         const uint32_t *d_meta, const ise_kf_results *d_out,
                                         uint32_t repeats, float *ms_out) {
    return kf_timed(c, repeats, ms_out, [&] {
        return ise_kf_check_2u1f_batch_device(c, n, stride, d_rounds, d_meta, d_out, nullptr);
    });
}

// Host records of 1-factorizations: edge-disjoint perfect matchings of the
// stated size; for the two-row canonizer also rounds in normal form (pairs
// (a, b) with a < b, pairs by first player) and every pair of rounds one
// Hamiltonian cycle.
static bool kf_records_ok(uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta, int mode) {
    const bool p1f = mode == KF_2P1F, normal = mode != KF_GENERAL;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t nr = meta[i] & 0xff, np = (meta[i] >> 8) & 0xff;
        if (np < 4 || np > ise::KF_MAX_PLAYERS || (np & 1) || nr < 1 || nr > np - 1 || nr > stride)
            return false;
        uint32_t used[ise::KF_MAX_PLAYERS] = {0};
        uint8_t part[ise::KF_MAX_ROUNDS][ise::KF_MAX_PLAYERS];
        for (uint32_t r = 0; r < nr; r++) {
            const uint64_t w = rounds[(uint64_t)i * stride + r];
            uint32_t seen = 0, prev = 0;
            for (uint32_t p = 0; p < np; p += 2) {
                const uint32_t x = (uint32_t)(w >> (4 * p)) & 0xf, y = (uint32_t)(w >> (4 * p + 4)) & 0xf;
                if (x >= np || y >= np || x == y || (seen >> x & 1) || (seen >> y & 1) || (used[x] >> y & 1))
                    return false;
                if (normal && (x > y || (p && x < prev))) return false;
                prev = x;
                seen |= (1u << x) | (1u << y);
                used[x] |= 1u << y;
                used[y] |= 1u << x;
                part[r][x] = (uint8_t)y;
                part[r][y] = (uint8_t)x;
            }
        }
        if (p1f)
            for (uint32_t a = 0; a < nr; a++)
                for (uint32_t b = a + 1; b < nr; b++) {
                    uint32_t v = 0, len = 0;
                    do {
                        v = part[b][part[a][v]];
                        len += 2;
                    } while (v != 0 && len <= np);
                    if (len != np) return false;
                }
    }
    return true;
}

static int kf_host_run(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta,
                       ise_kf_results *h, int mode, const ise::KfU1fMode *md3 = nullptr) {
    if (!c || !h || !h->verdict || (n && (!rounds || !meta)) || stride < 1 || stride > c->stride)
        return ISE_ERR_ARG;
    if (n > c->capacity) return ISE_ERR_CAPACITY;
    if (n == 0) return ISE_OK;
    if (!kf_records_ok(n, stride, rounds, meta, mode)) return ISE_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_rows, rounds, (size_t)n * stride * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_meta, meta, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    // the context's result buffers; the backjump position goes to the fail_row
    // array reinterpreted (one uint16 per candidate fits in perm_row's 32 bytes)
    ise_kf_results d{c->d.verdict, c->d_kf_witness, c->d_kf_auts, c->d_kf_tried,
                     h->player_index ? (uint16_t *)c->d.perm_row : nullptr};
    const int rc = mode != KF_GENERAL ? kf2_launch_on(c, n, stride, c->d_rows, c->d_meta, d, c->stream, mode, md3)
                                      : kf_launch_on(c, n, stride, c->d_rows, c->d_meta, d, c->d_head, c->stream);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(h->verdict, d.verdict, n, hipMemcpyDeviceToHost, c->stream));
    if (h->witness) HIPCHK(c, hipMemcpyAsync(h->witness, d.witness, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (h->n_auts) HIPCHK(c, hipMemcpyAsync(h->n_auts, d.n_auts, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (h->n_tried) HIPCHK(c, hipMemcpyAsync(h->n_tried, d.n_tried, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (h->player_index)
        HIPCHK(c, hipMemcpyAsync(h->player_index, d.player_index, n * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

int ise_kf_check(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta,
                 ise_kf_results *h) {
    return kf_host_run(c, n, stride, rounds, meta, h, KF_GENERAL);
}

int ise_kf_check_2p1f(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta,
                      ise_kf_results *h) {
    return kf_host_run(c, n, stride, rounds, meta, h, KF_2P1F);
}

int ise_kf_check_2u1f(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta,
                      ise_kf_results *h) {
    return kf_host_run(c, n, stride, rounds, meta, h, KF_2U1F);
}

int ise_kf_check_3u1f(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rounds, const uint32_t *meta,
                      const uint8_t *u1f_cycles, uint32_t n_types, ise_kf_results *h) {
    ise::KfU1fMode md;
    if (!h || h->player_index || !kf_u1f_mode(u1f_cycles, n_types, &md)) return ISE_ERR_ARG;
    return kf_host_run(c, n, stride, rounds, meta, h, KF_3U1F, &md);
}

int ise_set_order(ise_ctx *c, int mode, uint32_t hunt_budget, const uint8_t *d_order, uint32_t order_dim) {
    if (!c || mode < ISE_ORDER_PLAIN || mode > ISE_ORDER_LEARNED) return ISE_ERR_ARG;
    if (mode != ISE_ORDER_PLAIN && hunt_budget == 0) return ISE_ERR_ARG;
    if (mode == ISE_ORDER_LEARNED && (!d_order || order_dim < 1 || order_dim > ISE_MAX_ROWS)) return ISE_ERR_ARG;
    c->order_mode = mode;
    c->hunt_budget = hunt_budget;
    c->d_order = mode == ISE_ORDER_LEARNED ? d_order : nullptr;
    c->order_dim = mode == ISE_ORDER_LEARNED ? order_dim : 0;
    return ISE_OK;
}

int ise_set_defer(ise_ctx *c, uint32_t first_pass_perms, uint32_t second_pass_perms) {
    if (!c) return ISE_ERR_ARG;
    if (second_pass_perms && second_pass_perms <= first_pass_perms) return ISE_ERR_ARG;
    if (int rc = drop_graphs(c)) return rc;
    c->defer_perms[0] = first_pass_perms;
    c->defer_perms[1] = first_pass_perms ? second_pass_perms : 0u;
    return ensure_defer(c, c->capacity);
}

int ise_set_schedule(ise_ctx *c, int schedule) {
    if (!c || (schedule != ISE_SCHEDULE_QUEUE && schedule != ISE_SCHEDULE_DEEPEST_FIRST)) return ISE_ERR_ARG;
    if (int rc = drop_graphs(c)) return rc;
    c->schedule = schedule;
    return ensure_sched(c, c->capacity);
}

int ise_kf_set_split(ise_ctx *c, uint32_t first_chunks, uint32_t max_ranges) {
    if (!c || first_chunks > (1u << 20) || (first_chunks && (max_ranges < 1 || max_ranges > 65536)))
        return ISE_ERR_ARG;
    c->kf_first_chunks = first_chunks;
    c->kf_max_ranges = first_chunks ? max_ranges : c->kf_max_ranges;
    return ensure_kf_split(c, c->capacity);
}

int ise_report_columns(ise_ctx *c, int on) {
    if (!c) return ISE_ERR_ARG;
    c->report_columns = on != 0;
    return ISE_OK;
}

int ise_use_graphs(ise_ctx *c, int on) {
    if (!c) return ISE_ERR_ARG;
    c->use_graphs = on != 0;
    return ISE_OK;
}

int ise_graph_stats(const ise_ctx *c, uint32_t *graphs, uint64_t *replays) {
    if (!c) return ISE_ERR_ARG;
    if (graphs) *graphs = (uint32_t)c->graphs.size();
    if (replays) *replays = c->graph_replays;
    return ISE_OK;
}

int ise_graph_updates(const ise_ctx *c, uint64_t *updates) {
    if (!c || !updates) return ISE_ERR_ARG;
    *updates = c->graph_updates;
    return ISE_OK;
}

int ise_autotune(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *d_rows, const uint64_t *d_cells,
                 const uint32_t *d_meta, const ise_results *d_out, uint32_t *d_counter, uint32_t max_depth,
                 uint32_t max_cols, uint32_t repeats, float *table, ise_launch_cfg *best) {
    if (!c || !table || !d_out || n == 0 || repeats == 0 || max_depth < 1 || max_cols < 1 ||
        max_depth > stride || max_depth > ISE_MAX_ROWS || max_cols > ISE_MAX_COLS)
        return ISE_ERR_ARG;
    const uint32_t groups[3] = {16, 32, 64}, waves[4] = {1, 2, 4, 8};
    const ise_launch_cfg saved = c->cfg;
    ise_launch_cfg win = saved;
    float win_ms = -1.f;
    std::vector<float> ms(repeats);
    for (int gi = 0; gi < 3; gi++) {
        for (int wi = 0; wi < 4; wi++) {
            table[gi * 4 + wi] = -1.f;
            if (groups[gi] < max_depth || groups[gi] < max_cols) continue;
            c->cfg.group_lanes = groups[gi];
            c->cfg.waves_per_block = waves[wi];
            c->cfg.blocks = 0;
            int rc = ise_check_batch_device_timed(c, n, stride, d_rows, d_cells, d_meta, d_out, d_counter,
                                                  repeats, ms.data());
            if (rc) {
                c->cfg = saved;
                return rc;
            }
            float m = ms[0];
            for (float x : ms) m = x < m ? x : m;     // best of the repeats
            table[gi * 4 + wi] = m;
            if (win_ms < 0 || m < win_ms) {
                win_ms = m;
                win = c->cfg;
            }
        }
    }
    c->cfg = win;
    if (best) *best = win;
    return ISE_OK;
}

int ise_sync(ise_ctx *c) {
    if (!c) return ISE_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISE_OK;
}

const char *ise_last_error(const ise_ctx *c) { return c ? c->err.c_str() : "no context"; }

// Kernel time of the last collected submit, milliseconds (0 if none).
float ise_last_kernel_ms(const ise_ctx *c) { return c ? c->last_ms : 0.f; }

} // extern "C"
