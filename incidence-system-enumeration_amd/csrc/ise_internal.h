// ise_internal.h -- declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise.h"

namespace ise {

// Enqueue the queue-head reset and the canonicity kernel on `stream`.
hipError_t launch_check(int group_lanes, uint32_t blocks, uint32_t waves_per_block,
                        uint32_t n, uint32_t stride, const uint64_t *rows,
                        const uint64_t *cells, const uint32_t *meta,
                        const ise_results &out, uint32_t *head, uint32_t max_perms,
                        uint32_t hunt_budget, const uint8_t *order, uint32_t order_dim,
                        uint32_t *defer, const uint32_t *defer_budgets, uint32_t *sched,
                        hipStream_t stream, int pass_lo = 0, int pass_hi = 0);

// One pass of a gated launch on HBM-resident inputs (the learned order's
// batching): pass 1 runs the plain search on every candidate with a budget of
// `gate` permutations and lists those that exceed it; pass 2 runs the listed
// ones from the start, with the witness hunt in `order` when hunt_budget > 0.
// *list receives the list (its count: defer_count_ptr(head, 1)).
int launch_gated_pass(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                      const ise_results &out, uint32_t *head, uint32_t gate, int pass, uint32_t hunt_budget,
                      const uint8_t *order, uint32_t order_dim, const uint32_t **list);

// The learned order of the first row (ise_lead.h) on a batch: queue reset,
// schedule, then one wave per candidate.  frag/w: the network (A fragments,
// packed fp32 parameters); lead_in (optional [n][16]) replaces its order;
// lead_out / score_out (optional [n][16]) receive the order and scores used.
hipError_t launch_lead(uint32_t blocks, uint32_t waves_per_block, uint32_t n, uint32_t stride, const uint64_t *rows,
                       const uint32_t *meta, const ise_results &out, uint32_t *head, uint32_t max_perms, uint32_t gate,
                       const uint16_t *frag, const float *w, const uint8_t *lead_in, uint8_t *lead_out,
                       float *score_out, uint32_t *sched, hipStream_t stream);
hipError_t lead_occupancy(uint32_t waves_per_block, uint32_t stride, int *blocks_per_cu);
// The same with the context's launch shape, schedule and permutation limit
// (on its stream).
int launch_lead_ctx(ise_ctx *c, uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                    const ise_results &out, uint32_t *head, uint32_t gate, const uint16_t *frag, const float *w,
                    const uint8_t *lead_in, uint8_t *lead_out, float *score_out);

// Word of the queue buffer that counts the candidates pass `pass` deferred.
uint32_t *defer_count_ptr(uint32_t *head, int pass);

// Words of the schedule buffer beyond the n list entries (depth counters).
constexpr uint32_t SCHED_EXTRA_WORDS = 128;

// Passes of a deferred launch: up to two budgets below the limit, then the limit.
constexpr int DEFER_BUDGETS = 2;

// Resident workgroups per CU for the kernel variant.
hipError_t occupancy_blocks(int group_lanes, uint32_t waves_per_block, uint32_t stride, bool sets,
                            int *blocks_per_cu);

// Dynamic LDS of one workgroup (orbit geometry, or column sets, of every
// group's candidate).
size_t lds_bytes(int group_lanes, uint32_t waves_per_block, uint32_t stride, bool sets);

// Whether a launch with these outputs runs the column-set search (no column
// permutation reported, no witness hunt).
bool uses_sets(const ise_results &out, uint32_t hunt_budget);

// Smallest lane-group size that holds every candidate of a batch.
inline int fit_group_lanes(uint32_t max_depth, uint32_t max_cols) {
    const uint32_t need = max_depth > max_cols ? max_depth : max_cols;
    return need <= 16 ? 16 : need <= 32 ? 32 : 64;
}

// Geometry of a context (rows per candidate, candidates per submit).
uint32_t ctx_stride(const ise_ctx *c);
uint32_t ctx_capacity(const ise_ctx *c);
hipStream_t ctx_stream(const ise_ctx *c);
int ctx_device(const ise_ctx *c);
int ctx_num_cus(const ise_ctx *c);
void ctx_set_error(ise_ctx *c, const char *msg);
// Set whether ise_submit reports column permutations; returns the previous setting.
bool ctx_report_columns(ise_ctx *c, bool on);
void ctx_get_order(const ise_ctx *c, int *mode, uint32_t *budget, const uint8_t **order, uint32_t *dim);
// The lead order on every launch of the context (frag null: back to the plain
// order); cached graphs are dropped.  The network's buffers must outlive it.
int ctx_set_lead(ise_ctx *c, const uint16_t *frag, const float *w, uint32_t gate);
// The fragments the context's lead order uses (null when it is off).
const uint16_t *ctx_lead_frag(const ise_ctx *c);
// A context's search order, saved and restored around temporary changes.
struct OrderState {
    int mode;
    uint32_t budget;
    const uint8_t *order;
    uint32_t dim;
    const uint16_t *lead_frag;
    const float *lead_w;
    uint32_t lead_gate;
};
OrderState ctx_save_order(const ise_ctx *c);
int ctx_restore_order(ise_ctx *c, const OrderState &st);

} // namespace ise
