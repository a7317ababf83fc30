// ise_kernels.hip -- persistent canonicity kernels for gfx950.
//
// A grid of persistent workgroups drains a device work queue: the first lane
// of every G-lane group takes the next candidate index with one atomic add on
// a queue head in HBM, the group runs the full canonicity test of that
// candidate in its lanes (ise_wave.h) and writes one result record.  Records
// with an invalid shape are answered with ISE_VERDICT_INVALID instead of being
// read past their bounds.
#include <hip/hip_runtime.h>
#include "ise_wave.h"
#include "ise_lead.h"
#include "ise_internal.h"

namespace ise {

// The candidate range is cut into one contiguous slice per XCD (workgroups
// are dispatched to the XCDs round-robin, so workgroup w runs on XCD w % 8).
// Each slice has its own queue head on a separate 128-byte line; a group
// drains its home slice first and then takes from the others in turn, so
// neighbouring result records are written from one L2 and the queue atomics
// stay local until the tail of the batch.
constexpr int NUM_XCD = 8;
constexpr int HEAD_STRIDE = 32;   // words between queue heads
// Long searches are deferred: a first pass runs every candidate with a small
// permutation budget and lists those that exceed it; a later pass runs the
// listed ones from the start with a larger budget (listing again those that
// exceed it), and the last pass with the full budget.  The long searches of
// a batch thus start together instead of wherever they sat in the queue.
// Words of the queue buffer that hold, per pass p >= 1, the number of
// candidates it listed and the queue head of the pass that reads that list.
constexpr int DEFER_PASSES = 3;
__host__ __device__ constexpr int defer_count_word(int p) { return 16 + 64 * (p - 1); }
__host__ __device__ constexpr int defer_head_word(int p) { return 48 + 64 * (p - 1); }

__global__ void queue_reset_kernel(uint32_t *head) {
    const int t = (int)threadIdx.x;
    if (t < NUM_XCD) head[t * HEAD_STRIDE] = 0u;
    if (t >= NUM_XCD && t < NUM_XCD + DEFER_PASSES - 1) {
        head[defer_count_word(t - NUM_XCD + 1)] = 0u;
        head[defer_head_word(t - NUM_XCD + 1)] = 0u;
    }
}

template <int G>
__device__ __forceinline__ uint32_t next_candidate(const Grp<G> &g, uint32_t n, uint32_t *head, int &slice,
                                                   int home) {
    while (slice < home + NUM_XCD) {
        const int x = slice % NUM_XCD;
        const uint32_t lo = (uint32_t)(((uint64_t)n * x) / NUM_XCD);
        const uint32_t hi = (uint32_t)(((uint64_t)n * (x + 1)) / NUM_XCD);
        int taken = 0;
        if (g.ln == 0)
            taken = (int)atomicAdd(head + x * HEAD_STRIDE, 1u);
        const uint32_t k = (uint32_t)g.read(taken, 0);
        if (lo + k < hi)
            return lo + k;
        slice++;
    }
    return n;
}

// The same queue over a schedule (a list of candidate indices, deepest rows
// first): slice x holds list positions x, x+8, x+16, ... so that every XCD
// starts on the front of the schedule.  Returns a list position or n.
template <int G>
__device__ __forceinline__ uint32_t next_scheduled(const Grp<G> &g, uint32_t n, uint32_t *head, int &slice,
                                                   int home) {
    while (slice < home + NUM_XCD) {
        const int x = slice % NUM_XCD;
        int taken = 0;
        if (g.ln == 0)
            taken = (int)atomicAdd(head + x * HEAD_STRIDE, 1u);
        const uint64_t p = (uint64_t)(uint32_t)g.read(taken, 0) * NUM_XCD + (uint64_t)x;
        if (p < n)
            return (uint32_t)p;
        slice++;
    }
    return n;
}

// ---- schedule: candidate indices ordered by decreasing depth ---------------
// A search's length grows with the number of rows to permute, so deeper
// candidates are started first (longest-processing-time-first on the depth).
// aux words [0, 33): candidates per depth; [64, 97): fill cursors.
constexpr int SCHED_AUX_WORDS = 128;

__global__ void sched_reset_kernel(uint32_t *aux) {
    for (int t = (int)threadIdx.x; t < SCHED_AUX_WORDS; t += (int)blockDim.x) aux[t] = 0u;
}

__global__ __launch_bounds__(256) void sched_count_kernel(uint32_t n, const uint32_t *__restrict__ meta,
                                                          uint32_t *aux) {
    __shared__ uint32_t h[33];
    if (threadIdx.x < 33) h[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t d = meta[i] & 0xffu;
        atomicAdd(&h[d < 32u ? d : 32u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 33 && h[threadIdx.x]) atomicAdd(&aux[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void sched_fill_kernel(uint32_t n, const uint32_t *__restrict__ meta,
                                                         uint32_t *aux, uint32_t *list) {
    // per tile of 256 candidates: rank inside the block by LDS atomics, then
    // one global cursor atomic per depth present in the tile
    __shared__ uint32_t base[33], tile_cnt[33], tile_off[33];
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int d = 32; d >= 0; d--) { base[d] = acc; acc += aux[d]; }
    }
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        if (threadIdx.x < 33) tile_cnt[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t i = i0 + threadIdx.x;
        uint32_t d = 0, rank = 0;
        if (i < n) {
            const uint32_t m = meta[i] & 0xffu;
            d = m < 32u ? m : 32u;
            rank = atomicAdd(&tile_cnt[d], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 33 && tile_cnt[threadIdx.x])
            tile_off[threadIdx.x] = atomicAdd(&aux[64 + threadIdx.x], tile_cnt[threadIdx.x]);
        __syncthreads();
        if (i < n) list[base[d] + tile_off[d] + rank] = i;
        __syncthreads();
    }
}

// Waves per SIMD the register allocation of the canonicity kernel must leave
// room for (0: the compiler's choice).
#ifndef ISE_CHECK_WAVES_PER_EU
#define ISE_CHECK_WAVES_PER_EU 0
#endif
#if ISE_CHECK_WAVES_PER_EU > 0
#define ISE_CHECK_OCCUPANCY __attribute__((amdgpu_waves_per_eu(ISE_CHECK_WAVES_PER_EU)))
#else
#define ISE_CHECK_OCCUPANCY
#endif

// 32-bit words of LDS scratch per lane group: orbit geometry (one word per
// row and lane) or, for the column-set search, one 64-bit set per row and
// lane.
template <int G, bool SETS> __host__ __device__ constexpr uint32_t group_scratch_words(uint32_t rows_cap) {
    return SETS ? rows_cap * G * 2u : rows_cap * G;
}

template <int G, bool COUNT, bool HUNT, bool SETS>
__global__ __launch_bounds__(1024) ISE_CHECK_OCCUPANCY void check_kernel(uint32_t n, uint32_t stride,
                                                    const uint64_t *__restrict__ rows,
                                                    const uint64_t *__restrict__ cells,
                                                    const uint32_t *__restrict__ meta,
                                                    OutPtrs o, uint32_t *head, uint32_t max_perms,
                                                    uint32_t rows_cap, const uint8_t *__restrict__ order,
                                                    int od, uint32_t hunt_budget,
                                                    const uint32_t *__restrict__ in_list, uint32_t *out_list,
                                                    uint32_t budget, int pass,
                                                    const uint32_t *__restrict__ sched) {
    extern __shared__ uint32_t lds_geo[];
    Grp<G> g;
    // this group's orbit-geometry scratch: rows_cap rows of G words
    uint32_t *geo = lds_geo + ((threadIdx.x >> 6) * (64 / G) + (g.base / G)) * group_scratch_words<G, SETS>(rows_cap);
    const int home = (int)(blockIdx.x % NUM_XCD);
    int slice = home;
    // pass 0 or 1: the whole batch from the queue; pass p >= 2: the list of
    // pass p-1.  With out_list, searches past `budget` are listed for the
    // next pass instead of answered.
    const uint32_t listed = in_list ? head[defer_count_word(pass - 1)] : 0u;
    for (;;) {
        uint32_t ci;
        if (in_list) {
            int taken = 0;
            if (g.ln == 0) taken = (int)atomicAdd(head + defer_head_word(pass - 1), 1u);
            const uint32_t k = (uint32_t)g.read(taken, 0);
            if (k >= listed)
                break;
            ci = in_list[k];
        } else if (sched) {
            const uint32_t p = next_scheduled(g, n, head, slice, home);
            ci = p < n ? sched[p] : n;
        } else {
            ci = next_candidate(g, n, head, slice, home);
        }
        if (ci >= n)
            break;
        const uint32_t m = meta[ci];
        const int d = m & 0xff, b = (m >> 8) & 0xff, fixed = (m >> 16) & 0xff;
        const bool ok = d >= 1 && d <= G && d <= ISE_MAX_ROWS && (uint32_t)d <= stride &&
                        (uint32_t)d <= rows_cap &&
                        b >= 1 && b <= G && fixed < d;
        if (!ok) {
            if (g.ln == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        check_candidate<G, COUNT, HUNT, SETS>(g, ci, m, rows, cells, stride, o, budget, geo, order, od,
                                              hunt_budget, out_list,
                                              out_list ? head + defer_count_word(pass) : nullptr);
    }
}

// Learned order of the first row (ise_lead.h): one candidate per wave from
// the (scheduled) queue; candidates the lead order does not take run
// check_candidate's column-set or replayed search.
template <bool COUNT>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(ISE_LEAD_WAVES_PER_EU))) void lead_kernel(
    uint32_t n, uint32_t stride, const uint64_t *__restrict__ rows, const uint32_t *__restrict__ meta, OutPtrs o,
    uint32_t *head, uint32_t max_perms, uint32_t rows_cap, uint32_t gate, const bf16x4 *__restrict__ frag,
    const float *__restrict__ w, const uint8_t *__restrict__ lead_in, uint8_t *lead_out, float *score_out,
    const uint32_t *__restrict__ sched) {
    extern __shared__ uint32_t lds_geo[];
    Grp<64> g;
    uint32_t *lds = lds_geo + (threadIdx.x >> 6) * group_scratch_words<64, true>(rows_cap);
    uint64_t *stk = reinterpret_cast<uint64_t *>(lds);
    const int home = (int)(blockIdx.x % NUM_XCD);
    int slice = home;
    for (;;) {
        uint32_t ci;
        if (sched) {
            const uint32_t p = next_scheduled(g, n, head, slice, home);
            ci = p < n ? sched[p] : n;
        } else {
            ci = next_candidate(g, n, head, slice, home);
        }
        if (ci >= n)
            break;
        const uint32_t m = meta[ci];
        const int d = m & 0xff, b = (m >> 8) & 0xff, fixed = (m >> 16) & 0xff;
        const bool ok = d >= 1 && d <= ISE_MAX_ROWS && (uint32_t)d <= stride && (uint32_t)d <= rows_cap &&
                        b >= 1 && b <= 64 && fixed < d;
        if (!ok) {
            if (g.ln == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        const Staged c = load_candidate(g, ci, m, rows, nullptr, stride);
        if (d >= 2 && d <= LEAD_MAX_ROWS && fixed == 0 && b <= PACKED_COLS && (frag || lead_in) &&
            orbit_form(g, c, nullptr))
            lead_candidate<COUNT>(g, ci, c, o, max_perms, gate, stk, frag, w, lead_in, lead_out, score_out);
        else
            check_candidate<64, COUNT, false, true>(g, ci, m, rows, nullptr, stride, o, max_perms, lds, nullptr, 0,
                                                    0u, nullptr, nullptr);
    }
}

static OutPtrs to_out(const ise_results &r) {
    OutPtrs o;
    o.verdict = r.verdict; o.fail_row = r.fail_row; o.fail_col = r.fail_col;
    o.row_to_change = r.row_to_change; o.group_order = r.group_order;
    o.group_order_hi = r.group_order_hi;
    o.perm_row = r.perm_row; o.perm_col = r.perm_col; o.orbits = r.orbits;
    o.n_gens = r.n_gens; o.gens = r.gens; o.counters = r.counters;
    o.gen_cap = r.gens ? r.gen_cap : 0u;
    return o;
}

template <int G, bool SETS> static const void *kernel_ptr() {
    return (const void *)&check_kernel<G, true, false, SETS>;
}

// Rows of orbit geometry kept per candidate in LDS.
static uint32_t rows_cap_for(int group_lanes, uint32_t stride) {
    uint32_t cap = stride < (uint32_t)ISE_MAX_ROWS ? stride : (uint32_t)ISE_MAX_ROWS;
    return cap < (uint32_t)group_lanes ? cap : (uint32_t)group_lanes;
}

size_t lds_bytes(int group_lanes, uint32_t waves_per_block, uint32_t stride, bool sets) {
    const uint32_t cap = rows_cap_for(group_lanes, stride);
    const uint32_t groups = waves_per_block * (64u / (uint32_t)group_lanes);
    uint32_t words = 0;
    switch (group_lanes) {
    case 16: words = sets ? group_scratch_words<16, true>(cap) : group_scratch_words<16, false>(cap); break;
    case 32: words = sets ? group_scratch_words<32, true>(cap) : group_scratch_words<32, false>(cap); break;
    default: words = sets ? group_scratch_words<64, true>(cap) : group_scratch_words<64, false>(cap); break;
    }
    return (size_t)groups * words * sizeof(uint32_t);
}

bool uses_sets(const ise_results &out, uint32_t hunt_budget) {
    (void)hunt_budget;
    return out.perm_col == nullptr;
}

uint32_t *defer_count_ptr(uint32_t *head, int pass) { return head + defer_count_word(pass); }

hipError_t launch_check(int group_lanes, uint32_t blocks, uint32_t waves_per_block,
                        uint32_t n, uint32_t stride, const uint64_t *rows,
                        const uint64_t *cells, const uint32_t *meta,
                        const ise_results &out, uint32_t *head, uint32_t max_perms,
                        uint32_t hunt_budget, const uint8_t *order, uint32_t order_dim,
                        uint32_t *defer, const uint32_t *defer_budgets, uint32_t *sched,
                        hipStream_t stream, int pass_lo, int pass_hi) {
    // passes pass_lo..pass_hi of the launch (1-based; 0, 0: all); the queue
    // reset and the schedule belong to the first
    const bool from_start = pass_lo <= 1;
    if (from_start)
        hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
    if (sched && from_start) {
        // schedule list (n words) followed by its counters
        uint32_t *aux = sched + n;
        const uint32_t sb = (uint32_t)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
        hipLaunchKernelGGL(sched_reset_kernel, dim3(1), dim3(128), 0, stream, aux);
        hipLaunchKernelGGL(sched_count_kernel, dim3(sb), dim3(256), 0, stream, n, meta, aux);
        hipLaunchKernelGGL(sched_fill_kernel, dim3(sb), dim3(256), 0, stream, n, meta, aux, sched);
    }
    const dim3 grid(blocks), block(64 * waves_per_block);
    const OutPtrs o = to_out(out);
    const uint32_t cap = rows_cap_for(group_lanes, stride);
    // statistics counters and the witness hunt are compiled in only when asked for;
    // the column-set search runs when no column permutation is reported
    const bool count = out.counters != nullptr;
    const bool hunt = hunt_budget > 0;
    const bool sets = uses_sets(out, hunt_budget);
    const size_t lds = lds_bytes(group_lanes, waves_per_block, stride, sets);
    const int od = (int)order_dim;
    // passes: the budgets below the limit, in increasing order, then the limit;
    // pass p lists into half (p-1)&1 of `defer` (2n words) and pass p+1 reads it
    uint32_t budgets[DEFER_PASSES];
    int passes = 0;
    if (defer && defer_budgets)
        for (int i = 0; i < DEFER_PASSES - 1; i++) {
            const uint32_t b = defer_budgets[i];
            if (b > 0 && b < max_perms && (passes == 0 || b > budgets[passes - 1])) budgets[passes++] = b;
        }
    budgets[passes++] = max_perms;
#define ISE_LAUNCH_P(GL, C, H, S, P, IN, OUT, B)                                                         \
    hipLaunchKernelGGL((check_kernel<GL, C, H, S>), grid, block, lds, stream, n, stride, rows, cells, meta, o, \
                       head, max_perms, cap, order, od, hunt_budget, IN, OUT, B, P, sched)
#define ISE_LAUNCH_V(GL, C, H, S)                                                                        \
    do {                                                                                                 \
        if (passes == 1) {                                                                               \
            ISE_LAUNCH_P(GL, C, H, S, 0, nullptr, nullptr, max_perms);                                   \
        } else {                                                                                         \
            for (int p = 1; p <= passes; p++) {                                                          \
                if ((pass_lo && p < pass_lo) || (pass_hi && p > pass_hi)) continue;                      \
                const uint32_t *in = p > 1 ? defer + (size_t)((p - 2) & 1) * n : nullptr;                \
                uint32_t *outl = p < passes ? defer + (size_t)((p - 1) & 1) * n : nullptr;               \
                ISE_LAUNCH_P(GL, C, H, S, p, in, outl, budgets[p - 1]);                                  \
            }                                                                                            \
        }                                                                                                \
    } while (0)
#define ISE_LAUNCH(GL)                                          \
    if (count) {                                                \
        if (hunt && sets) ISE_LAUNCH_V(GL, true, true, true);   \
        else if (hunt) ISE_LAUNCH_V(GL, true, true, false);     \
        else if (sets) ISE_LAUNCH_V(GL, true, false, true);     \
        else ISE_LAUNCH_V(GL, true, false, false);              \
    } else {                                                    \
        if (hunt && sets) ISE_LAUNCH_V(GL, false, true, true);  \
        else if (hunt) ISE_LAUNCH_V(GL, false, true, false);    \
        else if (sets) ISE_LAUNCH_V(GL, false, false, true);    \
        else ISE_LAUNCH_V(GL, false, false, false);             \
    }
    switch (group_lanes) {
    case 16: ISE_LAUNCH(16) break;
    case 32: ISE_LAUNCH(32) break;
    case 64: ISE_LAUNCH(64) break;
    default:
        return hipErrorInvalidValue;
    }
#undef ISE_LAUNCH
#undef ISE_LAUNCH_V
#undef ISE_LAUNCH_P
    return hipGetLastError();
}

hipError_t launch_lead(uint32_t blocks, uint32_t waves_per_block, uint32_t n, uint32_t stride, const uint64_t *rows,
                       const uint32_t *meta, const ise_results &out, uint32_t *head, uint32_t max_perms, uint32_t gate,
                       const uint16_t *frag, const float *w, const uint8_t *lead_in, uint8_t *lead_out,
                       float *score_out, uint32_t *sched, hipStream_t stream) {
    if (out.perm_col) return hipErrorInvalidValue;   // the column-set form reports no column permutation
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
    if (sched) {
        uint32_t *aux = sched + n;
        const uint32_t sb = (uint32_t)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
        hipLaunchKernelGGL(sched_reset_kernel, dim3(1), dim3(128), 0, stream, aux);
        hipLaunchKernelGGL(sched_count_kernel, dim3(sb), dim3(256), 0, stream, n, meta, aux);
        hipLaunchKernelGGL(sched_fill_kernel, dim3(sb), dim3(256), 0, stream, n, meta, aux, sched);
    }
    const OutPtrs o = to_out(out);
    const uint32_t cap = rows_cap_for(64, stride);
    const size_t lds = lds_bytes(64, waves_per_block, stride, true);
    const bf16x4 *fr = reinterpret_cast<const bf16x4 *>(frag);
    if (out.counters)
        hipLaunchKernelGGL(lead_kernel<true>, dim3(blocks), dim3(64 * waves_per_block), lds, stream, n, stride, rows,
                           meta, o, head, max_perms, cap, gate, fr, w, lead_in, lead_out, score_out, sched);
    else
        hipLaunchKernelGGL(lead_kernel<false>, dim3(blocks), dim3(64 * waves_per_block), lds, stream, n, stride, rows,
                           meta, o, head, max_perms, cap, gate, fr, w, lead_in, lead_out, score_out, sched);
    return hipGetLastError();
}

hipError_t lead_occupancy(uint32_t waves_per_block, uint32_t stride, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)&lead_kernel<false>,
                                                        (int)(64 * waves_per_block),
                                                        lds_bytes(64, waves_per_block, stride, true));
}

hipError_t occupancy_blocks(int group_lanes, uint32_t waves_per_block, uint32_t stride, bool sets,
                            int *blocks_per_cu) {
    const void *f = sets ? (group_lanes == 16 ? kernel_ptr<16, true>() : group_lanes == 32 ? kernel_ptr<32, true>()
                                                                                          : kernel_ptr<64, true>())
                         : (group_lanes == 16 ? kernel_ptr<16, false>() : group_lanes == 32 ? kernel_ptr<32, false>()
                                                                                           : kernel_ptr<64, false>());
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, (int)(64 * waves_per_block),
                                                        lds_bytes(group_lanes, waves_per_block, stride, sets));
}

} // namespace ise
