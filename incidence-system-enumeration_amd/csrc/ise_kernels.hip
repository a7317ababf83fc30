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

__global__ void queue_reset_kernel(uint32_t *head) {
    if (threadIdx.x < NUM_XCD) head[threadIdx.x * HEAD_STRIDE] = 0u;
}

// Next candidate index for the group (uniform), or n when the batch is drained.
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
// lane plus one count byte per row and lane.
template <int G, bool SETS> __host__ __device__ constexpr uint32_t group_scratch_words(uint32_t rows_cap) {
    return SETS ? rows_cap * G * 2u + rows_cap * G / 4u : rows_cap * G;
}

template <int G, bool COUNT, bool HUNT, bool SETS>
__global__ __launch_bounds__(1024) ISE_CHECK_OCCUPANCY void check_kernel(uint32_t n, uint32_t stride,
                                                    const uint64_t *__restrict__ rows,
                                                    const uint64_t *__restrict__ cells,
                                                    const uint32_t *__restrict__ meta,
                                                    OutPtrs o, uint32_t *head, uint32_t max_perms,
                                                    uint32_t rows_cap, const uint8_t *__restrict__ order,
                                                    int od, uint32_t hunt_budget) {
    extern __shared__ uint32_t lds_geo[];
    Grp<G> g;
    // this group's orbit-geometry scratch: rows_cap rows of G words
    uint32_t *geo = lds_geo + ((threadIdx.x >> 6) * (64 / G) + (g.base / G)) * group_scratch_words<G, SETS>(rows_cap);
    const int home = (int)(blockIdx.x % NUM_XCD);
    int slice = home;
    for (;;) {
        const uint32_t ci = next_candidate(g, n, head, slice, home);
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
        check_candidate<G, COUNT, HUNT, SETS>(g, ci, m, rows, cells, stride, o, max_perms, geo, order, od,
                                              hunt_budget);
    }
}

static OutPtrs to_out(const ise_results &r) {
    OutPtrs o;
    o.verdict = r.verdict; o.fail_row = r.fail_row; o.fail_col = r.fail_col;
    o.row_to_change = r.row_to_change; o.group_order = r.group_order;
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
    return out.perm_col == nullptr && hunt_budget == 0;
}

hipError_t launch_check(int group_lanes, uint32_t blocks, uint32_t waves_per_block,
                        uint32_t n, uint32_t stride, const uint64_t *rows,
                        const uint64_t *cells, const uint32_t *meta,
                        const ise_results &out, uint32_t *head, uint32_t max_perms,
                        uint32_t hunt_budget, const uint8_t *order, uint32_t order_dim,
                        hipStream_t stream) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
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
#define ISE_LAUNCH_V(GL, C, H, S)                                                                        \
    hipLaunchKernelGGL((check_kernel<GL, C, H, S>), grid, block, lds, stream, n, stride, rows, cells, meta, o, \
                       head, max_perms, cap, order, od, hunt_budget)
#define ISE_LAUNCH(GL)                                          \
    if (count) {                                                \
        if (hunt) ISE_LAUNCH_V(GL, true, true, false);          \
        else if (sets) ISE_LAUNCH_V(GL, true, false, true);     \
        else ISE_LAUNCH_V(GL, true, false, false);              \
    } else {                                                    \
        if (hunt) ISE_LAUNCH_V(GL, false, true, false);         \
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
    return hipGetLastError();
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
