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

__global__ void queue_reset_kernel(uint32_t *head) { *head = 0u; }

template <int G>
__global__ __launch_bounds__(1024) void check_kernel(uint32_t n, uint32_t stride,
                                                    const uint64_t *__restrict__ rows,
                                                    const uint64_t *__restrict__ cells,
                                                    const uint32_t *__restrict__ meta,
                                                    OutPtrs o, uint32_t *head, uint32_t max_perms) {
    Grp<G> g;
    for (;;) {
        int taken = 0;
        if (g.ln == 0)
            taken = (int)atomicAdd(head, 1u);
        const uint32_t ci = (uint32_t)g.read(taken, 0);
        if (ci >= n)
            break;
        const uint32_t m = meta[ci];
        const int d = m & 0xff, b = (m >> 8) & 0xff, fixed = (m >> 16) & 0xff;
        const bool ok = d >= 1 && d <= G && d <= ISE_MAX_ROWS && (uint32_t)d <= stride &&
                        b >= 1 && b <= G && fixed < d;
        if (!ok) {
            if (g.ln == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        check_candidate<G>(g, ci, m, rows, cells, stride, o, max_perms);
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

template <int G> static const void *kernel_ptr() { return (const void *)&check_kernel<G>; }

hipError_t launch_check(int group_lanes, uint32_t blocks, uint32_t waves_per_block,
                        uint32_t n, uint32_t stride, const uint64_t *rows,
                        const uint64_t *cells, const uint32_t *meta,
                        const ise_results &out, uint32_t *head, uint32_t max_perms,
                        hipStream_t stream) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, stream, head);
    const dim3 grid(blocks), block(64 * waves_per_block);
    const OutPtrs o = to_out(out);
    switch (group_lanes) {
    case 16:
        hipLaunchKernelGGL(check_kernel<16>, grid, block, 0, stream, n, stride, rows, cells, meta, o, head, max_perms);
        break;
    case 32:
        hipLaunchKernelGGL(check_kernel<32>, grid, block, 0, stream, n, stride, rows, cells, meta, o, head, max_perms);
        break;
    case 64:
        hipLaunchKernelGGL(check_kernel<64>, grid, block, 0, stream, n, stride, rows, cells, meta, o, head, max_perms);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t occupancy_blocks(int group_lanes, uint32_t waves_per_block, int *blocks_per_cu) {
    const void *f = group_lanes == 16 ? kernel_ptr<16>() : group_lanes == 32 ? kernel_ptr<32>() : kernel_ptr<64>();
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, (int)(64 * waves_per_block), 0);
}

} // namespace ise
