// ise_canon.hip -- canonical form of 0/1 matrices on the GPU
// (CMatrixCanonChecker::CanonizeMatrix with k = 0,
// Source/MatrixCanonChecker.cpp:420-564).
//
// The reference alternates the canonicity test (row permutations saved) with
// rearranging the rows by the witness it returns and putting the columns back
// in orbit form (GenerateBinaryColumnOrbits, :186-263), until the test finds
// no witness.  Every round makes the column-sorted matrix strictly larger, so
// the loop ends at the unique largest form.  Here each round is one launch of
// the canonicity kernel over the matrices still open plus one launch of
// canon_step_kernel, which finishes the decided ones and writes the others,
// rearranged, into the next round's compact batch.
//
// canon_step_kernel: one wave per matrix; lane c owns column c.  The rows
// (d <= 32) are read once, rearranged by the witness, and column c's key is
// its d bits with row 0 most significant; the columns in decreasing key order
// are the orbit form (ones first inside every orbit of every row).  A column's
// new position is the number of larger keys plus the equal keys before it;
// ds_permute moves every key there and d ballots rebuild the rows.  With an
// initial column partition (the k-Sys form, whose first row of day labels
// makes every day a column orbit, MatrixCanonChecker.cpp:437-449) a column is
// ranked among the columns of its own block only, and lane i writes the
// orbit-start mask row i is compared under: the block starts plus every place
// an earlier row changes value.
#include <hip/hip_runtime.h>
#include "ise_canon.h"

namespace ise {

// Slot reservation: per pass a workgroup takes CANON_TILE consecutive matrices
// per wave, every wave ballots which of its matrices stay open, and one atomic
// per workgroup reserves their slots in the next batch (one per matrix would
// serialise on the counter's cache line); a wave's slots follow the earlier
// waves' of its workgroup.
constexpr uint32_t CANON_TILE = 8;
constexpr uint32_t CANON_WAVES = 4;   // waves per workgroup (256 threads)

// Lane i holds row i of a d x b matrix (i < d): the same matrix with its
// columns in decreasing order inside each block of `part` (bit c set: column
// c starts a block; part has bit 0 and no bit past b), row 0 most
// significant, ties by column -- the orbit form.  Lane c ranks column c's key
// among its block, ds_permute moves every key to its place, d ballots
// rebuild the rows.
__device__ inline uint64_t orbit_sort(uint32_t lane, uint32_t d, uint32_t b, uint64_t part, uint64_t row) {
    // column keys: bit (d-1-i) of lane c's key is entry (i, c)
    uint32_t key = 0;
    for (uint32_t i = 0; i < d; i++) {
        // both halves zero-extended (a signed 32-bit lane value would spill into the high word)
        const uint64_t r = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, (int)i) |
                           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(row >> 32), (int)i) << 32);
        key = (key << 1) | (uint32_t)((r >> lane) & 1ull);
    }
    // decreasing key order inside each block, ties by column: lanes past b keep their place
    uint32_t pos = lane;
    if (lane < b) {
        const uint64_t upto = lane >= 63 ? part : part & ((2ull << lane) - 1ull);
        const uint32_t mine = 63u - (uint32_t)__builtin_clzll(upto);   // first column of my block
        pos = mine;
        uint32_t cs = 0;
        for (uint32_t c = 0; c < b; c++) {
            if ((part >> c) & 1ull) cs = c;
            const uint32_t kc = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)c);
            pos += cs == mine && ((kc > key) || (kc == key && c < lane));
        }
    }
    const uint32_t sorted = (uint32_t)__builtin_amdgcn_ds_permute((int)(pos << 2), (int)key);
    uint64_t out = 0;
    for (uint32_t i = 0; i < d; i++) {
        const uint64_t w = __builtin_amdgcn_ballot_w64(lane < b && ((sorted >> (d - 1 - i)) & 1u));
        if (lane == i) out = w;
    }
    return out;
}

// Lane i's orbit-start mask (the column orbits row i is compared under): the
// block starts plus every place an earlier row changes value.
__device__ inline uint64_t orbit_cells(uint32_t lane, uint32_t d, uint64_t bmask, uint64_t part, uint64_t row) {
    uint64_t cm = part;
    for (uint32_t j = 0; j < d; j++) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, (int)j);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(row >> 32), (int)j);
        const uint64_t r = ((uint64_t)hi << 32) | lo;
        if (j < lane) cm |= r ^ (r << 1);
    }
    return cm & bmask;
}

// 0 while matrix k stays open, else its final status.
__device__ inline uint32_t step_status(const CanonStep &a, uint32_t k) {
    const uint32_t m = a.first ? a.meta_in[k] : a.meta_cur[k];
    const uint32_t d = m & 0xffu, b = (m >> 8) & 0xffu, fixed = (m >> 16) & 0xffu;
    const bool shape_ok = d >= 1 && d <= ISE_MAX_ROWS && d <= a.stride && d <= a.stride_in && b >= 1 &&
                          b <= ISE_MAX_COLS && fixed < d;
    if (!shape_ok) return ISE_VERDICT_INVALID;
    if (!a.first) {
        const uint32_t v = a.verdict[k];
        if (v != ISE_VERDICT_NONCANONICAL) return v;
        if (a.iter >= a.max_iters) return ISE_VERDICT_LIMIT;
    }
    return 0;
}

__global__ __launch_bounds__(64 * CANON_WAVES) void canon_step_kernel(CanonStep a) {
    __shared__ uint32_t wave_open[CANON_WAVES];
    __shared__ uint32_t group_base;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t n_cur = a.n_cur_dev ? *a.n_cur_dev : a.n_cur;   // every thread reads the same word
    const uint32_t per_pass = CANON_WAVES * CANON_TILE;
    const uint32_t passes = (n_cur + per_pass - 1) / per_pass;
    for (uint32_t p = blockIdx.x; p < passes; p += gridDim.x) {     // the same count for every wave
        const uint32_t k0 = p * per_pass + wave * CANON_TILE;
        const uint32_t kend = k0 >= n_cur ? k0 : (k0 + CANON_TILE < n_cur ? k0 + CANON_TILE : n_cur);
        const bool open_l = lane < kend - k0 && step_status(a, k0 + lane) == 0u;
        const uint64_t open_mask = __builtin_amdgcn_ballot_w64(open_l);
        if (lane == 0) wave_open[wave] = (uint32_t)__popcll(open_mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < CANON_WAVES; w++) total += wave_open[w];
            group_base = total ? atomicAdd(a.count_next, total) : 0u;
        }
        __syncthreads();
        uint32_t base = group_base;
        for (uint32_t w = 0; w < wave; w++) base += wave_open[w];
        __syncthreads();   // read before the next pass writes them
        for (uint32_t k = k0; k < kend; k++) {
            const uint32_t orig = a.first ? k : a.idx_cur[k];
            const uint32_t m = a.first ? a.meta_in[k] : a.meta_cur[k];
            const uint32_t d = m & 0xffu, b = (m >> 8) & 0xffu;
            const uint64_t *src = a.first ? a.rows_in + (size_t)k * a.stride_in : a.rows_cur + (size_t)k * a.stride;
            const uint32_t status = step_status(a, k);   // 0: open, else the final status of this matrix
            if (status) {
                // decided: the matrix of this round (or nothing, for an invalid record)
                if (lane < a.stride_out)
                    a.canon[(size_t)orig * a.stride_out + lane] =
                        status == ISE_VERDICT_INVALID || lane >= d ? 0ull : src[lane];
                if (lane == 0) {
                    a.status[orig] = (uint8_t)status;
                    if (a.aut) a.aut[orig] = status == ISE_VERDICT_CANONICAL ? a.group_order[k] : 0ull;
                    if (a.iters) a.iters[orig] = a.iter;
                }
                continue;
            }
            // rows rearranged by the witness: new row i = row perm[i] (identity in the first round)
            uint64_t row = 0;
            if (lane < d) {
                const uint32_t from = a.first ? lane : a.perm_row[(size_t)k * ISE_MAX_ROWS + lane];
                row = src[from < d ? from : lane];
            }
            const uint64_t bmask = b >= 64 ? ~0ull : ((1ull << b) - 1ull);
            const uint64_t part = (a.part | 1ull) & bmask;
            const uint64_t out = orbit_sort(lane, d, b, part, row);
            const uint32_t slot = base + (uint32_t)__popcll(open_mask & ((1ull << (k - k0)) - 1ull));
            if (lane < a.stride) a.rows_next[(size_t)slot * a.stride + lane] = lane < d ? out : 0ull;
            if (a.cells_next) {
                const uint64_t cm = orbit_cells(lane, d, bmask, part, out);   // every lane takes part
                if (lane < a.stride) a.cells_next[(size_t)slot * a.stride + lane] = lane < d ? cm : 0ull;
            }
            if (lane == 0) {
                a.meta_next[slot] = (m & 0x00ffffffu) | (ISE_SAVE_ROW_PERMUTATIONS << 24);
                a.idx_next[slot] = orig;
            }
        }
    }
}

__global__ void canon_count_reset_kernel(uint32_t *count) {
    if (threadIdx.x == 0) *count = 0u;
}

// Slots [*count, bound) of the next batch: empty meta word (depth 0).
__global__ void canon_tail_kernel(uint32_t *meta_next, const uint32_t *count, uint32_t bound) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < bound && i >= *count) meta_next[i] = 0u;
}

hipError_t canon_step_launch(const CanonStep &a, uint32_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL(canon_count_reset_kernel, dim3(1), dim3(64), 0, stream, a.count_next);
    if (a.n_cur) hipLaunchKernelGGL(canon_step_kernel, dim3(blocks), dim3(64 * CANON_WAVES), 0, stream, a);
    if (a.n_cur_dev && a.n_cur)
        hipLaunchKernelGGL(canon_tail_kernel, dim3((a.n_cur + 255) / 256), dim3(256), 0, stream, a.meta_next,
                           a.count_next, a.n_cur);
    return hipGetLastError();
}

} // namespace ise
