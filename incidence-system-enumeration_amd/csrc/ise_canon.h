// ise_canon.h -- one round of the batched CanonizeMatrix loop (ise_canon.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise.h"

namespace ise {

// Arguments of canon_step_kernel.  Round 0 (first != 0) reads the caller's
// batch (rows_in / meta_in, stride_in words per matrix) and only puts the
// columns in orbit form.  Later rounds read the compact batch of the round
// before (rows_cur / meta_cur / idx_cur = original index, n_cur matrices) with
// that round's test results (verdict, perm_row, group_order): decided
// matrices go to canon / status / aut / iters at their original index, open
// ones are rearranged by their witness and appended to the next compact batch
// (rows_next / meta_next / idx_next, count_next).  `iter` = tests run so far.
// part != 0: an initial column partition (bit c set: column c starts a block,
// the orbits of a first row of labels); columns stay inside their block and
// cells_next receives every row's orbit-start mask for the next test.
// n_cur_dev (later rounds, may be null): the device word holding the number
// of matrices of the current batch, written by the round before; n_cur is
// then only an upper bound, and the slots of the next batch from its count up
// to n_cur get an empty meta word (depth 0), so that a test launched over
// n_cur slots marks them invalid at once and nothing reads them.
struct CanonStep {
    uint32_t first, n_cur, stride, stride_in, stride_out, iter, max_iters;
    uint64_t part;
    uint64_t *cells_next;
    const uint32_t *n_cur_dev;
    const uint64_t *rows_in;
    const uint32_t *meta_in;
    const uint64_t *rows_cur;
    const uint32_t *meta_cur;
    const uint32_t *idx_cur;
    const uint8_t *verdict;
    const uint8_t *perm_row;
    const uint64_t *group_order;
    uint64_t *rows_next;
    uint32_t *meta_next;
    uint32_t *idx_next;
    uint32_t *count_next;
    uint64_t *canon;
    uint8_t *status;
    uint64_t *aut;
    uint32_t *iters;
};

// Reset count_next, then the step kernel over n_cur matrices.
hipError_t canon_step_launch(const CanonStep &a, uint32_t blocks, hipStream_t stream);

} // namespace ise
