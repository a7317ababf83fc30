// ise_solution.cpp -- host helpers that turn a witness into the inputs of the
// reference's CRowSolution bookkeeping.
//
// When the last row of a partial matrix makes it non-canonical, the reference
// builds an "improved" solution vector for that row
// (CCanonicityChecker::reconstructSolution, Source/CanonicityChecker.cpp:
// 227-277) and hands it to CRowSolution::moveNoncanonicalSolutions
// (Source/RowSolution.cpp:158-281), which moves the current solution's orbit
// next to the canonical solution it is equivalent to.  The vector holds, per
// column orbit of the last row, the number of ones that the witness puts
// there: the ones of row perm_row[last] among the columns the column
// permutation places in that orbit.  Orbits before the deciding one compared
// equal, so their counts are the current solution's; the deciding orbit has
// more ones, so the vector is lexicographically larger than the solution.
#include <stdint.h>
#include <string.h>
#include "ise.h"

extern "C" int ise_improved_solution(const uint64_t *rows, uint32_t depth, uint32_t ncols, const uint64_t *cells,
                                     const uint8_t *perm_row, const uint8_t *perm_col, uint32_t fail_row,
                                     uint8_t *solution_out, uint32_t *n_orbits) {
    if (!rows || !perm_row || !perm_col || !solution_out || depth < 1 || depth > ISE_MAX_ROWS || ncols < 1 ||
        ncols > ISE_MAX_COLS || fail_row != depth - 1)
        return ISE_ERR_ARG;
    const uint64_t full = ncols == 64 ? ~0ull : (1ull << ncols) - 1;
    // orbit starts of the last row: given, or from the rows above it
    uint64_t starts = 1;
    if (cells) {
        starts = (cells[depth - 1] | 1ull) & full;
    } else {
        for (uint32_t i = 0; i + 1 < depth; i++) starts |= (rows[i] ^ (rows[i] << 1)) & full;
        starts &= full;
    }
    // the witness is a permutation of the rows and of the columns
    uint64_t seen = 0;
    for (uint32_t i = 0; i < depth; i++) {
        if (perm_row[i] >= depth || (seen >> perm_row[i] & 1)) return ISE_ERR_ARG;
        seen |= 1ull << perm_row[i];
    }
    seen = 0;
    for (uint32_t c = 0; c < ncols; c++) {
        if (perm_col[c] >= ncols || (seen >> perm_col[c] & 1)) return ISE_ERR_ARG;
        seen |= 1ull << perm_col[c];
    }
    const uint64_t src = rows[perm_row[depth - 1]];
    uint32_t o = 0;
    for (uint32_t c = 0; c < ncols;) {
        uint32_t end = c + 1;
        while (end < ncols && !(starts >> end & 1)) end++;
        uint32_t ones = 0;
        for (uint32_t q = c; q < end; q++) ones += (uint32_t)(src >> perm_col[q] & 1);
        solution_out[o++] = (uint8_t)ones;
        c = end;
    }
    if (n_orbits) *n_orbits = o;
    return ISE_OK;
}
