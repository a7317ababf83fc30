"""reconstructSolution (Source/CanonicityChecker.cpp:227-277): the improved
solution handed to CRowSolution::moveNoncanonicalSolutions when the last row
of a partial matrix makes it non-canonical.

The oracle's vector (captured inside its search at the failure) is pinned by
what it must be: lexicographically larger than the current solution (the
counts of the last row per column orbit), equal to it up to the deciding
orbit, with the weight of the witness's source row, each entry at most its
orbit's length -- and it must be realised by an actual row: some row of the
matrix, arranged inside the orbits, has exactly those counts.  The library's
host helper (ise_improved_solution, a set formulation: the source row's ones
among the columns the witness's column permutation places in each orbit)
must give the same vector from the oracle's witness."""
import numpy as np

import ise_amd
from _helpers import unpack

SETS = [(15, 3, 1), (13, 3, 1), (16, 6, 2), (19, 9, 4), (25, 4, 1)]


def orbits_of_last_row(mat):
    d, b = mat.shape
    starts = [0] + [c for c in range(1, b) if any(mat[i, c] != mat[i, c - 1] for i in range(d - 1))]
    return [(s, (starts[i + 1] if i + 1 < len(starts) else b) - s) for i, s in enumerate(starts)]


def last_row_failures(oracle, v, k, lam, count, seed):
    b = ise_amd.design_params(v, k, lam)[1]
    rows, meta = ise_amd.random_extensions(v, k, lam, count, 3, v, seed)
    res = oracle.check_batch(rows, meta)
    out = []
    for i in range(len(meta)):
        d = int(meta[i]) & 0xff
        if res["verdict"][i] == 0 and res["fail_row"][i] == d - 1:
            out.append((rows[i], d, b, res["perm_row"][i], res["perm_col"][i]))
    return out


def test_oracle_improved_solution_is_a_larger_realised_solution(oracle):
    n = 0
    for v, k, lam in SETS:
        for rows, d, b, prow, pcol in last_row_failures(oracle, v, k, lam, 1500, 0x5A + v)[:150]:
            mat = unpack(rows[:d], b)
            imp = oracle.improved_solution(mat)
            orbs = orbits_of_last_row(mat)
            assert imp is not None and len(imp) == len(orbs)
            cur = np.array([mat[d - 1, s:s + L].sum() for s, L in orbs], np.uint8)
            diff = np.nonzero(imp != cur)[0]
            assert len(diff) and imp[diff[0]] > cur[diff[0]]          # larger, first at the deciding orbit
            assert imp.sum() == mat[int(prow[d - 1])].sum()           # the witness's source row
            assert all(imp[o] <= L for o, (_, L) in enumerate(orbs))
            # realised: the source row through the witness's column permutation
            src = mat[int(prow[d - 1])]
            got = [int(src[pcol[s:s + L]].sum()) for s, L in orbs]
            assert list(imp) == got
            n += 1
    assert n >= 300


def test_library_helper_equals_oracle(oracle):
    n = 0
    for v, k, lam in SETS:
        for rows, d, b, prow, pcol in last_row_failures(oracle, v, k, lam, 1500, 0x7B + v)[:200]:
            imp = ise_amd.improved_solution(rows[:d], d, b, prow[:d], pcol[:b], d - 1)
            assert np.array_equal(imp, oracle.improved_solution(unpack(rows[:d], b)))
            n += 1
    assert n >= 250
    # not the last row: no vector
    assert ise_amd.improved_solution(np.zeros(4, np.uint64), 4, 6, np.arange(4), np.arange(6), 2) is None
