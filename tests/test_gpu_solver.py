"""Row extension on the GPU (ise_solve_rows_device) against the oracle's row
solver, and the device-resident enumeration (ise_enumerate_bibd_device)
against the host-solver enumeration and the published counts."""
import numpy as np
import pytest

import ise_amd
from _helpers import counts_for, hist

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("v,k,lam", [(7, 3, 1), (13, 3, 1), (15, 3, 1), (15, 7, 3), (16, 6, 2), (19, 9, 4),
                                     (25, 4, 1), (6, 3, 12)])
def test_row_solutions_match_oracle(gpu_ctx, oracle, v, k, lam):
    print(f"case gpu row solver {(v, k, lam)}", flush=True)
    rows, meta = ise_amd.random_extensions(v, k, lam, 600, 1, v - 1, 19, stride=32)
    if v <= 16:
        _, _, _, (r2, m2) = oracle.enumerate(v, k, lam, capture=(4000, 32, 1, v - 1))
        rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    depth = meta & 0xff
    for d in np.unique(depth):
        sel = np.nonzero(depth == d)[0]
        kids, counts = gpu_ctx.solve_rows_device(v, k, lam, rows[sel], int(d))
        want, at = [], 0
        for i, c in enumerate(sel):
            sol = oracle.row_solutions(v, k, lam, rows[c], int(d))
            assert counts[i] == len(sol)
            want.append(sol)
        want = np.concatenate(want) if want else np.zeros(0, np.uint64)
        assert len(kids) == len(want)
        assert np.array_equal(kids[:, int(d)], want)
        parent_of = np.repeat(sel, counts)
        assert np.array_equal(kids[:, :int(d)], rows[parent_of, :int(d)])
        assert not kids[:, int(d) + 1:].any()


def test_row_solver_capacity_error(gpu_ctx):
    rows, meta = ise_amd.random_extensions(15, 3, 1, 64, 3, 3, 1, stride=32)
    with pytest.raises(ise_amd.ISEError) as e:
        gpu_ctx.solve_rows_device(15, 3, 1, rows, 3, cap=4)
    assert e.value.code == ise_amd.ERR_CAPACITY


@pytest.mark.parametrize("v,k,lam", [(7, 3, 1), (9, 3, 1), (9, 3, 2), (13, 3, 1), (15, 3, 1), (15, 7, 3),
                                     (16, 6, 2), (19, 9, 4), (6, 3, 12)])
def test_device_enumeration_equals_host_enumeration(gpu_ctx, v, k, lam):
    print(f"case device enumeration {(v, k, lam)}", flush=True)
    d1, a1, s1, l1 = gpu_ctx.enumerate_bibd(v, k, lam)
    d2, a2, s2, l2 = gpu_ctx.enumerate_bibd_device(v, k, lam)
    assert np.array_equal(d1, d2) and np.array_equal(a1, a2)
    assert {k: s1[k] for k in s2} == s2
    assert np.array_equal(l1, l2)
    if (v, k, lam) != (6, 3, 12):
        assert hist(a2) == counts_for(v, k, lam)["aut_hist"]
