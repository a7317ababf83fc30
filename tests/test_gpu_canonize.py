"""GPU CanonizeMatrix (ise_canonize_batch / ise_canonize_matrix) against the
oracle's: canonical matrix, |Aut| and the number of tests run, on random
relabellings of the reference's printed canonical matrices (which must come
back unchanged), random 0/1 matrices and seeded partial designs (up to 50
columns), plus invalid records."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden, unpack

pytestmark = pytest.mark.gpu

STRIDE = 32


def pack(mats):
    rows = np.zeros((len(mats), STRIDE), np.uint64)
    meta = np.zeros(len(mats), np.uint32)
    for i, m in enumerate(mats):
        d, b = m.shape
        rows[i, :d] = (m.astype(np.uint64) << np.arange(b, dtype=np.uint64)).sum(axis=1)
        meta[i] = d | (b << 8)
    return rows, meta


def test_canonize_batch_parity(gpu_ctx, oracle):
    print("case canonize batch", flush=True)
    rng = np.random.default_rng(0xC0)
    mats, expect = [], []
    printed = [(np.array([[int(c) for c in r] for r in m["rows"]], np.uint8), m["aut"])
               for blk in golden("printed_bibds.json") for m in blk["matrices"]]
    for i in rng.choice(len(printed), 300, replace=False):
        mat, aut = printed[i]
        mats.append(np.ascontiguousarray(mat[rng.permutation(mat.shape[0])][:, rng.permutation(mat.shape[1])]))
        expect.append((len(mats) - 1, mat, aut))
    for _ in range(600):
        d, b = int(rng.integers(2, 13)), int(rng.integers(2, 41))
        mats.append((rng.random((d, b)) < rng.choice([0.2, 0.4, 0.6])).astype(np.uint8))
    for v, k, lam in ((15, 3, 1), (25, 4, 1), (19, 9, 4)):
        r, m = ise_amd.random_extensions(v, k, lam, 200, 4, min(v, 16), 0xC1 + v)
        b = ise_amd.design_params(v, k, lam)[1]
        for i in range(len(m)):
            mat = unpack(r[i, :int(m[i]) & 0xff], b)
            mats.append(np.ascontiguousarray(mat[rng.permutation(mat.shape[0])][:, rng.permutation(b)]))
    rows, meta = pack(mats)
    g = gpu_ctx.canonize_batch(rows, meta)
    assert (g["status"] == 1).all()
    for i, mat in enumerate(mats):
        d, b = mat.shape
        out, aut, iters = oracle.canonize(mat)
        got = unpack(g["canon"][i, :d], b)
        assert np.array_equal(got, out), (i, mat.tolist())
        assert int(g["aut"][i]) == aut and int(g["iters"][i]) == iters, (i, g["aut"][i], aut, g["iters"][i], iters)
        assert (g["canon"][i, d:] == 0).all()
    for i, mat, aut in expect:
        assert np.array_equal(unpack(g["canon"][i, :mat.shape[0]], mat.shape[1]), mat) and g["aut"][i] == aut


def test_canonize_matrix_and_invalid_records(gpu_ctx, oracle):
    print("case canonize matrix + invalid", flush=True)
    rng = np.random.default_rng(7)
    for _ in range(20):
        mat = (rng.random((int(rng.integers(2, 10)), int(rng.integers(2, 30)))) < 0.5).astype(np.uint8)
        out, aut, iters = gpu_ctx.canonize_matrix(mat)
        o_out, o_aut, o_iters = oracle.canonize(mat)
        assert np.array_equal(out, o_out) and aut == o_aut and iters == o_iters
    mats = [(rng.random((6, 9)) < 0.5).astype(np.uint8) for _ in range(4)]
    rows, meta = pack(mats)
    meta[1] = 6 | (70 << 8)        # 70 columns: too many
    meta[2] = 6 | (9 << 8) | (6 << 16)   # lenStab = depth
    g = gpu_ctx.canonize_batch(rows, meta)
    assert list(g["status"]) == [1, 3, 3, 1]
    assert (g["canon"][[1, 2]] == 0).all() and (g["aut"][[1, 2]] == 0).all()
    for i in (0, 3):
        out, aut, _ = oracle.canonize(mats[i])
        assert np.array_equal(unpack(g["canon"][i, :6], 9), out) and g["aut"][i] == aut
