"""The oracle's CanonizeMatrix (k = 0, Source/MatrixCanonChecker.cpp:420-564)
pinned by what the mathematics fixes: the result is the largest column-sorted
form over all row orders (brute force), the same byte for byte for every row
and column relabelling of the input, a fixed point of itself, and every
canonical matrix the reference prints is its own canonical form with the
printed |Aut(M)|."""
import numpy as np

from _helpers import golden


def printed():
    for block in golden("printed_bibds.json"):
        for m in block["matrices"]:
            yield block, np.array([[int(c) for c in row] for row in m["rows"]], np.uint8), m["aut"]


def relabel(rng, mat):
    return np.ascontiguousarray(mat[rng.permutation(mat.shape[0])][:, rng.permutation(mat.shape[1])])


def test_canonize_is_the_bruteforce_maximum(oracle):
    rng = np.random.default_rng(0xCA)
    for t in range(400):
        d, b = int(rng.integers(2, 8)), int(rng.integers(2, 11))
        mat = (rng.random((d, b)) < rng.choice([0.3, 0.5, 0.7])).astype(np.uint8)
        out, aut, iters = oracle.canonize(mat)
        best, n_auts = oracle.bruteforce(mat)
        assert np.array_equal(out, best), (t, mat.tolist())
        assert aut == n_auts, t
        assert iters >= 1


def test_printed_matrices_are_fixed_points(oracle):
    n = 0
    for block, mat, aut in printed():
        out, got_aut, iters = oracle.canonize(mat)
        assert np.array_equal(out, mat), block["source"]
        assert got_aut == aut and iters == 1, block["source"]
        n += 1
    assert n >= 900


def test_relabellings_canonize_to_the_printed_matrix(oracle):
    rng = np.random.default_rng(0x929)
    mats = list(printed())
    for i in rng.choice(len(mats), 240, replace=False):
        block, mat, aut = mats[i]
        for _ in range(2):
            out, got_aut, iters = oracle.canonize(relabel(rng, mat))
            assert np.array_equal(out, mat), block["source"]
            assert got_aut == aut, block["source"]


def test_canonical_form_is_invariant_and_idempotent(oracle):
    rng = np.random.default_rng(0x1D)
    for t in range(200):
        d, b = int(rng.integers(3, 12)), int(rng.integers(3, 24))
        mat = (rng.random((d, b)) < 0.4).astype(np.uint8)
        out, aut, _ = oracle.canonize(mat)
        for _ in range(3):
            o2, a2, _ = oracle.canonize(relabel(rng, mat))
            assert np.array_equal(o2, out) and a2 == aut, t
        o3, a3, it3 = oracle.canonize(out)
        assert np.array_equal(o3, out) and a3 == aut and it3 == 1, t
