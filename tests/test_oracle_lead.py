"""The oracle's lead-order search (the learned order of the first row) on
the CPU: with the rows of position 0 in increasing order it is the plain
walk, field by field and counter by counter, wherever the walk meets no
automorphism; with any order it keeps every verdict, gives canonical
candidates exactly the plain results and non-canonical ones a valid witness;
the row features it is trained on hold their defining identities."""
import numpy as np
import pytest

from _helpers import check_witness, unpack

FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "orbits", "n_gens")


def streams(oracle):
    import ise_amd
    out = []
    _, _, _, (r, m) = oracle.enumerate(13, 3, 1, capture=(4096, 32, 2, 13))
    out.append(("2-(13,3,1) enumerator", r, m))
    _, _, _, (r, m) = oracle.enumerate(16, 6, 2, capture=(8192, 32, 2, 16))
    out.append(("2-(16,6,2) enumerator", r, m))
    r, m, _ = oracle.enumerate_sampled(25, 4, 1, 8, 0x5151, 8, 16, stride=32, cap=1 << 16)
    out.append(("2-(25,4,1) sampled", r[:6000], m[:6000]))
    r, m = ise_amd.random_extensions(15, 3, 1, 3000, 4, 15, 77, stride=32)
    out.append(("2-(15,3,1) random", r, m))
    return out


@pytest.fixture(scope="module")
def cases(oracle):
    return streams(oracle)


def test_ascending_first_rows_are_the_plain_walk(oracle, cases):
    """With the rows of position 0 in increasing order and no gate the lead
    search is the walk of next_permutation before its first automorphism:
    every field and every counter equal the plain search's on the candidates
    whose walk finds no automorphism."""
    for name, rows, meta in cases:
        plain = oracle.check_batch(rows, meta, gen_cap=4)
        asc = np.tile(np.arange(16, dtype=np.uint8), (len(meta), 1))
        res, mode = oracle.check_batch_lead(rows, meta, asc, 0, gen_cap=4)
        d = meta & 0xff
        take = (plain["n_gens"] == 0) & (d >= 2) & (d <= 16)
        assert take.sum() > 100, name
        assert (mode[take] >= 2).all(), name
        for f in FIELDS + ("counters",):
            assert np.array_equal(res[f][take], plain[f][take]), (name, f)


@pytest.mark.parametrize("gate", [0, 16, 32])
def test_any_first_row_order_keeps_the_verdicts(oracle, cases, gate):
    rng = np.random.default_rng(gate + 1)
    for name, rows, meta in cases:
        plain = oracle.check_batch(rows, meta, gen_cap=4)
        lead = np.stack([rng.permutation(16) for _ in range(len(meta))]).astype(np.uint8)
        res, mode = oracle.check_batch_lead(rows, meta, lead, gate, gen_cap=4)
        assert np.array_equal(res["verdict"], plain["verdict"]), name
        canon = plain["verdict"] == 1
        for f in FIELDS:
            assert np.array_equal(res[f][canon], plain[f][canon]), (name, f)
        # exhausted lead searches: canonical, trivial group, the walk ended on the descending permutation
        ex = np.nonzero(mode == 3)[0]
        assert (plain["verdict"][ex] == 1).all() and (plain["n_gens"][ex] == 0).all(), name
        d = (meta & 0xff).astype(int)
        for c in ex[:200]:
            assert list(plain["perm_row"][c, :d[c]]) == list(range(d[c] - 1, -1, -1)), name
        # lead witnesses are valid
        b = (meta >> 8) & 0xff
        wit = np.nonzero(mode == 2)[0]
        assert len(wit) > 0 or gate > 0, name
        for c in wit[:300]:
            mat = unpack(rows[c, :d[c]], int(b[c]))
            check_witness(mat, res["perm_row"][c], int(res["fail_row"][c]))


def test_gate_decides_short_walks_unchanged(oracle, cases):
    """Candidates whose walk ends within the gate keep every plain field, counters included."""
    name, rows, meta = cases[3]
    plain = oracle.check_batch(rows, meta, gen_cap=4)
    lead = np.tile(np.arange(15, -1, -1, dtype=np.uint8), (len(meta), 1))
    res, mode = oracle.check_batch_lead(rows, meta, lead, 32, gen_cap=4)
    short = mode == 1
    assert short.sum() > 1000
    assert (plain["counters"][short, 0] <= 32).all()
    for f in FIELDS + ("counters",):
        assert np.array_equal(res[f][short], plain[f][short]), f


def test_lead_features_of_a_small_matrix(oracle):
    """Three lines of the Fano plane through point 0: every row holds one
    column of weight 3 and two of weight 1, shares its heavy column with
    rows 0 and 2, and all three rows tie."""
    rows = np.zeros((1, 32), np.uint64)
    for i, bits in enumerate(("1110000", "1001100", "1000011")):
        rows[0, i] = int(bits[::-1], 2)
    meta = np.array([3 | (7 << 8)], np.uint32)
    X = oracle.lead_features(rows, meta)[0] * 16
    for j in range(3):
        assert list(X[j, 0:4]) == [1, 0, 2, 0]
        assert list(X[j, 4:9]) == [0, 0, 0, 0, 0]
        assert X[j, 9] == 16 * (j == 0) and X[j, 10] == 16 * (j == 2)
        assert X[j, 11] == j and X[j, 12] == 3
        assert X[j, 13] == (0 if j == 0 else 3) and X[j, 14] == (0 if j == 2 else 3)
        assert X[j, 15] == 3
    assert not X[3:].any()


def test_lead_features_identities(oracle, cases):
    name, rows, meta = cases[2]
    X = oracle.lead_features(rows[:500], meta[:500]) * 16
    d = (meta[:500] & 0xff).astype(int)
    for c in range(500):
        w = [bin(int(x)).count("1") for x in rows[c, :d[c]]]
        assert np.array_equal(X[c, :d[c], 0:4].sum(1), w)                    # every column counted once
        assert np.array_equal(X[c, :d[c], 4:8], X[c, :d[c], 0:4] - X[c, 0, 0:4])
        assert (X[c, :d[c], 15] >= 1).all() and X[c, 0, 8] == 0
        assert not X[c, d[c]:].any()


def test_witness_prefix_labels():
    """Plain logs label the rows up to the witness's first row: that row 1,
    the lower ones 0, the higher ones unlabelled; only non-canonical
    candidates past the gate without automorphisms are labelled."""
    from ise_amd import learned
    meta = np.array([6 | (10 << 8)] * 4, np.uint32)
    verdict = np.array([0, 0, 1, 0], np.uint8)
    perm = np.zeros((4, 32), np.uint8)
    perm[0, 0], perm[1, 0], perm[3, 0] = 3, 0, 2
    counters = np.array([[40, 1, 1, 1], [40, 1, 1, 1], [40, 1, 1, 1], [8, 1, 1, 1]], np.uint32)
    ngen = np.zeros(4, np.uint16)
    y, w = learned.lead_labels(meta, verdict, perm, counters, ngen, 16)
    assert list(w[0, :6]) == [1, 1, 1, 1, 0, 0] and list(y[0, :6]) == [0, 0, 0, 1, 0, 0]
    assert list(w[1, :3]) == [1, 0, 0] and y[1, 0] == 1
    assert not w[2].any() and not w[3].any()     # canonical; decided within the gate
