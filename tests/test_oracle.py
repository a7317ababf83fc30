"""The CPU oracle pinned against the reference's published results, the
canonical matrices it prints, and brute force on small matrices."""
import itertools

import numpy as np
import pytest

from _helpers import (check_witness, col_sorted, counts_for, design_key, golden, hist, lex_cmp,
                      unpack)

ORACLE_CONFIGS = [(7, 3, 1), (6, 3, 2), (9, 3, 1), (13, 3, 1), (9, 3, 2), (15, 7, 3), (16, 6, 2),
                  (10, 4, 2), (11, 5, 2), (13, 4, 1)]


@pytest.mark.parametrize("v,k,lam", ORACLE_CONFIGS)
def test_enumeration_matches_published_counts(oracle, v, k, lam):
    des, auts, st, _ = oracle.enumerate(v, k, lam)
    ref = counts_for(v, k, lam)
    assert len(des) == ref["designs"]
    assert hist(auts) == ref["aut_hist"]


@pytest.mark.parametrize("v,k,lam", [(15, 3, 1), (19, 9, 4)])
def test_enumeration_matches_published_counts_large(oracle, v, k, lam):
    des, auts, st, _ = oracle.enumerate(v, k, lam)
    ref = counts_for(v, k, lam)
    assert len(des) == ref["designs"]
    assert hist(auts) == ref["aut_hist"]


def test_designs_are_distinct_and_valid(oracle):
    v, k, lam = 9, 3, 2
    des, auts, _, _ = oracle.enumerate(v, k, lam)
    assert len({design_key(d) for d in des}) == len(des)
    b = 24
    for d in des:
        m = unpack(d, b)
        assert (m.sum(axis=0) == k).all() and (m.sum(axis=1) == 8).all()
        g = m.astype(int) @ m.T.astype(int)
        assert (g[~np.eye(v, dtype=bool)] == lam).all()


def test_printed_canonical_matrices(oracle):
    """Every canonical matrix the reference prints is canonical with its |Aut|."""
    total = 0
    for block in golden("printed_bibds.json"):
        for m in block["matrices"]:
            mat = np.array([[int(c) for c in row] for row in m["rows"]], np.uint8)
            r = oracle.test_canonicity(mat, out_info=1 | 4)
            assert r["verdict"][0] == 1, (block["source"], m)
            assert int(r["group_order"][0]) == m["aut"], block["source"]
            total += 1
    assert total >= 900


def test_printed_2_9_3_2_are_enumerated(oracle):
    des, auts, _, _ = oracle.enumerate(9, 3, 2)
    mine = {design_key(d): int(a) for d, a in zip(des, auts)}
    block = [b for b in golden("printed_bibds.json") if (b["v"], b["k"], b["lambda"]) == (9, 3, 2)][0]
    for m in block["matrices"]:
        key = tuple(sum(1 << c for c, ch in enumerate(row) if ch == "1") for row in m["rows"])
        assert mine[key] == m["aut"]


def _random_orbit_matrix(rng, d, b, density=0.5):
    mat = (rng.random((d, b)) < density).astype(np.uint8)
    return col_sorted(mat)


@pytest.mark.parametrize("seed", range(6))
def test_verdict_matches_bruteforce(oracle, seed):
    """canonical <=> the matrix equals its brute-force lexicographic maximum;
    |Aut| = number of row permutations fixing the column-sorted matrix."""
    rng = np.random.default_rng(seed)
    for _ in range(60):
        d = int(rng.integers(1, 7))
        b = int(rng.integers(1, 11))
        mat = _random_orbit_matrix(rng, d, b, float(rng.uniform(0.2, 0.8)))
        canon, n_auts = oracle.bruteforce(mat)
        r = oracle.test_canonicity(mat, out_info=1 | 4)
        is_canon = bool((canon == mat).all())
        assert bool(r["verdict"][0]) == is_canon
        if is_canon:
            assert int(r["group_order"][0]) == n_auts
        else:
            check_witness(mat, r["perm_row"][0], int(r["fail_row"][0]))
            perm = r["perm_row"][0][: int(r["fail_row"][0]) + 1]
            assert int(r["row_to_change"][0]) == max(int(r["fail_row"][0]), int(perm.max()))


def test_bruteforce_auts_of_canonical_forms(oracle):
    """Brute-force canonical forms are accepted, with the brute-force |Aut|."""
    rng = np.random.default_rng(99)
    for _ in range(200):
        d, b = int(rng.integers(2, 7)), int(rng.integers(2, 10))
        mat = _random_orbit_matrix(rng, d, b)
        canon, n_auts = oracle.bruteforce(mat)
        r = oracle.test_canonicity(canon, out_info=1)
        assert r["verdict"][0] == 1
        assert int(r["group_order"][0]) == n_auts


def _close_group(gens, d):
    ident = tuple(range(d))
    seen = {ident}
    frontier = [ident]
    while frontier:
        nxt = []
        for p in frontier:
            for g in gens:
                q = tuple(g[p[i]] for i in range(d))
                if q not in seen:
                    seen.add(q)
                    nxt.append(q)
        frontier = nxt
    return len(seen)


def test_row_generators_close_to_group_order(oracle):
    """The automorphisms found generate a group of order |Aut| (orbit-stabiliser)."""
    des, auts, _, _ = oracle.enumerate(13, 3, 1)
    for d, a in zip(des, auts):
        mat = unpack(d, 26)
        r = oracle.test_canonicity(mat, out_info=1, gen_cap=64)
        ng = int(r["n_gens"][0])
        assert 0 < ng <= 64
        gens = [tuple(int(x) for x in r["gens"][0, i, :13]) for i in range(ng)]
        for g in gens:
            assert (col_sorted(mat[list(g)]) == mat).all()
        assert _close_group(gens, 13) == int(a)


def test_partial_generators_are_orbit_permutations(oracle):
    des, _, _, cands = oracle.enumerate(9, 3, 1, capture=(4096, 32, 3, 8))
    rows, meta = cands
    res = oracle.check_batch(rows, meta, gen_cap=16)
    for i in np.nonzero(res["verdict"] == 1)[0]:
        ng = min(int(res["n_gens"][i]), 16)
        d = int(meta[i]) & 0xff
        cells = oracle.cell_masks(np.append(rows[i][:d], 0), d + 1, 12)
        norb = bin(int(cells[d])).count("1")
        for gi in range(ng):
            g = res["gens"][i, gi]
            assert sorted(g[:norb].tolist()) == list(range(norb))
            assert (g[norb:] == 255).all()


def test_cell_masks_given_or_derived_agree(oracle):
    _, _, _, cands = oracle.enumerate(9, 3, 2, capture=(3000, 32, 1, 9))
    rows, meta = cands
    cells = np.zeros_like(rows)
    for i in range(len(rows)):
        d, b = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
        cells[i, :d] = oracle.cell_masks(rows[i][:d], d, b)
    a = oracle.check_batch(rows, meta)
    c = oracle.check_batch(rows, meta, cells=cells)
    for f in a:
        assert np.array_equal(a[f], c[f]), f


@pytest.mark.parametrize("seed", range(3))
def test_hunt_never_changes_a_verdict(oracle, seed):
    """A witness hunt in any row order (random here) followed by the plain
    search gives exactly the plain verdicts and, for canonical candidates, the
    same |Aut|; every hunt witness makes its matrix larger."""
    import ise_amd
    r1, m1 = ise_amd.random_extensions(13, 3, 1, 1500, 3, 13, seed)
    _, _, _, (r2, m2) = oracle.enumerate(9, 3, 2, capture=(1 << 13, 32, 3, 9))
    rows, meta = np.concatenate([r1, r2]), np.concatenate([m1, m2])
    rng = np.random.default_rng(seed)
    dim = 16
    order = np.zeros((len(meta), dim, dim), np.uint8)
    for c in range(len(meta)):
        d = int(meta[c]) & 0xff
        for i in range(d):
            order[c, i, :d] = rng.permutation(d)
    plain = oracle.check_batch(rows, meta)
    for kw in (dict(order=order, order_dim=dim, budget=40), dict(budget=3)):
        h = oracle.check_batch_hunt(rows, meta, **kw)
        assert np.array_equal(h["verdict"], plain["verdict"])
        canon = plain["verdict"] == 1
        assert np.array_equal(h["group_order"][canon], plain["group_order"][canon])
        for i in np.nonzero(h["verdict"] == 0)[0][:200]:
            d, b = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
            check_witness(unpack(rows[i, :d], b), h["perm_row"][i], int(h["fail_row"][i]))


def _value_class_matrix(counts):
    """Rows of a few values, orbit form: (11) x counts[0], (10) x counts[1],
    (00) x counts[2] in two columns (one column when counts has two entries:
    (1) then (0))."""
    if len(counts) == 2:
        return np.array([[1]] * counts[0] + [[0]] * counts[1], np.uint8)
    return np.array([[1, 1]] * counts[0] + [[1, 0]] * counts[1] + [[0, 0]] * counts[2], np.uint8)


GROUP_ORDER_CASES = [(21, 0), (25, 0), (32, 0), (32, 1), (32, 8), (32, 16), (24, 12)]
GROUP_ORDER_CASES3 = [(12, 10, 10), (16, 8, 8), (1, 30, 1)]


def group_order_cases():
    """(matrix, |Aut|): rows of one value permute freely and neither column
    order can move a row to another value, so |Aut| (row action) is the
    product of the factorials of the value multiplicities, past 2^64 here."""
    from math import factorial
    out = []
    for d, a in GROUP_ORDER_CASES:
        counts = [d - a, a]
        out.append((_value_class_matrix(counts), factorial(d - a) * factorial(a)))
    for counts in GROUP_ORDER_CASES3:
        out.append((_value_class_matrix(list(counts)),
                    factorial(counts[0]) * factorial(counts[1]) * factorial(counts[2])))
    return out


def test_group_orders_past_64_bits(oracle):
    """|Aut| as the reference's 128-bit UInt holds it (Global.h:197-211):
    group_order the low word, group_order_hi the high one."""
    for mat, aut in group_order_cases():
        r = oracle.test_canonicity(mat, out_info=1)
        assert r["verdict"][0] == 1
        got = int(r["group_order"][0]) | (int(r["group_order_hi"][0]) << 64)
        assert got == aut, (mat.shape, got, aut)
    assert any(aut >= 1 << 64 for _, aut in group_order_cases())
