"""The oracle's k-Sys binary canonizer (runCanonizer, Source/CDTools.cpp:25-51)
pinned against a brute force over every player order and day order, and by the
invariances a canonical form has.

The brute force works from the reference's own matrix: row 0 the day labels
(day j of `days` labelled days - 1 - j, C_InSys::prepareFirstMatrixRow,
Source/matrix.h:192-197), the player rows below (convertToBinaryMatrix,
:198-210), the columns of each label kept together in decreasing order.  It
does not use the oracle's label-row encoding.
"""
import itertools
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KF = json.load(open(os.path.join(HERE, "golden", "kfact.json")))


@pytest.fixture(scope="module")
def orc(oracle):
    return oracle


def round_robin(v):
    """1-factorization of K_v (circle method): v - 1 days of v/2 pairs."""
    days = []
    for r in range(v - 1):
        pairs = [(v - 1, r)]
        for i in range(1, v // 2):
            pairs.append(((r + i) % (v - 1), (r - i) % (v - 1)))
        days.append([p for pr in pairs for p in pr])
    return np.array(days, np.uint8)


def affine_plane_3():
    """The four parallel classes of AG(2,3): 9 players, groups of 3."""
    pts = [(x, y) for x in range(3) for y in range(3)]
    idx = {p: i for i, p in enumerate(pts)}
    days = []
    for slope in (None, 0, 1, 2):
        groups = []
        for c in range(3):
            if slope is None:
                line = [(c, y) for y in range(3)]
            else:
                line = [(x, (slope * x + c) % 3) for x in range(3)]
            groups.extend(idx[p] for p in line)
        days.append(groups)
    return np.array(days, np.uint8)


def kf_fixture(name, i=0):
    return np.array(KF[name]["matrices"][i]["rounds"], np.uint8)


def brute_force(mat, v, k, days, walk):
    """Largest matrix over player orders (and day orders when walk)."""
    G = v // k
    b = days * G
    perms = np.array(list(itertools.permutations(range(v))), np.int64)   # row i holds player perms[:, i]
    pos = np.argsort(perms, axis=1)                                       # pos[:, p] = row of player p
    best = None
    for order in (itertools.permutations(range(days)) if walk else [tuple(range(days))]):
        cols = []
        for j, d in enumerate(order):
            day = mat[d]
            vals = []
            for g in range(G):
                x = np.zeros(len(perms), np.int64)
                for n in range(k):
                    x |= np.int64(1) << (v - 1 - pos[:, day[g * k + n]])
                vals.append(x)
            vals = np.sort(np.stack(vals, 1), axis=1)[:, ::-1]            # decreasing inside the day
            cols.append(vals)
        colv = np.concatenate(cols, 1)                                    # (perms, b), row 0 = MSB
        rows = np.zeros((len(perms), v), np.int64)
        for i in range(v):
            bits = (colv >> (v - 1 - i)) & 1
            rows[:, i] = (bits << (b - 1 - np.arange(b))).sum(1)          # column 0 most significant
        top = np.lexsort([rows[:, i] for i in reversed(range(v))])[-1]
        cand = tuple(int(x) for x in rows[top])
        if best is None or cand > best:
            best = cand
    out = np.zeros((v, b), np.uint8)
    for i, r in enumerate(best):
        for q in range(b):
            out[i, q] = (r >> (b - 1 - q)) & 1
    return out


def relabel(mat, v, k, rng, shuffle_days):
    """Same k-system with players renamed, groups and their members shuffled
    (and the days reordered when shuffle_days)."""
    pl = rng.permutation(v).astype(np.uint8)
    out = []
    for day in mat:
        groups = [list(pl[day[g * k:(g + 1) * k]]) for g in range(v // k)]
        for gr in groups:
            rng.shuffle(gr)
        rng.shuffle(groups)
        out.append([p for gr in groups for p in gr])
    out = np.array(out, np.uint8)
    if shuffle_days:
        out = out[rng.permutation(len(out))]
    return out


@pytest.mark.parametrize("v", [4, 6])
def test_full_days_match_brute_force_k2(orc, v):
    mat = round_robin(v)
    rng = np.random.default_rng(v)
    for t in range(3):
        m = relabel(mat, v, 2, rng, shuffle_days=True) if t else mat
        got = orc.cd_run_canonizer(m, v, 2)
        assert np.array_equal(got, brute_force(m, v, 2, v - 1, walk=True))


@pytest.mark.parametrize("day", [1, 2, 3, 4])
def test_partial_days_match_brute_force(orc, day):
    """Fewer days than the maximum: the day order stays (numClasses = 0)."""
    v = 6
    rng = np.random.default_rng(100 + day)
    mat = relabel(round_robin(v), v, 2, rng, shuffle_days=True)[:day]
    got = orc.cd_run_canonizer(mat, v, 2, day)
    assert got.shape == (v, day * 3)
    assert np.array_equal(got, brute_force(mat, v, 2, day, walk=False))


def test_partial_days_keep_the_day_order(orc):
    """With days fixed the label of day 0 is the largest, so the day order is
    part of the input: every order of four days of a K_8 factorization
    matches the brute force, the orders do not all give one matrix (they give
    four), and the largest of them is the form over all day orders."""
    v = 8
    mat = kf_fixture("u8")[:4]
    forms = []
    for o in itertools.permutations(range(4)):
        m = mat[list(o)]
        got = orc.cd_run_canonizer(m, v, 2, 4)
        assert np.array_equal(got, brute_force(m, v, 2, 4, walk=False))
        forms.append(got)
    assert len({f.tobytes() for f in forms}) == 4
    best = max(forms, key=lambda x: tuple(x.reshape(-1)))
    assert np.array_equal(best, brute_force(mat, v, 2, 4, walk=True))


def test_groups_of_three_affine_plane(orc):
    v, k = 9, 3
    mat = affine_plane_3()
    rng = np.random.default_rng(9)
    m = relabel(mat, v, k, rng, shuffle_days=True)
    got = orc.cd_run_canonizer(m, v, k)
    assert np.array_equal(got, orc.cd_run_canonizer(mat, v, k))
    assert np.array_equal(got, brute_force(m, v, k, 4, walk=True))
    for day in (2, 3):
        assert np.array_equal(orc.cd_run_canonizer(m[:day], v, k, day), brute_force(m[:day], v, k, day, walk=False))


@pytest.mark.parametrize("name", ["u8"])
def test_full_days_invariant_under_relabelling(orc, name):
    mat = kf_fixture(name)
    v = mat.shape[1]
    ref = orc.cd_run_canonizer(mat, v, 2)
    rng = np.random.default_rng(8)
    for _ in range(2):
        assert np.array_equal(orc.cd_run_canonizer(relabel(mat, v, 2, rng, shuffle_days=True), v, 2), ref)
    # a k-system incidence: k ones per column, one per day in each player row
    assert (ref.sum(0) == 2).all() and (ref.sum(1) == v - 1).all()


@pytest.mark.parametrize("name,day", [("k10", 4), ("p10", 6), ("u12", 5)])
def test_partial_days_invariant_under_relabelling(orc, name, day):
    mat = kf_fixture(name)[:day]
    v = mat.shape[1]
    ref = orc.cd_run_canonizer(mat, v, 2, day)
    rng = np.random.default_rng(day)
    for _ in range(2):
        assert np.array_equal(orc.cd_run_canonizer(relabel(mat, v, 2, rng, shuffle_days=False), v, 2, day), ref)
    assert ref.shape == (v, day * v // 2)


def test_day_starts(orc):
    assert orc.cd_day_starts(3, 5) == 1 | 1 << 5 | 1 << 10
    assert orc.cd_day_starts(1, 7) == 1


def test_one_block_is_the_plain_canonical_form(orc):
    """part = 1 (one block) is CanonizeMatrix with k = 0; a block per column
    keeps every column in place, so only the rows are sorted (decreasing)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        mat = (rng.random((7, 11)) < 0.4).astype(np.uint8)
        a, aut_a, _ = orc.canonize_part(mat, 1)
        b, aut_b, _ = orc.canonize(mat)
        assert np.array_equal(a, b) and aut_a == aut_b
        c, _, _ = orc.canonize_part(mat, (1 << 11) - 1)
        keys = [tuple(r) for r in mat]
        assert [tuple(r) for r in c] == sorted(keys, reverse=True)


def test_rejects_invalid_input(orc):
    with pytest.raises(ValueError):
        orc.cd_run_canonizer(np.full((3, 4), 7, np.uint8), 4, 2)        # player 7 of 4
    with pytest.raises(ValueError):
        orc.cd_run_canonizer(round_robin(12), 12, 2)                      # 66 columns


def test_canonize_orders_over_every_order_is_the_walk(orc):
    """The timing helper: the largest form over all day orders given
    explicitly equals the full-day run; over one order, that order's form."""
    mat = kf_fixture("u8")
    every = np.array(list(itertools.permutations(range(7))), np.uint8)
    assert np.array_equal(orc.cd_canonize_orders(mat, 8, 2, every), orc.cd_run_canonizer(mat, 8, 2))
    one = np.array([[3, 1, 4, 0, 6, 5, 2]], np.uint8)
    identity = np.arange(7, dtype=np.uint8)[None]
    reordered = np.ascontiguousarray(mat[one[0]])
    assert np.array_equal(orc.cd_canonize_orders(mat, 8, 2, one), orc.cd_canonize_orders(reordered, 8, 2, identity))


def _aut_from_day_orders(orc, mat, v, k):
    """|Aut(M)| of a full k-system from its canonical forms: the automorphisms
    are a day permutation times a player permutation; the day orders whose
    form is the largest are the day parts (one per coset of the automorphisms
    that fix every day), and those are counted by the largest form's |Aut|
    with the days fixed (the canonical-form loop's row-action order)."""
    days = mat.shape[0]
    best = orc.cd_run_canonizer(mat, v, k)
    reach = sum(np.array_equal(orc.cd_canonize_orders(mat, v, k, np.array([o], np.uint8)), best)
                for o in itertools.permutations(range(days)))
    _, fixed_aut, _ = orc.canonize_part(best, orc.cd_day_starts(days, v // k))
    return reach * fixed_aut, best


@pytest.mark.parametrize("name,index", [("u8", 0), ("k15x3", 6)])
def test_full_day_forms_give_the_printed_aut(orc, name, index):
    """The printed |Aut(M)| of the reference's k-systems (K_8 uniform
    1-factorization: 1344; a Kirkman schoolgirl solution: 21) equals (day
    orders reaching the largest form, counted order by order) x (its |Aut|
    with days fixed)."""
    m = KF[name]["matrices"][index]
    mat = np.array(m["rounds"], np.uint8)
    v = mat.shape[1]
    k = 2 if name == "u8" else 3
    aut, best = _aut_from_day_orders(orc, mat, v, k)
    assert aut == m["aut"]
    # the walk's own count gives the same orders
    form, day_order, fixed_order = orc.cd_run_canonizer_orders(mat, v, k)
    assert np.array_equal(form, best) and day_order * fixed_order == m["aut"]


def test_kirkman_solutions(orc):
    """The seven printed 15-player k-systems in groups of three are the seven
    non-isomorphic Kirkman schoolgirl solutions: seven different canonical
    matrices; the walk's day-order count x fixed-day |Aut| is each printed
    |Aut(M)| (24, 12, 24, 12, 168, 168, 21); a relabelled, regrouped,
    reordered copy reaches the same matrix."""
    forms = []
    for m in KF["k15x3"]["matrices"]:
        mat = np.array(m["rounds"], np.uint8)
        f, day_order, fixed_order = orc.cd_run_canonizer_orders(mat, 15, 3)
        assert day_order * fixed_order == m["aut"]
        forms.append(f)
    assert len({f.tobytes() for f in forms}) == 7
    rng = np.random.default_rng(15)
    mat = np.array(KF["k15x3"]["matrices"][1]["rounds"], np.uint8)
    assert np.array_equal(orc.cd_run_canonizer(relabel(mat, 15, 3, rng, shuffle_days=True), 15, 3), forms[1])


@pytest.mark.parametrize("name", ["p15x3"])
def test_walk_group_orders_give_the_printed_aut(orc, name):
    """Every printed full k-system of the fixture: the walk's day-order count
    times the fixed-day |Aut| is the printed |Aut(M)|; partial days report
    one day order (the printed K_10 ones are checked on the GPU)."""
    k = 3 if name.endswith("x3") else 2
    for m in KF[name]["matrices"]:
        mat = np.array(m["rounds"], np.uint8)
        v = mat.shape[1]
        _, day_order, fixed_order = orc.cd_run_canonizer_orders(mat, v, k)
        assert day_order * fixed_order == m["aut"], (name, day_order, fixed_order)
    _, day_order, _ = orc.cd_run_canonizer_orders(np.ascontiguousarray(mat[:3]), v, k, 3)
    assert day_order == 1


def test_partitioned_canonical_form_is_the_brute_force_maximum(orc):
    """orc_canonize_part on random small matrices and random partitions: the
    largest matrix over every row order, with the columns of each block in
    decreasing order (row 0 most significant) -- which is the largest over
    the column orders inside the blocks for that row order."""
    rng = np.random.default_rng(31)
    for _ in range(40):
        d, b = int(rng.integers(2, 7)), int(rng.integers(2, 11))
        mat = (rng.random((d, b)) < rng.uniform(0.25, 0.6)).astype(np.uint8)
        starts = [0] + sorted(int(x) for x in rng.choice(np.arange(1, b), int(rng.integers(0, b)), replace=False))
        part = sum(1 << s for s in starts)
        bounds = starts + [b]
        best = None
        for perm in itertools.permutations(range(d)):
            m = mat[list(perm)]
            cols = []
            for lo, hi in zip(bounds[:-1], bounds[1:]):
                block = [tuple(m[:, q]) for q in range(lo, hi)]
                cols.extend(sorted(block, reverse=True))
            cand = tuple(tuple(c[i] for c in cols) for i in range(d))
            if best is None or cand > best:
                best = cand
        got, _, _ = orc.canonize_part(mat, part)
        assert [tuple(r) for r in got] == list(best), (mat.tolist(), starts)
