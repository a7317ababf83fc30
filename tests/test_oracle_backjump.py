"""The k-Sys backjump position (alldata::setPlayerIndex / setPlayerIndexByPos,
k-Factorization/TripleSys/Source/kmSupport.cpp:303-356) the oracle derives from
a test's first smaller relabelling, pinned by what it means to the k-Sys
enumerator: it names round s and position j such that every matrix that agrees
with this one on rounds 0..s-1 and on round s up to position j is not
canonical either, so the enumerator may go straight back to that position.

Checked here on the general canonizer, on random partial factorizations and
on relabelled prefixes of the 396 printed factorizations of K_10 (rounds in
the enumerator's shape: round i pairs player 0 with i+1): the tail of round s
past position j (whole pairs) is re-drawn at random as a valid matching,
later rounds dropped, and the result must still be non-canonical.  (The
compared prefix of the smaller row names distinct players, so j is at least
the position of the first difference, and the kept positions cover
everything compared.)  Canonical matrices keep the starting position
nrows * n - 3 (cnv.cpp:213)."""
import numpy as np

import ise_amd


def redraw_tail(mat, s, j, rng):
    """Rounds 0..s of mat with round s's pairs after position j (rounded up to
    a pair) re-paired at random, avoiding pairs of rounds 0..s-1; None if the
    draw gets stuck."""
    n = mat.shape[1]
    keep = (j | 1) + 1
    used = set()
    for r in range(s):
        for q in range(0, n, 2):
            used.add(frozenset((int(mat[r, q]), int(mat[r, q + 1]))))
    for q in range(0, keep, 2):
        used.add(frozenset((int(mat[s, q]), int(mat[s, q + 1]))))
    rest = sorted(int(x) for x in mat[s, keep:])
    for _ in range(50):
        order = list(rest)
        pairs, ok = [], True
        free = list(order)
        while free:
            a = min(free)
            free.remove(a)
            cand = [b for b in free if frozenset((a, b)) not in used]
            if not cand:
                ok = False
                break
            b = cand[int(rng.integers(len(cand)))]
            free.remove(b)
            pairs.append((a, b))
        if ok:
            row = list(mat[s, :keep]) + [x for p in pairs for x in p]
            new = np.array(list(mat[:s]) + [row], np.uint8)
            return new
    return None


def backjump_cases(oracle, rng):
    for n, rmin, rmax, seed in ((10, 3, 8, 0x10), (12, 3, 10, 0x12)):
        rows, meta = ise_amd.kf_random(n, rmin, rmax, 1500, seed)
        for i in range(len(meta)):
            yield ise_amd.kf_unpack(rows[i, :int(meta[i]) & 0xff], n)
    from _helpers import golden
    k10 = [np.array(m["rounds"], np.uint8) for m in golden("kfact.json")["k10"]["matrices"]]
    for _ in range(600):
        full = oracle.kf_normalize(rng.permutation(10).astype(np.uint8)[k10[int(rng.integers(len(k10)))]])
        yield full[:int(rng.integers(3, 10))]


def test_backjump_is_safe_for_the_enumerator(oracle):
    rng = np.random.default_rng(0xB7)
    checked = changed = 0
    for mat in backjump_cases(oracle, rng):
        if True:
            nr, n = mat.shape
            pi, row, s, was_set = oracle.kf_player_index(mat, 0)
            v, _, _, _ = oracle.kf_check(mat)
            start = nr * n - 3
            assert 0 <= pi <= start
            if v == 1:
                assert pi == start and not was_set
                continue
            if not was_set:
                assert pi == start
                continue
            assert pi // n == s or pi == start
            if row > s or pi == start:
                continue
            j = pi % n
            for _ in range(3):
                new = redraw_tail(mat, s, j, rng)
                if new is None:
                    continue
                changed += not np.array_equal(new, mat[:s + 1])
                assert oracle.kf_check(new)[0] == 0, (mat.tolist(), pi, new.tolist())
                checked += 1
    assert checked >= 200 and changed >= 50, (checked, changed)


def test_two_row_canonizers_start_and_bound(oracle):
    from _helpers import golden
    kf = golden("kfact.json")
    src, smeta = ise_amd.kf_pack([m["rounds"] for m in kf["p14"]["matrices"]])
    rows, meta = ise_amd.kf_relabel_p1f(src, smeta, 300, 3, 0xB14)
    for mode, (r, m) in ((1, (rows, meta)),
                         (2, ise_amd.kf_relabel_p1f(*ise_amd.kf_pack([kf["u12"]["matrices"][0]["rounds"]]), 300, 3,
                                                    0xB12, uniform=True))):
        n = (int(m[0]) >> 8) & 0xff
        for i in range(len(m)):
            nr = int(m[i]) & 0xff
            mat = ise_amd.kf_unpack(r[i, :nr], n)
            pi, row, s, was_set = oracle.kf_player_index(mat, mode)
            assert 0 <= pi <= nr * n - 3
            if was_set:
                assert pi // n == s
