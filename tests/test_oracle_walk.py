"""The oracle's plain row-permutation walk and its first witness, pinned by
textbook order and brute force rather than by the reference's own code.

* With no automorphism recorded (every row its own orbit) the walk
  (CGroupOrder::next_permutation, k-Factorization/Utils/Include/GroupOrder.h:
  136-248) must visit the permutations that fix the first lenStab rows in
  lexicographic order (Narayana's successor, the algorithm that file cites),
  and a refuted prefix p[0..at] must take it to the first lexicographically
  later permutation whose prefix of that length differs.
  One exception is the reference's own: a refuted prefix whose tail is
  exhausted falls back to the plain successor with no fixed rows
  (GroupOrder.h:183-186 recurses with the default lenStab = 0), and the BIBD
  checker's guard (`nRow < len_stab`, CanonicityChecker.h:330) reads
  lenStabilizer() = 0 (:102), so where the textbook walk would end the walk
  goes on to the next permutation of all rows, which moves a fixed row.  The
  tests pin that too: it happens exactly where the textbook walk ends.
* On matrices whose test records no automorphism, the walk is therefore plain
  lexicographic order with refuted prefixes skipped, and the reported witness
  must be the first permutation in itertools.permutations order whose
  column-sorted image is lexicographically larger than the matrix, with the
  failing row its first differing row (CanonicityChecker.h:326-416).
"""
import itertools

import numpy as np
import pytest

from _helpers import column_sorted, random_orbit_form_batch

EMAX = 255


def lex_perms(n, fixed):
    """Permutations of range(n) fixing the first `fixed` points, lexicographic."""
    head = tuple(range(fixed))
    return [head + t for t in itertools.permutations(range(fixed, n))]


def fresh_state(n):
    # as the test starts: stabiliser point n-1, no automorphism yet
    return np.array([n - 1, 255], np.uint16)


@pytest.mark.parametrize("n", range(1, 8))
def test_walk_is_lexicographic_without_automorphisms(oracle, n):
    for fixed in range(n):
        perm = np.arange(n, dtype=np.uint8)
        orb = np.arange(n, dtype=np.uint8)
        st = fresh_state(n)
        seen = []
        prev = tuple(perm)
        while True:
            r = oracle.walk_step(perm, orb, EMAX, fixed, st)
            if r == EMAX:
                break
            cur = tuple(int(x) for x in perm)
            # the returned position is the first one whose image changed
            first = next(i for i in range(n) if cur[i] != prev[i])
            assert r == first, (n, fixed, prev, cur, r)
            seen.append(cur)
            prev = cur
        assert seen == lex_perms(n, fixed)[1:], (n, fixed)


def next_with_other_prefix(order, i, cur, at):
    """Index of the first permutation after order[i] whose prefix of length
    at + 1 differs from cur's (len(order) if none)."""
    j = i + 1
    while j < len(order) and order[j][:at + 1] == cur[:at + 1]:
        j += 1
    return j


@pytest.mark.parametrize("seed", range(6))
def test_walk_skips_exactly_the_refuted_prefixes(oracle, seed):
    rng = np.random.default_rng(seed)
    left_stabiliser = 0
    for trial in range(80):
        n = int(rng.integers(2, 8))
        fixed = int(rng.integers(0, n))
        order = lex_perms(n, fixed)
        index = {p: i for i, p in enumerate(order)}
        full = lex_perms(n, 0)
        full_index = {p: i for i, p in enumerate(full)}
        perm = np.arange(n, dtype=np.uint8)
        orb = np.arange(n, dtype=np.uint8)
        st = fresh_state(n)
        at, last = EMAX, fixed
        while True:
            cur = tuple(int(x) for x in perm)
            i = index[cur]
            expect = i + 1 if at == EMAX else next_with_other_prefix(order, i, cur, at)
            r = oracle.walk_step(perm, orb, at, fixed, st)
            nxt = tuple(int(x) for x in perm)
            if expect >= len(order):
                if at == EMAX or r == EMAX:
                    # the textbook walk ends here, and so does this one
                    assert r == EMAX, (n, fixed, cur, at)
                else:
                    # the reference's fall-back: next permutation of all rows
                    # after the refuted prefix, moving a fixed row
                    j = next_with_other_prefix(full, full_index[cur], cur, at)
                    assert j < len(full) and nxt == full[j], (n, fixed, cur, at, nxt)
                    assert r < fixed and r == next(q for q in range(n) if nxt[q] != cur[q])
                    left_stabiliser += 1
                break
            assert nxt == order[expect], (n, fixed, cur, at, nxt, order[expect])
            assert r == next(q for q in range(n) if nxt[q] != cur[q])
            # the search refutes a prefix at a row at or after the changed one
            last = r
            at = EMAX if rng.random() < 0.4 else int(rng.integers(last, n))
    assert left_stabiliser > 0


def first_larger(mat, fixed):
    """(index in lexicographic order, permutation, first differing row) of the
    first non-identity row permutation fixing `fixed` rows whose column-sorted
    image is larger than mat, or None."""
    d, b = mat.shape
    perms = np.array(lex_perms(d, fixed), np.int64)
    imgs = column_sorted(mat[perms]).reshape(len(perms), d * b)
    diff = imgs - mat.reshape(1, d * b).astype(np.int8)
    nz = diff != 0
    has = nz.any(axis=1)
    pos = np.argmax(nz, axis=1)
    larger = has & (diff[np.arange(len(perms)), pos] > 0)
    larger[0] = False                      # the identity
    if not larger.any():
        return None
    k = int(np.argmax(larger))
    return k, tuple(int(x) for x in perms[k]), int(pos[k]) // b


def test_first_witness_is_the_first_larger_permutation(oracle):
    mats, rows, meta = random_orbit_form_batch(0x51DE, 3600)
    stabs = [int(x) >> 16 & 0xff for x in meta]
    res = oracle.check_batch(rows, meta, gen_cap=1)
    checked = {"noncanonical": 0, "canonical": 0, "with_stab": 0, "left_stabiliser": 0}
    for i, (m, f) in enumerate(zip(mats, stabs)):
        bf = first_larger(m, f)
        d = m.shape[0]
        got = tuple(int(x) for x in res["perm_row"][i, :d])
        if bf is None and res["verdict"][i] == 0:
            # only where the walk falls back past the stabiliser (see above):
            # the witness moves a fixed row and makes the matrix larger,
            # first at fail_row
            assert f > 0 and got[:f] != tuple(range(f)), (i, f, got)
            img = column_sorted(m[np.array(got)][None])[0].reshape(-1)
            diff = img - m.reshape(-1)
            pos = int(np.argmax(diff != 0))
            assert diff[pos] > 0 and pos // m.shape[1] == res["fail_row"][i], (i, got)
            checked["left_stabiliser"] += 1
            continue
        # otherwise the verdict is the brute force's, whatever was pruned
        assert (res["verdict"][i] == 0) == (bf is not None), i
        if res["n_gens"][i] != 0:
            continue
        if bf is None:
            checked["canonical"] += 1
            continue
        _, wperm, wrow = bf
        assert got == wperm, (i, m.tolist(), f, got, wperm)
        assert res["fail_row"][i] == wrow, (i, res["fail_row"][i], wrow)
        assert res["row_to_change"][i] == max(max(wperm[:wrow + 1]), wrow)
        checked["noncanonical"] += 1
        checked["with_stab"] += f > 0
    assert checked["noncanonical"] + checked["canonical"] >= 2000, checked
    assert checked["noncanonical"] >= 1000 and checked["with_stab"] >= 200, checked
