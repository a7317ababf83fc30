"""The 1-factorization oracle pinned against the reference's printed canonical
factorizations (K_10: 396 with |Aut(M)|; the uniform K_12 one with |Aut| 660),
brute force over all relabellings for n <= 8, and the published QuickPerm
visiting order."""
import itertools

import numpy as np
import pytest

import ise_amd
from _helpers import golden, p1f_candidates, p1f_relabel, u1f_candidates


def test_k10_printed_factorizations(oracle):
    block = golden("kfact.json")["k10"]
    assert len(block["matrices"]) == 396
    for m in block["matrices"]:
        v, w, auts, tried = oracle.kf_check(m["rounds"])
        assert v == 1 and auts == m["aut"]
        assert tried == 9 * 120 * 32


def test_u12_uniform_factorization(oracle):
    m = golden("kfact.json")["u12"]["matrices"][0]
    v, w, auts, tried = oracle.kf_check(m["rounds"])
    assert v == 1 and auts == 660 == m["aut"]


def test_quickperm_order_matches_published_sequence(oracle):
    # Fuchs' countdown QuickPerm lists ABC, BAC, CAB, ACB, BCA, CBA
    seq = oracle.kf_quickperm(3) // 2
    assert ["".join("ABC"[x] for x in row) for row in seq] == ["ABC", "BAC", "CAB", "ACB", "BCA", "CBA"]
    for k in range(2, 8):
        t = oracle.kf_quickperm(k)
        assert len({tuple(r) for r in t}) == len(t)      # every arrangement exactly once
        assert (np.sort(t, axis=1) == 2 * np.arange(k)).all()


def _random_factorization(rng, n, nr):
    while True:
        used = set()
        rounds = []
        ok = True
        for _ in range(nr):
            for _ in range(200):
                free = list(range(n))
                pairs = []
                while free:
                    x = free.pop(0)
                    cand = [y for y in free if (min(x, y), max(x, y)) not in used]
                    if not cand:
                        break
                    y = cand[rng.integers(len(cand))]
                    free.remove(y)
                    pairs.append((x, y))
                if len(pairs) == n // 2:
                    break
            else:
                ok = False
                break
            if len(pairs) != n // 2:
                ok = False
                break
            for x, y in pairs:
                used.add((min(x, y), max(x, y)))
            rounds.append([v for p in pairs for v in p])
        if ok:
            return np.array(rounds, np.uint8)


@pytest.mark.parametrize("n", [4, 6, 8])
def test_matches_bruteforce(oracle, n):
    """Canonical <=> no player relabelling gives a smaller normal form, and the
    automorphisms counted equal those found by brute force over all n!."""
    rng = np.random.default_rng(n)
    seen_canonical = 0
    for _ in range(40 if n < 8 else 12):
        nr = int(rng.integers(1, n))
        raw = _random_factorization(rng, n, nr)
        # relabel so that round 0 is the identity pairing, then normal form
        lab = np.empty(n, np.uint8)
        lab[raw[0]] = np.arange(n)
        mat = oracle.kf_normalize(lab[raw])
        v, w, auts, tried = oracle.kf_check(mat)
        bv, bauts = oracle.kf_bruteforce(mat)
        assert v == bv
        if v:
            seen_canonical += 1
            assert auts == bauts
    assert seen_canonical > 0


def test_random_k12_generator_is_valid_and_normal(oracle):
    rows, meta = ise_amd.kf_random(12, 3, 10, 500, 7)
    for i in range(500):
        nr = int(meta[i]) & 0xff
        mat = ise_amd.kf_unpack(rows[i, :nr], 12)
        assert (mat[0] == np.arange(12)).all()
        edges = set()
        for r in mat:
            assert sorted(r) == list(range(12))
            for x, y in zip(r[::2], r[1::2]):
                assert x < y and (x, y) not in edges
                edges.add((x, y))
        assert (oracle.kf_normalize(mat) == mat).all()


# ---- two-row canonizer of perfect 1-factorizations (cnvCheck2P1F) ----------

def _p1f_decode(idx, nr, n):
    """Relabelling index of the two-row order -> (round sent to 0, round sent to 1, start player)."""
    k, q = idx % n, idx // n
    swap, pi = q & 1, q >> 1
    last = 1
    while (last + 1) * last // 2 <= pi:
        last += 1
    ind = pi - last * (last - 1) // 2
    return (last, ind, k) if swap else (ind, last, k)


@pytest.mark.parametrize("key,tried", [("p10", 9 * 8 * 10), ("p14", 13 * 12 * 14)])
def test_p1f_printed_factorizations(oracle, key, tried):
    """The reference's printed perfect 1-factorizations of K_10 (1) and K_14
    (23) are canonical for the two-row canonizer with their printed |Aut(M)|,
    after trying every ordered pair of rounds from every start player."""
    block = golden("kfact.json")[key]
    assert len(block["matrices"]) == (1 if key == "p10" else 23)
    for m in block["matrices"]:
        v, w, auts, t = oracle.kf_check_2p1f(m["rounds"])
        assert v == 1 and auts == m["aut"] and t == tried


def test_p1f_agrees_with_the_general_canonizer(oracle):
    """On prefixes of relabelled perfect 1-factorizations of K_10 (rounds 0, 1
    the standard pair, round r pairing 0 with r + 1) the two-row canonizer and
    the general one give the same verdict and, when canonical, the same |Aut|;
    the two-row witness relabelling gives a lexicographically smaller normal
    form (checked with an independent relabelling)."""
    cands = p1f_candidates(golden("kfact.json")["p10"], 400, 0x2F1)
    canon = 0
    for mat in cands:
        v, w, auts, t = oracle.kf_check_2p1f(mat)
        gv, _, gauts, _ = oracle.kf_check(mat)
        assert v == gv
        if v:
            canon += 1
            assert auts == gauts
        else:
            nr, n = mat.shape
            r0, r1, k = _p1f_decode(w, nr, n)
            rel = p1f_relabel(mat, r0, r1, k)
            assert bytes(rel.reshape(-1)) < bytes(mat.reshape(-1))
            assert t == w + 1
    assert 0 < canon < len(cands)


def test_p1f_two_rounds_and_rejection(oracle):
    """Two rounds (one Hamiltonian cycle): all 2n relabellings are
    automorphisms (the dihedral group of the cycle); a factorization with a
    pair of rounds that is not one cycle is rejected."""
    p = np.array(golden("kfact.json")["p14"]["matrices"][0]["rounds"], np.uint8)
    v, w, auts, t = oracle.kf_check_2p1f(p[:2])
    assert v == 1 and auts == t == 2 * 14
    v, w, auts, t = oracle.kf_check_2p1f(p[:1])
    assert v == 1 and t == 0
    not_p1f = [np.array(m["rounds"], np.uint8) for m in golden("kfact.json")["k10"]["matrices"]]
    bad = 0
    for m in not_p1f:
        try:
            oracle.kf_check_2p1f(m)
        except ValueError:
            bad += 1
    assert bad == len(not_p1f) - 1   # the one perfect factorization among the 396


def test_p1f_relabel_generator(oracle):
    """ise_kf_relabel_p1f: every candidate is a perfect partial factorization in
    normal form with the sources' rounds 0, 1 and round r pairing 0 with r + 1,
    and a relabelling of a prefix-closed part of a printed matrix."""
    block = golden("kfact.json")["p14"]
    src_rows, src_meta = ise_amd.kf_pack([m["rounds"] for m in block["matrices"]])
    rows, meta = ise_amd.kf_relabel_p1f(src_rows, src_meta, 600, 2, 0x5EED)
    rows2, meta2 = ise_amd.kf_relabel_p1f(src_rows, src_meta, 600, 2, 0x5EED)
    assert np.array_equal(rows, rows2) and np.array_equal(meta, meta2)
    std = np.array(block["matrices"][0]["rounds"], np.uint8)[:2]
    canon = 0
    for i in range(len(meta)):
        nr = int(meta[i]) & 0xff
        assert 2 <= nr <= 13 and (int(meta[i]) >> 8) == 14
        mat = ise_amd.kf_unpack(rows[i, :nr], 14)
        assert (mat[:2] == std).all()
        assert (mat[:, 1] == np.arange(1, nr + 1)).all()
        assert (oracle.kf_normalize(mat) == mat).all()
        canon += oracle.kf_check_2p1f(mat)[0]   # raises unless perfect
    assert 0 < canon < len(meta)


# ---- two-row canonizer of uniform 1-factorizations (cnvCheck2U1F) ----------

@pytest.mark.parametrize("key,cycles,tried", [("u8", (4, 4), 7 * 6 * 2 * 4 * 4), ("u10", (4, 6), 9 * 8 * 4 * 6),
                                              ("u12", (6, 6), 11 * 10 * 2 * 6 * 6), ("p10", (10,), 9 * 8 * 10)])
def test_u1f_printed_factorizations(oracle, key, cycles, tried):
    """The printed uniform 1-factorizations of K_8 (cycles 4:4), K_10 (4:6) and
    K_12 (6:6), and the perfect one of K_10 (one 10-cycle), are canonical for
    the two-row canonizer with their printed |Aut(M)|, after every mapping of
    every ordered pair of rounds onto rounds 0, 1 (per pair: the cycle
    lengths' product times the permutations of equal lengths)."""
    for m in golden("kfact.json")[key]["matrices"]:
        v, w, auts, t = oracle.kf_check_2u1f(m["rounds"])
        assert v == 1 and auts == m["aut"] and t == tried


@pytest.mark.parametrize("key", ["u8", "u10"])
def test_u1f_agrees_with_the_general_canonizer(oracle, key):
    """On relabelled prefixes of the printed uniform factorizations the two-row
    canonizer and the general one agree on the verdict and, when canonical, on
    |Aut|; on perfect ones it gives the two-row P1F canonizer's |Aut|."""
    cands = u1f_candidates(golden("kfact.json")[key], 300, 0x2A + len(key))
    canon = 0
    for mat in cands:
        v, w, auts, t = oracle.kf_check_2u1f(mat)
        gv, _, gauts, _ = oracle.kf_check(mat)
        assert v == gv
        if v:
            canon += 1
            assert auts == gauts
        else:
            assert w == t - 1
    # every relabelling of the K_8 one that keeps its cycle type reproduces it
    assert canon > 0 and (key == "u8" or canon < len(cands))
    for mat in p1f_candidates(golden("kfact.json")["p10"], 100, 0x2B):
        a, b = oracle.kf_check_2u1f(mat), oracle.kf_check_2p1f(mat)
        assert a[0] == b[0] and (a[0] == 0 or a[2] == b[2])


def test_u1f_other_cycle_type_is_not_canonical(oracle):
    """A pair of rounds whose cycle lengths differ from those of rounds 0, 1
    makes the matrix not canonical at once, with no witness relabelling."""
    m = np.array(golden("kfact.json")["k10"]["matrices"][0]["rounds"], np.uint8)
    v, w, auts, t = oracle.kf_check_2u1f(m)
    assert v == 0 and w == (1 << 64) - 1


def test_u1f_relabel_generator(oracle):
    """ise_kf_relabel_p1f(uniform=True): candidates in normal form with the
    source's rounds 0, 1, every pair of rounds of their cycle type, some
    canonical and some not."""
    block = golden("kfact.json")["u12"]
    src_rows, src_meta = ise_amd.kf_pack([m["rounds"] for m in block["matrices"]])
    rows, meta = ise_amd.kf_relabel_p1f(src_rows, src_meta, 400, 2, 0x7A, uniform=True)
    std = np.array(block["matrices"][0]["rounds"], np.uint8)[:2]
    verdicts = []
    for i in range(len(meta)):
        nr = int(meta[i]) & 0xff
        mat = ise_amd.kf_unpack(rows[i, :nr], 12)
        assert (mat[:2] == std).all() and (oracle.kf_normalize(mat) == mat).all()
        v, w, auts, t = oracle.kf_check_2u1f(mat)
        assert v == 1 or w != (1 << 64) - 1      # no pair of another cycle type
        verdicts.append(v)
    assert 0 < sum(verdicts) < len(verdicts)


@pytest.mark.parametrize("fn", ["kf_check", "kf_check_2p1f", "kf_check_2u1f"])
def test_canonizers_reject_what_is_not_a_factorization(oracle, fn):
    """A round that is not a perfect matching of the players, or a pair met in
    two rounds, is an argument error for every canonizer (the device path's
    host side rejects the same records before a launch)."""
    m = np.array(golden("kfact.json")["p10"]["matrices"][0]["rounds"], np.uint8)
    check = getattr(oracle, fn)
    check(m[:3])                                   # accepted as given
    repeated = np.concatenate([m[:2], m[:1]])      # round 2 = round 0: five 2-cycles with round 0
    bad_player = m[:3].copy()
    bad_player[2, 3] = 10                          # no player 10 in K_10
    twice = m[:3].copy()
    twice[1, 0:2] = twice[1, 2:4]                  # a player twice in one round
    for bad in (repeated, bad_player, twice):
        with pytest.raises(ValueError):
            check(bad)


# ---- cnvCheck3U1F for groups of 2 -------------------------------------------

@pytest.mark.parametrize("key,cycles,tried", [("u8", (4, 4), 7 * 6 * 2 * 4 * 4), ("u10", (4, 6), 9 * 8 * 4 * 6),
                                              ("u12", (6, 6), 11 * 10 * 2 * 6 * 6), ("p10", (10,), 9 * 8 * 10)])
def test_3u1f_printed_factorizations(oracle, key, cycles, tried):
    """With their own cycle type allowed, the printed uniform and perfect
    1-factorizations are canonical with their printed |Aut(M)| (each
    automorphism kept once), every ordered pair's relabellings tried 32 times
    over; with no type given (one cycle through every player) only the perfect
    one passes, the others are not canonical before any relabelling."""
    for m in golden("kfact.json")[key]["matrices"]:
        v, w, auts, t = oracle.kf_check_3u1f(m["rounds"], [cycles])
        assert v == 1 and auts == m["aut"] and t == 32 * tried and w == (1 << 64) - 1
        v, w, auts, t = oracle.kf_check_3u1f(m["rounds"])
        if key == "p10":
            assert v == 1 and auts == m["aut"] and t == 32 * tried
        else:
            assert (v, w, auts, t) == (0, (1 << 64) - 1, 0, 0)
        v, w, auts, t = oracle.kf_check_3u1f(m["rounds"], [(4, 4, 4), cycles])
        assert v == 1 and t == 32 * tried


def test_3u1f_matches_bruteforce_on_k8(oracle):
    """Relabelled prefixes of the uniform 1-factorization of K_8 (every pair of
    rounds of type 4:4, rounds 0, 1 those of the printed one): verdict and, for
    the canonical ones, the number of automorphisms equal those of all 8!
    player relabellings."""
    cands = u1f_candidates(golden("kfact.json")["u8"], 150, 0x3A, rmin=2)
    seen = set()
    for mat in cands:
        v, w, auts, t = oracle.kf_check_3u1f(mat, [(4, 4)])
        bv, bauts = oracle.kf_bruteforce(mat)
        assert v == bv
        if v:
            assert auts == bauts
        seen.add(v)
    assert seen == {0, 1} or seen == {1}


def _per_pair(lengths):
    """Relabellings per ordered pair: the product of the cycle lengths (start
    offset x direction per cycle) times the orders of cycles of equal length."""
    from collections import Counter
    from math import factorial, prod
    return prod(lengths) * prod(factorial(c) for c in Counter(lengths).values())


def _assert_2u1f_tried_32_times(oracle, mat, types):
    v2, w2, a2, t2 = oracle.kf_check_2u1f(mat)
    v3, w3, a3, t3 = oracle.kf_check_3u1f(mat, types)
    assert v3 == v2 and a3 == a2
    if v2 == 1 or w2 == (1 << 64) - 1:
        assert w3 == w2 and t3 == 32 * t2
    else:
        T = _per_pair(_cycle_lengths(mat[0], mat[1]))
        p, q = divmod(w2, T)
        assert w3 == p * 32 * T + q and t3 == w3 + 1
    return v3


def test_3u1f_is_2u1f_tried_32_times(oracle):
    """Against the two-row U1F canonizer, rounds 0, 1's type allowed: the same
    verdict and automorphisms; a witness of that canonizer at ordered pair p,
    position q becomes (p * 32) * T + q, T the relabellings per pair, and the
    relabellings tried are that index + 1 (canonical: 32 times all); a pair of
    another cycle type ends both at once, after 32 times the relabellings of
    the pairs before it."""
    kf = golden("kfact.json")
    verdicts = set()
    for key, cyc, count in (("u10", (4, 6), 600), ("u12", (6, 6), 600), ("p10", (10,), 300)):
        for mat in u1f_candidates(kf[key], count, 0x3B + count, rmin=1):
            verdicts.add(_assert_2u1f_tried_32_times(oracle, mat, [cyc]))
    rows, meta = ise_amd.kf_random(12, 2, 9, 300, 0x3C)
    for i in range(len(meta)):
        mat = ise_amd.kf_unpack(rows[i, :int(meta[i]) & 0xff], 12)
        _assert_2u1f_tried_32_times(oracle, mat, [tuple(sorted(_cycle_lengths(mat[0], mat[1])))])
    assert verdicts == {0, 1}


def _cycle_lengths(r0, r1):
    """Cycle lengths (players) of two perfect matchings given as pair lists."""
    n = len(r0)
    pa, pb = [0] * n, [0] * n
    for r, p in ((r0, pa), (r1, pb)):
        for j in range(0, n, 2):
            p[r[j]], p[r[j + 1]] = r[j + 1], r[j]
    seen, out = set(), []
    for k0 in range(n):
        if k0 in seen:
            continue
        k, length = k0, 0
        while True:
            k1 = pa[k]
            seen.update((k, k1))
            length += 2
            k = pb[k1]
            if k == k0:
                break
        out.append(length)
    return out
