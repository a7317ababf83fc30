"""Helpers shared by the tests: golden data, matrix utilities, invariants."""
import json
import os
from collections import Counter

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def counts_for(v, k, lam):
    for s in golden("bibd_counts.json"):
        if (s["v"], s["k"], s["lambda"]) == (v, k, lam):
            return s
    raise KeyError((v, k, lam))


def hist(auts):
    return {str(int(a)): int(c) for a, c in sorted(Counter(int(x) for x in auts).items())}


def unpack(words, ncols):
    words = np.asarray(words, dtype=np.uint64)
    return ((words[:, None] >> np.arange(ncols, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)


def col_sorted(mat):
    """Columns in decreasing lexicographic order (row 0 most significant)."""
    mat = np.asarray(mat, dtype=np.uint8)
    keys = [tuple(mat[:, c]) for c in range(mat.shape[1])]
    order = sorted(range(mat.shape[1]), key=lambda c: keys[c], reverse=True)
    return mat[:, order]


def lex_cmp(a, b):
    """-1/0/1 comparing matrices row by row, column 0 first, 1 > 0."""
    fa, fb = np.asarray(a).ravel(), np.asarray(b).ravel()
    diff = np.nonzero(fa != fb)[0]
    if len(diff) == 0:
        return 0
    i = diff[0]
    return 1 if fa[i] > fb[i] else -1


def design_key(words):
    return tuple(int(x) for x in words)


def check_witness(mat, perm_row, fail_row):
    """The reported witness makes the matrix larger, first at fail_row."""
    d = mat.shape[0]
    perm = [int(x) for x in perm_row[:d]]
    assert sorted(perm) == list(range(d))
    permuted = col_sorted(mat[perm])
    assert lex_cmp(permuted, mat) > 0
    first = np.nonzero((permuted != mat).any(axis=1))[0][0]
    assert first == fail_row


def p1f_relabel(mat, r0, r1, k):
    """Relabel a perfect 1-factorization so that rounds r0, r1 become its rounds
    0, 1 (the cycle of r0, r1 walked from player k against the cycle of rounds
    0, 1 walked from player 0), then put it in normal form (pairs sorted, pairs
    by first player, rounds by the partner of player 0)."""
    mat = np.asarray(mat, dtype=np.int64)
    nr, n = mat.shape
    part = np.zeros((nr, n), np.int64)
    for r in range(nr):
        part[r, mat[r, 0::2]] = mat[r, 1::2]
        part[r, mat[r, 1::2]] = mat[r, 0::2]
    tr = np.full(n, -1, np.int64)
    v1, it = k, 0
    for _ in range(n // 2):
        tr[v1] = it
        v2 = part[r0, v1]
        it = part[0, it]
        tr[v2] = it
        it = part[1, it]
        v1 = part[r1, v2]
    assert (np.sort(tr) == np.arange(n)).all()
    rows = []
    for r in range(nr):
        pairs = np.sort(tr[mat[r].reshape(-1, 2)], axis=1)
        pairs = pairs[np.argsort(pairs[:, 0])]
        rows.append(pairs.reshape(-1))
    rows.sort(key=lambda row: row[1])
    return np.array(rows, np.uint8)


def p1f_candidates(block, count, seed, rmin=3):
    """Prefixes (rmin.. rounds) of the printed perfect 1-factorizations of one
    block relabelled through a random pair of rounds and start player: rounds 0
    and 1 stay the standard pair, the partner of player 0 in round r is r + 1."""
    rng = np.random.default_rng(seed)
    mats = [np.array(m["rounds"], np.uint8) for m in block["matrices"]]
    out = []
    for _ in range(count):
        m = mats[int(rng.integers(len(mats)))]
        nr, n = m.shape
        r0, r1 = rng.choice(nr, 2, replace=False)
        full = p1f_relabel(m, int(r0), int(r1), int(rng.integers(n)))
        out.append(full[:int(rng.integers(rmin, nr + 1))])
    return out


def _cycles(pa, pb):
    """Cycles of two rounds (partner arrays): from the smallest player not yet
    met, alternately the pa- and pb-partner; players in order."""
    n, seen, out = len(pa), set(), []
    for k0 in range(n):
        if k0 in seen:
            continue
        seq, x = [k0], k0
        while True:
            y = int(pa[x])
            seq.append(y)
            z = int(pb[y])
            if z == k0:
                break
            seq.append(z)
            x = z
        seen.update(seq)
        out.append(seq)
    return sorted(out, key=len)


def u1f_candidates(block, count, seed, rmin=3):
    """Prefixes of the printed uniform 1-factorizations of one block
    relabelled so that a random ordered pair of rounds goes onto rounds 0, 1
    (its cycles onto theirs: equal lengths in a random assignment, random start
    and direction), in normal form."""
    rng = np.random.default_rng(seed)
    mats = [np.array(m["rounds"], np.int64) for m in block["matrices"]]
    out = []
    for _ in range(count):
        mat = mats[int(rng.integers(len(mats)))]
        nr, n = mat.shape
        part = np.zeros((nr, n), np.int64)
        for r in range(nr):
            part[r, mat[r, 0::2]] = mat[r, 1::2]
            part[r, mat[r, 1::2]] = mat[r, 0::2]
        r0, r1 = (int(x) for x in rng.choice(nr, 2, replace=False))
        ref, cyc = _cycles(part[0], part[1]), _cycles(part[r0], part[r1])
        assert [len(c) for c in ref] == [len(c) for c in cyc]
        tr = np.full(n, -1, np.int64)
        lens = sorted(set(len(c) for c in ref))
        for L in lens:
            P = [c for c in ref if len(c) == L]
            Q = [cyc[i] for i in rng.permutation([i for i, c in enumerate(cyc) if len(c) == L])]
            for p, q in zip(P, Q):
                d = 2 * int(rng.integers(L // 2))
                back = bool(rng.integers(2))
                for m in range(L):
                    tr[q[(d + 1 - m) % L] if back else q[(d + m) % L]] = p[m]
        assert (np.sort(tr) == np.arange(n)).all()
        rows = []
        for r in range(nr):
            pairs = np.sort(tr[mat[r].reshape(-1, 2)], axis=1)
            rows.append(pairs[np.argsort(pairs[:, 0])].reshape(-1))
        rows.sort(key=lambda row: row[1])
        full = np.array(rows, np.uint8)
        out.append(full[:int(rng.integers(rmin, nr + 1))])
    return out


def column_sorted(mats):
    """Largest column order of each matrix (columns sorted decreasingly with
    row 0 as the most significant bit); mats [m][d][b] -> same shape, int8."""
    m, d, b = mats.shape
    w = (1 << np.arange(d - 1, -1, -1)).astype(np.int64)
    keys = np.einsum("mdb,d->mb", mats.astype(np.int64), w)
    keys = -np.sort(-keys, axis=1)
    return ((keys[:, None, :] >> np.arange(d - 1, -1, -1)[None, :, None]) & 1).astype(np.int8)


def random_orbit_form_batch(seed, count, stride=8, max_rows=7, max_cols=10, stab_frac=0.35):
    """Seeded random 0/1 matrices of 2..max_rows rows and 2..max_cols columns
    in orbit form (columns in decreasing order: ones first inside every column
    orbit, as MakeRow writes rows), a share of them with lenStab > 0.
    Returns (list of matrices, rows [count][stride] uint64, meta [count])."""
    rng = np.random.default_rng(seed)
    mats = []
    rows = np.zeros((count, stride), np.uint64)
    meta = np.zeros(count, np.uint32)
    for i in range(count):
        d = int(rng.integers(2, max_rows + 1))
        b = int(rng.integers(2, max_cols + 1))
        density = float(rng.choice([0.3, 0.5, 0.7]))
        m = column_sorted((rng.random((1, d, b)) < density).astype(np.int8))[0]
        f = int(rng.integers(0, d)) if rng.random() < stab_frac else 0
        mats.append(m)
        rows[i, :d] = (m.astype(np.uint64) << np.arange(b, dtype=np.uint64)).sum(axis=1)
        meta[i] = d | (b << 8) | (f << 16)
    return mats, rows, meta


def orbit_cells(rows, d, b):
    """(start, length) of the column orbits of row d (defined by rows 0..d-1)."""
    m, full = 1, (1 << b) - 1
    for i in range(d):
        r = int(rows[i])
        m |= (r ^ (r << 1)) & full
    st = [c for c in range(b) if (m >> c) & 1]
    return [(s, (st[i + 1] if i + 1 < len(st) else b) - s) for i, s in enumerate(st)]


def largest_in_orbit(sol, gens):
    """Whether the solution (ones per orbit) is the largest of its orbit under
    the group the permutations gens generate (image[i] = sol[g[i]])."""
    seen, todo = {sol}, [sol]
    while todo:
        x = todo.pop()
        for g in gens:
            y = tuple(x[g[i]] for i in range(len(x)))
            if y > sol:
                return False
            if y not in seen:
                seen.add(y)
                todo.append(y)
    return True


def oracle_pruned_levels(oracle, v, k, lam, gen_cap=32):
    """Candidates per depth 1..v of a level enumeration built on the oracle
    (its row solver, its test and the generators it returns) in which only
    the orbit-largest solutions of each canonical matrix are extended."""
    import ise_amd
    b = ise_amd.design_params(v, k, lam)[1]
    parents = [np.zeros(32, np.uint64)]
    levels = []
    for depth in range(v):
        res = None
        if depth:
            res = oracle.check_batch(np.array(parents), np.full(len(parents), depth | (b << 8), np.uint32),
                                     gen_cap=gen_cap)
        kids = []
        for pi, p in enumerate(parents):
            cells = orbit_cells(p, depth, b) if depth else None
            gens = []
            if depth:
                ng = min(int(res["n_gens"][pi]), gen_cap)
                gens = [list(res["gens"][pi, g, :len(cells)]) for g in range(ng)]
            for s in oracle.row_solutions(v, k, lam, p, depth):
                if gens:
                    sol = tuple(bin((int(s) >> a) & ((1 << L) - 1)).count("1") for a, L in cells)
                    if not largest_in_orbit(sol, gens):
                        continue
                kid = p.copy()
                kid[depth] = s
                kids.append(kid)
        levels.append(len(kids))
        if not kids:
            break
        oi = (1 | 4) << 24 if depth + 1 == v else 0
        r = oracle.check_batch(np.array(kids), np.full(len(kids), (depth + 1) | (b << 8) | oi, np.uint32))
        parents = [kids[i] for i in range(len(kids)) if r["verdict"][i] == 1]
    return levels
