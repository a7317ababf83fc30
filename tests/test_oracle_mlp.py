"""CPU pins of the learned scorer's oracle: the gradient against central
finite differences of the loss, the bf16 rounding mode against the exact
one, the pair features against the canonicity oracle, and the sampled
level enumeration against the depth-first one."""
import numpy as np

from _helpers import col_sorted, lex_cmp, unpack

P = 20752
IN, HID, OPAD = 16, 128, 16
B1, W2 = IN * HID, IN * HID + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * OPAD


def random_params(rng, scale=0.3):
    p = (rng.standard_normal(P) * scale).astype(np.float32)
    p[W3:B3].reshape(HID, OPAD)[:, 1:] = 0
    p[B3 + 1:] = 0
    return p


def test_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(3)
    M = 24
    p = random_params(rng)
    X = rng.standard_normal((M, IN)).astype(np.float32)
    y = (rng.random(M) < 0.4).astype(np.float32)
    w = rng.random(M).astype(np.float32) + 0.1
    g, _ = oracle.mlp_gradient(p, X, y, w, bf16=False)
    assert np.all(g[W3:B3].reshape(HID, OPAD)[:, 1:] == 0) and np.all(g[B3 + 1:] == 0)
    idx = np.concatenate([rng.choice(np.arange(0, B1), 12, replace=False),
                          rng.choice(np.arange(B1, W2), 6, replace=False),
                          rng.choice(np.arange(W2, B2), 12, replace=False),
                          rng.choice(np.arange(B2, W3), 6, replace=False),
                          W3 + OPAD * rng.choice(HID, 6, replace=False), [B3]])
    for i in idx:
        e = np.float32(1e-3)
        hi, lo = p.copy(), p.copy()
        hi[i] = p[i] + e
        lo[i] = p[i] - e
        _, lh = oracle.mlp_gradient(hi, X, y, w, bf16=False)
        _, ll = oracle.mlp_gradient(lo, X, y, w, bf16=False)
        fd = (lh - ll) / (float(hi[i]) - float(lo[i]))
        assert abs(fd - g[i]) <= 1e-4 + 2e-3 * abs(g[i]), (int(i), fd, g[i])


def test_bf16_rounding_mode_stays_close(oracle):
    rng = np.random.default_rng(4)
    M = 64
    p = random_params(rng)
    X = rng.standard_normal((M, IN)).astype(np.float32)
    y = (rng.random(M) < 0.3).astype(np.float32)
    w = np.ones(M, np.float32)
    ge, le = oracle.mlp_gradient(p, X, y, w, bf16=False)
    gq, lq = oracle.mlp_gradient(p, X, y, w, bf16=True)
    assert abs(le - lq) < 0.05 * abs(le)
    assert np.linalg.norm(ge - gq) < 0.05 * np.linalg.norm(ge)
    ze = oracle.mlp_forward(p, X, bf16=False)
    zq = oracle.mlp_forward(p, X, bf16=True)
    assert np.max(np.abs(ze - zq)) < 0.05 * np.max(np.abs(ze))


def test_pair_features_agree_with_the_canonicity_oracle(oracle):
    import ise_amd
    rows, meta = ise_amd.random_extensions(13, 3, 1, 400, 3, 13, 21)
    _, _, _, (r2, m2) = oracle.enumerate(9, 3, 1, capture=(2000, 32, 2, 9))
    rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    od = 13
    X = oracle.mlp_features(rows, meta, od)
    witnessed = nolarger = 0
    res = oracle.check_batch(rows, meta)
    for c in range(len(meta)):
        d, b = int(meta[c]) & 0xff, (int(meta[c]) >> 8) & 0xff
        f = X[c, :d, :d]
        # a row compared with itself is equal; intersection with itself is r
        assert np.all(f[np.arange(d), np.arange(d), 0] == 0)
        mat = unpack(rows[c, :d], b)
        assert np.allclose(f[np.arange(d), np.arange(d), 9], mat.sum(axis=1) / 16)
        # feature 0 == +1 with j > i names a witness (identity prefix, then
        # row j at position i): the candidate must be non-canonical, and a
        # canonical one has no such pair
        larger = [(i, j) for i, j in np.argwhere(f[:, :, 0] == 1) if j > i]
        if res["verdict"][c] == 1:
            assert not larger
        for i, j in larger[:3]:
            perm = list(range(i)) + [j] + [x for x in range(i, d) if x != j]
            assert lex_cmp(col_sorted(mat[perm]), mat) > 0
        if res["verdict"][c] == 0 and not larger:
            nolarger += 1
        if larger:
            witnessed += 1
        assert np.all(X[c, d:] == 0) and np.all(X[c, :, d:] == 0)
    # both kinds of non-canonical candidate occur, so the check above bites
    assert witnessed > 50 and nolarger > 0


def test_sampled_enumeration_without_beam_is_the_full_enumeration(oracle):
    rows, meta, lc = oracle.enumerate_sampled(13, 3, 1, 0, 1, 1, 13, cap=1 << 16)
    _, _, st, (r2, m2) = oracle.enumerate(13, 3, 1, capture=(1 << 16, 32, 1, 13))
    assert len(meta) == len(m2) == int(st[0])
    key = lambda r, m: sorted(zip(m.tolist(), map(tuple, r.tolist())))
    assert key(rows, meta) == key(r2, m2)
    assert int(lc.sum()) == int(st[0])


def test_sampled_enumeration_keeps_at_most_beam(oracle):
    full = oracle.enumerate_sampled(15, 3, 1, 0, 1, 1, 15, cap=1 << 20)[2]
    rows, meta, lc = oracle.enumerate_sampled(15, 3, 1, 4, 99, 1, 15, cap=1 << 20)
    # the levels before the first one with more than 4 canonical parents match
    assert lc[1] == full[1] and lc[2] == full[2]
    # a level tests at most the children of 4 parents
    depth = meta & 0xff
    res = oracle.check_batch(rows, meta)
    for d in range(1, 15):
        canon = int(((depth == d) & (res["verdict"] == 1)).sum())
        kids = int((depth == d + 1).sum())
        if canon > 4:
            assert kids < lc[d + 1] or lc[d + 1] == 0 or kids <= full[d + 1]
    again = oracle.enumerate_sampled(15, 3, 1, 4, 99, 1, 15, cap=1 << 20)
    assert np.array_equal(again[0], rows) and np.array_equal(again[1], meta)
