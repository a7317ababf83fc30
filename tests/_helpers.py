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
