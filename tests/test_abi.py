"""CPU-side checks of the C ABI library: it loads, exports every entry point
declared in include/ise.h, and its host-side producers behave."""
import ctypes
import os
import re

import numpy as np
import pytest

import ise_amd
from _helpers import unpack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "ise.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ise_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ise_amd.lib()
    names = declared_entry_points()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ise_amd.ISEError) as e:
        ise_amd.Context(capacity=16, stride=32)
    assert e.value.code == ise_amd.ERR_NODEV


@pytest.mark.parametrize("v,k,lam,dmin,dmax", [(15, 3, 1, 4, 15), (19, 9, 4, 5, 19), (25, 4, 1, 8, 16),
                                               (16, 6, 2, 5, 16)])
def test_random_extensions_are_valid_and_seeded(v, k, lam, dmin, dmax):
    _, b, r, _, _ = ise_amd.design_params(v, k, lam)
    rows, meta = ise_amd.random_extensions(v, k, lam, 2000, dmin, dmax, 1234)
    rows2, meta2 = ise_amd.random_extensions(v, k, lam, 2000, dmin, dmax, 1234)
    rows3, _ = ise_amd.random_extensions(v, k, lam, 2000, dmin, dmax, 1235)
    assert np.array_equal(rows, rows2) and np.array_equal(meta, meta2)
    assert not np.array_equal(rows, rows3)
    depths = meta & 0xff
    assert depths.min() == dmin and depths.max() <= dmax
    assert ((meta >> 8) & 0xff == b).all()
    for i in range(0, 2000, 7):
        d = int(depths[i])
        m = unpack(rows[i, :d], b)
        assert (m.sum(axis=1) == r).all()
        assert (m.sum(axis=0) <= k).all()
        g = m.astype(int) @ m.T.astype(int)
        assert (g[~np.eye(d, dtype=bool)] <= lam).all()
        cols = [tuple(m[:, c]) for c in range(b)]
        assert cols == sorted(cols, reverse=True)
        assert (rows[i, d:] == 0).all()


def test_pack_roundtrip():
    rng = np.random.default_rng(0)
    mat = (rng.random((9, 50)) < 0.4).astype(np.uint8)
    assert np.array_equal(ise_amd.unpack_rows(ise_amd.pack_rows(mat), 50), mat)
    assert ise_amd.pack_meta(13, 26, 2, 5) == 13 | (26 << 8) | (2 << 16) | (5 << 24)


def _targets(n, dmin, dmax, seed):
    """Each candidate's depth as the generator draws it: its own splitmix64
    state, seeded from (seed, index), first output skipped."""
    M = (1 << 64) - 1

    def splitmix(s):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return s, z ^ (z >> 31)

    out = []
    for i in range(n):
        s = seed ^ ((0xD1342543DE82EF95 * (i + 1)) & M)
        s, _ = splitmix(s)
        s, z = splitmix(s)
        out.append(dmin + z % (dmax - dmin + 1))
    return np.array(out)


def test_rejection_sampling_reaches_every_depth():
    """2-(15,3,1) stress batch, depth uniform in [4, 15]: with rejection
    sampling every candidate has the depth drawn for it (depth 15 = a
    complete Steiner triple system, every pair of rows meeting once), and a
    candidate whose first walk reached its depth is the one-walk candidate."""
    n, seed = 480, 0x15031
    one, meta1 = ise_amd.random_extensions(15, 3, 1, n, 4, 15, seed, stride=16)
    rows, meta = ise_amd.random_extensions(15, 3, 1, n, 4, 15, seed, stride=16, attempts=4096)
    want = _targets(n, 4, 15, seed)
    assert np.array_equal(meta & 0xff, want)
    assert ((meta1 & 0xff) <= want).all() and ((meta1 & 0xff) < want).any()
    same = (meta1 & 0xff) == want
    assert np.array_equal(rows[same], one[same]) and np.array_equal(meta[same], meta1[same])
    for i in range(n):
        d = int(want[i])
        m = unpack(rows[i, :d], 35).astype(int)
        assert (m.sum(axis=1) == 7).all() and (m.sum(axis=0) <= 3).all()
        g = m @ m.T
        off = g[~np.eye(d, dtype=bool)]
        assert (off <= 1).all()
        if d == 15:
            assert (off == 1).all() and (m.sum(axis=0) == 3).all()
    with pytest.raises(ise_amd.ISEError):
        ise_amd.random_extensions(15, 3, 1, 4, 4, 15, seed, attempts=ise_amd.RANDOM_MAX_ATTEMPTS + 1)
