"""The automorphism group after the test (host part): closure of the
generators (CPermutStorage::constructGroup), restriction to column orbits
(adjustGenerators).  The closure is pinned by group orders with closed forms
and by the |Aut| the canonicity test reports for the published designs,
whose automorphism generators the oracle records."""
import math

import numpy as np
import pytest

import ise_amd
from _helpers import counts_for, hist


def compose(x, g):
    return x[g]


def cycle(n, shift=1):
    return np.roll(np.arange(n, dtype=np.uint8), -shift)


def test_symmetric_and_dihedral_orders():
    n = 6
    swap = np.arange(n, dtype=np.uint8)
    swap[[0, 1]] = [1, 0]
    G = ise_amd.group_closure(np.stack([swap, cycle(n)]))
    assert len(G) == math.factorial(n)
    assert np.array_equal(G[0], np.arange(n))
    assert len({g.tobytes() for g in G}) == len(G)
    refl = np.arange(8, dtype=np.uint8)[::-1].copy()
    assert len(ise_amd.group_closure(np.stack([cycle(8), refl]))) == 16
    # cyclic group of order lcm of cycle lengths
    p = np.array([1, 2, 0, 4, 3, 6, 7, 8, 9, 5], np.uint8)   # 3-cycle, 2-cycle, 5-cycle
    assert len(ise_amd.group_closure(p[None])) == 30


def test_closure_is_closed_and_identity_only():
    rng = np.random.default_rng(3)
    gens = np.stack([rng.permutation(7).astype(np.uint8) for _ in range(2)])
    G = ise_amd.group_closure(gens)
    S = {g.tobytes() for g in G}
    for a in G[:: max(1, len(G) // 40)]:
        for b in G[:: max(1, len(G) // 40)]:
            assert compose(a, b).tobytes() in S
    assert len(G) in (2520, 5040) or math.factorial(7) % len(G) == 0
    e = ise_amd.group_closure(np.zeros((0, 5), np.uint8))
    assert e.shape == (1, 5) and np.array_equal(e[0], np.arange(5))


def test_closure_errors():
    with pytest.raises(ise_amd.ISEError):
        ise_amd.group_closure(np.array([[0, 0, 1]], np.uint8))           # not a permutation
    swap = np.array([1, 0, 2, 3, 4], np.uint8)
    with pytest.raises(ise_amd.ISEError):
        ise_amd.group_closure(np.stack([swap, cycle(5)]), cap=100)       # |S_5| = 120 > cap


def test_adjust_generators():
    g = np.array([[1, 0, 3, 2, 4], [0, 1, 2, 3, 4]], np.uint8)
    out = ise_amd.adjust_generators(g, [2, 3])
    assert out.tolist() == [[1, 0], [0, 1]]
    with pytest.raises(ise_amd.ISEError):
        ise_amd.adjust_generators(g, [1, 2])                            # {1, 2} not invariant


@pytest.mark.parametrize("v,k,lam", [(7, 3, 1), (9, 3, 1), (13, 3, 1), (15, 3, 1), (15, 7, 3), (16, 6, 2)])
def test_generators_close_to_the_automorphism_group(oracle, v, k, lam):
    designs, auts, _, _ = oracle.enumerate(v, k, lam)
    b = v * (lam * (v - 1) // (k - 1)) // k
    rows = np.zeros((len(designs), 32), np.uint64)
    rows[:, :v] = designs
    meta = np.full(len(designs), v | (b << 8) | (ise_amd.SAVE_ROW_PERMUTATIONS << 24), np.uint32)
    res = oracle.check_batch(rows, meta, gen_cap=64)
    orders = []
    for c in range(len(designs)):
        ng = int(res["n_gens"][c])
        assert ng <= 64
        gens = res["gens"][c, :ng, :v]
        G = ise_amd.group_closure(gens, cap=1 << 16)
        assert len(G) == int(res["group_order"][c]) == int(auts[c])
        orders.append(len(G))
    if (v, k, lam) != (9, 3, 1):
        assert hist(orders) == counts_for(v, k, lam)["aut_hist"]
