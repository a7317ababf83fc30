"""Witness-hunt search orders on the GPU: every field equal to the oracle's
hunt-then-plain restatement, verdicts equal to the plain search's, witnesses
valid, and canonical results (|Aut|) unchanged."""
import numpy as np
import pytest

import ise_amd
from _helpers import check_witness, unpack

pytestmark = pytest.mark.gpu

FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "perm_col",
          "orbits", "n_gens", "counters")


def same(g, o, what):
    for f in FIELDS:
        a, b = g[f], o[f]
        eq = (a == b) if a.ndim == 1 else (a == b).all(axis=tuple(range(1, a.ndim)))
        if not eq.all():
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"{what}: {f} differs at {i}: gpu={a[i]} oracle={b[i]}")


def random_orders(meta, dim, seed):
    rng = np.random.default_rng(seed)
    order = np.zeros((len(meta), dim, dim), np.uint8)
    for c in range(len(meta)):
        d = int(meta[c]) & 0xff
        for i in range(d):
            order[c, i, :d] = rng.permutation(d)
    return order


@pytest.mark.parametrize("v,k,lam,budget", [(13, 3, 1, 64), (9, 3, 2, 8), (15, 7, 3, 32), (16, 6, 2, 4)])
def test_hunt_increasing_order_parity(gpu_ctx, oracle, v, k, lam, budget):
    print(f"case hunt ascending {(v, k, lam)} budget {budget}", flush=True)
    _, _, _, (rows, meta) = oracle.enumerate(v, k, lam, capture=(1 << 15, 32, 1, v))
    gpu_ctx.set_order(ise_amd.ORDER_HUNT, budget)
    try:
        g = gpu_ctx.check_batch(rows, meta)
    finally:
        gpu_ctx.set_order(ise_amd.ORDER_PLAIN)
    same(g, oracle.check_batch_hunt(rows, meta, budget=budget), f"hunt {(v, k, lam)}")
    plain = oracle.check_batch(rows, meta)
    assert np.array_equal(g["verdict"], plain["verdict"])
    canon = plain["verdict"] == 1
    assert np.array_equal(g["group_order"][canon], plain["group_order"][canon])


def test_hunt_learned_order_table_parity(gpu_ctx, oracle):
    print("case hunt order table 2-(15,3,1) random + enumerator", flush=True)
    r1, m1 = ise_amd.random_extensions(15, 3, 1, 3000, 4, 15, 11)
    _, _, _, (r2, m2) = oracle.enumerate(13, 3, 1, capture=(1 << 12, 32, 4, 13))
    rows, meta = np.concatenate([r1, r2]), np.concatenate([m1, m2])
    dim = 16
    order = random_orders(meta, dim, 5)
    n = len(meta)
    d_order = gpu_ctx.device_array(order.shape, np.uint8, order)
    gpu_ctx.set_order(ise_amd.ORDER_LEARNED, 48, d_order, dim)
    try:
        g = gpu_ctx.check_batch(rows, meta)
    finally:
        gpu_ctx.set_order(ise_amd.ORDER_PLAIN)
    same(g, oracle.check_batch_hunt(rows, meta, order=order, order_dim=dim, budget=48), "order table")
    plain = oracle.check_batch(rows, meta)
    assert np.array_equal(g["verdict"], plain["verdict"])
    for i in np.nonzero(g["verdict"] == 0)[0][:300]:
        d, b = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
        check_witness(unpack(rows[i, :d], b), g["perm_row"][i], int(g["fail_row"][i]))


def test_order_mode_arguments_checked(gpu_ctx):
    with pytest.raises(ise_amd.ISEError):
        gpu_ctx.set_order(ise_amd.ORDER_HUNT, 0)
    with pytest.raises(ise_amd.ISEError):
        gpu_ctx.set_order(ise_amd.ORDER_LEARNED, 16)
    gpu_ctx.set_order(ise_amd.ORDER_PLAIN)


SET_FIELDS = tuple(f for f in FIELDS if f != "perm_col")


@pytest.mark.parametrize("v,k,lam,budget", [(13, 3, 1, 64), (9, 3, 2, 8), (15, 7, 3, 32), (16, 6, 2, 4)])
def test_hunt_on_column_sets_parity(gpu_ctx, oracle, v, k, lam, budget):
    """The hunt in column-set form (no column permutation asked for): every
    other field equal to the oracle's hunt-then-plain restatement."""
    print(f"case hunt column sets {(v, k, lam)} budget {budget}", flush=True)
    _, _, _, (rows, meta) = oracle.enumerate(v, k, lam, capture=(1 << 15, 32, 1, v))
    r2, m2 = ise_amd.random_extensions(v, k, lam, 2000, 2, v, 0x4E + budget)
    rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    dim = 32
    order = random_orders(meta, dim, budget)
    d_order = gpu_ctx.device_array(order.shape, np.uint8, order)
    gpu_ctx.set_order(ise_amd.ORDER_LEARNED, budget, d_order, dim)
    try:
        g = gpu_ctx.check_batch(rows, meta, fields=SET_FIELDS)
    finally:
        gpu_ctx.set_order(ise_amd.ORDER_PLAIN)
    o = oracle.check_batch_hunt(rows, meta, order=order, order_dim=dim, budget=budget)
    for f in SET_FIELDS:
        a, b = g[f], o[f]
        eq = (a == b) if a.ndim == 1 else (a == b).all(axis=tuple(range(1, a.ndim)))
        if not eq.all():
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"{(v, k, lam)}: {f} differs at {i}: gpu={a[i]} oracle={b[i]}")
    d_order.free()
