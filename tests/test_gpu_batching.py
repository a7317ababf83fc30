"""Uneven batches (three parameter sets interleaved in one work queue), graph
replay per enumeration level, and the launch autotuner -- all checked against
the oracle."""
import numpy as np
import pytest

import ise_amd
from _helpers import counts_for, design_key, hist

pytestmark = pytest.mark.gpu

SETS = [(15, 7, 3), (16, 6, 2), (19, 9, 4)]


def mixed_batch(n_each, seed):
    """Seeded 1:1:1 interleaving of random extensions at depths 5..v."""
    parts = [ise_amd.random_extensions(v, k, lam, n_each, 5, v, seed + v) for v, k, lam in SETS]
    rows = np.concatenate([p[0] for p in parts])
    meta = np.concatenate([p[1] for p in parts])
    order = np.random.default_rng(seed).permutation(len(meta))
    return rows[order], meta[order]


def test_mixed_queue_parity(gpu_ctx, oracle):
    print("case mixed 1:1:1 queue", flush=True)
    rows, meta = mixed_batch(3000, 0xC3)
    g = gpu_ctx.check_batch(rows, meta, gen_cap=4)
    o = oracle.check_batch(rows, meta, gen_cap=4)
    for f in ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "perm_col",
              "orbits", "n_gens", "counters"):
        assert np.array_equal(g[f], o[f]), f


def test_autotuner_picks_a_fitting_config_and_stays_exact(gpu_ctx, oracle):
    print("case autotune", flush=True)
    rows, meta = mixed_batch(4000, 0xA7)
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    dev, s = gpu_ctx.device_results(n, ("verdict", "fail_row", "perm_row", "group_order"))
    best, table = gpu_ctx.autotune(n, 32, d_rows, d_meta, s, d_head, max_depth=19, max_cols=19, repeats=2)
    assert best["group_lanes"] in (32, 64)
    assert (table[0] == -1).all()            # 16 lanes cannot hold 19 columns
    assert (table[1:] > 0).all()
    gpu_ctx.check_batch_device(n, 32, d_rows, d_meta, s, d_head)
    gpu_ctx.sync()
    o = oracle.check_batch(rows, meta)
    for f in ("verdict", "fail_row", "perm_row", "group_order"):
        assert np.array_equal(dev[f].to_host(), o[f]), f
    gpu_ctx.set_launch()


@pytest.mark.parametrize("v,k,lam", SETS + [(13, 3, 1)])
def test_enumeration_with_graph_per_level(v, k, lam, oracle):
    print(f"case graph enumeration {(v, k, lam)}", flush=True)
    with ise_amd.Context(capacity=1 << 14, stride=32) as ctx:
        ctx.use_graphs(True)
        des, auts, stats, levels = ctx.enumerate_bibd(v, k, lam)
        graphs, replays = ctx.graph_stats()
        updates = ctx.graph_updates()
    ref = counts_for(v, k, lam)
    assert len(des) == ref["designs"] and hist(auts) == ref["aut_hist"]
    odes, oauts, _, _ = oracle.enumerate(v, k, lam)
    assert sorted(map(design_key, des)) == sorted(map(design_key, odes))
    # one graph for the whole enumeration: every level (and every tail chunk)
    # replays it, updated in place when the batch size changes
    chunks = sum(-(-int(c) // (1 << 14)) for c in levels if c)
    assert graphs == 1 and replays == chunks
    assert 1 <= updates < replays


def same_results(g, o, fields, what):
    """Field by field; generators only up to the number found (and stored)."""
    for f in fields:
        if f == "gens":
            cap = g["gens"].shape[1]
            for i in range(len(g["verdict"])):
                ng = min(int(o["n_gens"][i]), cap)
                assert np.array_equal(g["gens"][i, :ng], o["gens"][i, :ng]), (what, "gens", i)
        else:
            assert np.array_equal(g[f], o[f]), (what, f)


def test_graph_cache_survives_buffer_growth(oracle):
    """Captured graphs address context buffers: a larger gen_cap reallocates
    the generator buffer and a device batch above the context capacity the
    schedule list, so the cache must not replay a graph holding the old
    addresses.  gen_cap 2 -> 4 -> 2, then a device batch above capacity, then
    the first shape again; every result is checked against the oracle."""
    print("case graph cache buffer growth", flush=True)
    rows, meta = ise_amd.random_extensions(15, 7, 3, 1500, 5, 15, 0x77)
    big_rows, big_meta = ise_amd.random_extensions(15, 7, 3, 3000, 5, 15, 0x78)
    fields = ("verdict", "fail_row", "row_to_change", "group_order", "perm_row", "orbits", "n_gens", "gens")
    with ise_amd.Context(capacity=2048, stride=32) as ctx:
        ctx.use_graphs(True)
        for step, cap in enumerate((2, 4, 2)):
            g = ctx.check_batch(rows, meta, gen_cap=cap, fields=fields)
            o = oracle.check_batch(rows, meta, gen_cap=cap)
            same_results(g, o, fields, (step, cap))
        n = len(big_meta)
        d_rows = ctx.device_array(big_rows.shape, np.uint64, big_rows)
        d_meta = ctx.device_array(big_meta.shape, np.uint32, big_meta)
        d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
        dev, s = ctx.device_results(n, ("verdict", "perm_row", "group_order"))
        ctx.check_batch_device(n, 32, d_rows, d_meta, s, d_head)
        ctx.sync()
        ob = oracle.check_batch(big_rows, big_meta)
        for f in ("verdict", "perm_row", "group_order"):
            assert np.array_equal(dev[f].to_host(), ob[f]), ("device batch", f)
        for step, cap in enumerate((2, 4)):
            g = ctx.check_batch(rows, meta, gen_cap=cap, fields=fields)
            o = oracle.check_batch(rows, meta, gen_cap=cap)
            same_results(g, o, fields, ("after growth", step, cap))
        ctx.set_defer(16, 0)
        g = ctx.check_batch(rows, meta, gen_cap=2, fields=fields)
        o = oracle.check_batch(rows, meta, gen_cap=2)
        same_results(g, o, fields, "after set_defer")
        ctx.set_defer(0, 0)


def test_graph_replays_between_eager_work(oracle):
    """The same captured batch replayed three times, with eager launches of
    another batch in between; every replay is checked."""
    print("case graph replays", flush=True)
    a_rows, a_meta = ise_amd.random_extensions(15, 3, 1, 2048, 4, 15, 1)
    b_rows, b_meta = ise_amd.random_extensions(13, 3, 1, 1000, 3, 13, 2)
    oa = oracle.check_batch(a_rows, a_meta)
    ob = oracle.check_batch(b_rows, b_meta)
    with ise_amd.Context(capacity=4096, stride=32) as ctx:
        for rep in range(3):
            ctx.use_graphs(True)
            ga = ctx.check_batch(a_rows, a_meta)
            ctx.use_graphs(False)
            gb = ctx.check_batch(b_rows, b_meta)
            for f in ("verdict", "fail_row", "perm_row", "perm_col", "group_order", "orbits"):
                assert np.array_equal(ga[f], oa[f]), (rep, f)
                assert np.array_equal(gb[f], ob[f]), (rep, f)
        graphs, replays = ctx.graph_stats()
    assert graphs == 1 and replays == 3
