"""GPU parity of the column-set search: when no column permutation is asked
for, the kernel keeps the set of source columns of every column orbit instead
of the permutation itself.  Every other result must equal the oracle's (the
reference's scalar checker restated) field by field: verdict, witness row
permutation, fail row and column, rowToChange, |Aut|, row orbits, generators
and the search counters.  Candidates not in orbit form take the replayed
search inside the same kernel."""
import ctypes

import numpy as np
import pytest

import ise_amd
from _helpers import check_witness, col_sorted, golden, unpack

pytestmark = pytest.mark.gpu

FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "group_order_hi", "perm_row",
          "orbits", "n_gens", "counters", "gens")


def assert_same(gpu, orc, gen_cap, what):
    for f in FIELDS[:-1]:
        a, b = gpu[f], orc[f]
        eq = (a == b) if a.ndim == 1 else (a == b).all(axis=tuple(range(1, a.ndim)))
        if not eq.all():
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"{what}: field {f} differs at candidate {i}: gpu={a[i]} oracle={b[i]}")
    assert "perm_col" not in gpu
    if gen_cap:
        for i in range(len(gpu["verdict"])):
            ng = min(int(orc["n_gens"][i]), gen_cap)
            assert np.array_equal(gpu["gens"][i, :ng], orc["gens"][i, :ng]), f"{what}: gens {i}"


def run(gpu_ctx, oracle, rows, meta, what, cells=None, gen_cap=8, group=0):
    """Counted launch (every row the walk compares is counted, so no row is
    skipped) and the launch without counters, which moves past the rows a
    level's first orbits refute without comparing them again: both must give
    the oracle's results."""
    gpu_ctx.set_launch(group_lanes=group)
    try:
        g = gpu_ctx.check_batch(rows, meta, cells=cells, gen_cap=gen_cap, fields=FIELDS)
        u = gpu_ctx.check_batch(rows, meta, cells=cells, gen_cap=gen_cap,
                                fields=tuple(f for f in FIELDS if f != "counters"))
    finally:
        gpu_ctx.set_launch()
    o = oracle.check_batch(rows, meta, cells=cells, gen_cap=gen_cap)
    assert (g["verdict"] <= 1).all()
    assert_same(g, o, gen_cap, what)
    u["counters"] = o["counters"]
    assert_same(u, o, gen_cap, what + " (no counters)")
    return g


@pytest.mark.parametrize("v,k,lam", [(7, 3, 1), (9, 3, 1), (13, 3, 1), (9, 3, 2), (15, 7, 3), (16, 6, 2)])
@pytest.mark.parametrize("group", [0, 64])
def test_enumerator_candidates(gpu_ctx, oracle, v, k, lam, group):
    print(f"case sets enumerator {(v, k, lam)} group={group}", flush=True)
    _, _, _, (rows, meta) = oracle.enumerate(v, k, lam, capture=(1 << 16, 32, 1, v))
    run(gpu_ctx, oracle, rows, meta, f"enum {(v, k, lam)}", group=group)


@pytest.mark.parametrize("with_cells", [False, True])
@pytest.mark.parametrize("v,k,lam,group", [(15, 3, 1, 0), (15, 3, 1, 64), (13, 3, 1, 32), (9, 3, 1, 16),
                                           (19, 9, 4, 0), (16, 6, 2, 16), (25, 4, 1, 0)])
def test_random_extensions(gpu_ctx, oracle, with_cells, v, k, lam, group):
    print(f"case sets random {(v, k, lam)} cells={with_cells} group={group}", flush=True)
    b = ise_amd.design_params(v, k, lam)[1]
    rows, meta = ise_amd.random_extensions(v, k, lam, 1 << 13, 4, v, 0x5E7 + v + group)
    cells = None
    if with_cells:
        cells = np.zeros_like(rows)
        for i in range(len(rows)):
            d = int(meta[i]) & 0xff
            cells[i, :d] = oracle.cell_masks(rows[i, :d], d, b)
    g = run(gpu_ctx, oracle, rows, meta, "random", cells=cells, gen_cap=4, group=group)
    for i in np.nonzero(g["verdict"] == 0)[0][:100]:
        d = int(meta[i]) & 0xff
        check_witness(unpack(rows[i, :d], b), g["perm_row"][i], int(g["fail_row"][i]))


def test_sampled_2_25_4_1_stream(gpu_ctx, oracle):
    """Beam-sampled enumerator candidates of 2-(25,4,1) at depth 8-16, the
    candidates that carry the long searches of the bench stream."""
    print("case sets sampled 2-(25,4,1)", flush=True)
    rows, meta = oracle.enumerate_sampled(25, 4, 1, 8, 0x25041, 8, 16, stride=32, cap=1 << 18)[:2]
    run(gpu_ctx, oracle, rows, meta, "sampled 25_4_1", gen_cap=4)


def test_edge_shapes_and_unsorted_matrices(gpu_ctx, oracle):
    """Orbit-form matrices of extreme shapes, and random matrices whose
    columns are not in orbit form (those take the replayed search)."""
    print("case sets edge shapes", flush=True)
    rng = np.random.default_rng(11)
    rows, meta = [], []
    for d, b in [(1, 1), (1, 64), (32, 1), (2, 64), (32, 64), (20, 64), (32, 40), (16, 16), (17, 16), (6, 9)]:
        for sort in (True, False):
            for _ in range(6):
                mat = (rng.random((d, b)) < 0.5).astype(np.uint8)
                if sort:
                    mat = col_sorted(mat)
                w = np.zeros(32, np.uint64)
                w[:d] = ise_amd.pack_rows(mat)
                rows.append(w)
                meta.append(ise_amd.pack_meta(d, b, int(rng.integers(0, max(1, d // 2))), int(rng.integers(0, 2)) | 4))
    run(gpu_ctx, oracle, np.array(rows), np.array(meta, np.uint32), "edges", gen_cap=4)


def test_full_matrices_row_permutations(gpu_ctx, oracle):
    """Complete designs with t_saveRowPermutations: |Aut| and row generators."""
    print("case sets printed matrices", flush=True)
    rows, meta, auts = [], [], []
    for block in golden("printed_bibds.json"):
        for m in block["matrices"]:
            mat = np.array([[int(c) for c in r] for r in m["rows"]], np.uint8)
            w = np.zeros(32, np.uint64)
            w[:mat.shape[0]] = ise_amd.pack_rows(mat)
            rows.append(w)
            meta.append(ise_amd.pack_meta(mat.shape[0], mat.shape[1], 0, 1 | 4))
            auts.append(m["aut"])
    g = run(gpu_ctx, oracle, np.array(rows), np.array(meta, np.uint32), "printed", gen_cap=8)
    assert (g["group_order"] == np.array(auts, np.uint64)).all()


def test_collect_refuses_unreported_columns(gpu_ctx):
    print("case sets collect refuses perm_col", flush=True)
    L = gpu_ctx._L
    rows, meta = ise_amd.random_extensions(13, 3, 1, 64, 4, 13, 5)
    gpu_ctx._check(L.ise_report_columns(gpu_ctx.h, 0))
    try:
        gpu_ctx._check(L.ise_submit(gpu_ctx.h, len(meta), gpu_ctx.stride, rows.ctypes.data, None,
                                    meta.ctypes.data, 0))
        r = ise_amd._abi.alloc_results(len(meta))
        s = ise_amd._abi.results_struct(r)
        with pytest.raises(ise_amd.ISEError) as e:
            gpu_ctx._check(L.ise_collect(gpu_ctx.h, ctypes.byref(s)))
        assert e.value.code == ise_amd.ERR_STATE
        r = ise_amd._abi.alloc_results(len(meta), fields=("verdict",))
        s = ise_amd._abi.results_struct(r)
        gpu_ctx._check(L.ise_collect(gpu_ctx.h, ctypes.byref(s)))
    finally:
        gpu_ctx._check(L.ise_report_columns(gpu_ctx.h, 1))


def test_device_path_without_columns(gpu_ctx, oracle):
    """HBM-resident batch with no perm_col output (the bench's configuration)."""
    print("case sets device path", flush=True)
    rows, meta = oracle.enumerate_sampled(25, 4, 1, 4, 0x2541, 8, 16, stride=16, cap=1 << 16)[:2]
    r2, m2 = ise_amd.random_extensions(25, 4, 1, 4096, 8, 16, 0x2541, stride=16)
    rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    fields = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "orbits", "counters")
    dev, s = gpu_ctx.device_results(n, fields)
    gpu_ctx.check_batch_device(n, 16, d_rows, d_meta, s, d_head)
    gpu_ctx.sync()
    rows32 = np.zeros((n, 32), np.uint64)
    rows32[:, :16] = rows
    o = oracle.check_batch(rows32, meta)
    for f in fields:
        assert np.array_equal(dev[f].to_host(), o[f]), f
    for b in list(dev.values()) + [d_rows, d_meta, d_head]:
        b.free()


@pytest.mark.parametrize("first,second", [(0, 0), (1, 0), (1, 8), (7, 64), (64, 4096)])
@pytest.mark.parametrize("columns", [False, True])
def test_deferred_long_searches_same_results(gpu_ctx, oracle, first, second, columns):
    """Launches in passes (searches past `first`, then `second` permutations
    deferred and rerun from the start) give the oracle's results, with and
    without the column permutation."""
    print(f"case deferral first={first} second={second} columns={columns}", flush=True)
    _, _, _, (r1, m1) = oracle.enumerate(16, 6, 2, capture=(1 << 14, 32, 1, 16))
    r2, m2 = ise_amd.random_extensions(19, 9, 4, 4096, 5, 19, 0xDEF)
    r3, m3 = oracle.enumerate_sampled(25, 4, 1, 4, 0xDEF, 8, 16, stride=32, cap=1 << 15)[:2]
    rows, meta = np.concatenate([r1, r2, r3]), np.concatenate([m1, m2, m3])
    gpu_ctx.set_defer(first, second)
    try:
        fields = FIELDS + (("perm_col",) if columns else ())
        g = gpu_ctx.check_batch(rows, meta, gen_cap=4, fields=fields)
    finally:
        gpu_ctx.set_defer()
    o = oracle.check_batch(rows, meta, gen_cap=4)
    if columns:
        assert np.array_equal(g.pop("perm_col"), o["perm_col"])
    assert_same(g, o, 4, f"deferral {first}")


@pytest.mark.parametrize("deepest_first", [True, False])
def test_schedule_after_smaller_batches(gpu_ctx, oracle, deepest_first):
    """Deepest-first schedules over consecutive batches of growing size and
    shuffled depths (every wave of the schedule kernels sees several depths)
    leave no candidate out and give the oracle's results."""
    print(f"case schedule deepest_first={deepest_first}", flush=True)
    gpu_ctx.set_schedule(deepest_first)
    try:
        for v, k, lam, n in ((7, 3, 1, 40), (9, 3, 2, 700), (16, 6, 2, 5000)):
            _, _, _, (rows, meta) = oracle.enumerate(v, k, lam, capture=(1 << 16, 32, 1, v))
            r2, m2 = ise_amd.random_extensions(v, k, lam, n, 2, v, 0x5C + n)
            rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
            idx = np.random.default_rng(n).permutation(len(meta))
            rows, meta = rows[idx], meta[idx]
            g = gpu_ctx.check_batch(rows, meta, gen_cap=4, fields=FIELDS)
            o = oracle.check_batch(rows, meta, gen_cap=4)
            assert_same(g, o, 4, f"schedule {(v, k, lam)}")
    finally:
        gpu_ctx.set_schedule(True)


def test_group_orders_past_64_bits(gpu_ctx, oracle):
    """|Aut| beyond 2^64 (rows of a few values: products of factorials up to
    32!) in both words, on the column-set search, the uncounted launch and the
    replayed search, equal to the closed form and to the oracle."""
    print("case sets group orders past 2^64", flush=True)
    from test_oracle import group_order_cases
    cases = group_order_cases()
    rows = np.zeros((len(cases), 32), np.uint64)
    meta = np.zeros(len(cases), np.uint32)
    for i, (mat, _) in enumerate(cases):
        rows[i, :mat.shape[0]] = ise_amd.pack_rows(mat)
        meta[i] = ise_amd.pack_meta(mat.shape[0], mat.shape[1], 0, 1)
    want = [aut for _, aut in cases]
    g = run(gpu_ctx, oracle, rows, meta, "group orders", gen_cap=4)
    full = gpu_ctx.check_batch(rows, meta)            # column permutation reported: replayed search
    for r in (g, full):
        got = [int(lo) | (int(hi) << 64) for lo, hi in zip(r["group_order"], r["group_order_hi"])]
        assert got == want
