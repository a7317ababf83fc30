"""GPU parity: the HIP canonicity kernel against the CPU oracle, field by field,
through the C ABI (libise_amd.so)."""
import ctypes

import numpy as np
import pytest

import ise_amd
from _helpers import (check_witness, col_sorted, counts_for, design_key, golden, hist,
                      random_orbit_form_batch, unpack)

pytestmark = pytest.mark.gpu

FIELDS_CMP = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "group_order_hi", "perm_row",
              "perm_col", "orbits", "n_gens", "counters")


def assert_same(gpu, orc, gen_cap=0, what=""):
    for f in FIELDS_CMP:
        a, b = gpu[f], orc[f]
        eq = (a == b) if a.ndim == 1 else (a == b).all(axis=tuple(range(1, a.ndim)))
        if not eq.all():
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"{what}: field {f} differs at candidate {i}: gpu={a[i]} oracle={b[i]}")
    if gen_cap:
        for i in range(len(gpu["verdict"])):
            ng = min(int(orc["n_gens"][i]), gen_cap)
            assert np.array_equal(gpu["gens"][i, :ng], orc["gens"][i, :ng]), f"{what}: gens {i}"


def stride32(rows):
    out = np.zeros((len(rows), 32), np.uint64)
    out[:, :rows.shape[1]] = rows
    return out


@pytest.mark.parametrize("v,k,lam", [(7, 3, 1), (9, 3, 1), (13, 3, 1), (9, 3, 2), (15, 7, 3), (16, 6, 2)])
@pytest.mark.parametrize("group", [0, 64])
def test_enumerator_candidates_parity(gpu_ctx, oracle, v, k, lam, group):
    print(f"case enumerator candidates {(v, k, lam)} group={group}", flush=True)
    _, _, _, cands = oracle.enumerate(v, k, lam, capture=(1 << 16, 32, 1, v))
    rows, meta = cands
    gpu_ctx.set_launch(group_lanes=group)
    try:
        g = gpu_ctx.check_batch(rows, meta, gen_cap=8)
    finally:
        gpu_ctx.set_launch()
    o = oracle.check_batch(rows, meta, gen_cap=8)
    assert (g["verdict"] <= 1).all()
    assert_same(g, o, 8, f"{(v, k, lam)}")


@pytest.mark.parametrize("with_cells", [False, True])
@pytest.mark.parametrize("v,k,lam,group", [(15, 3, 1, 0), (15, 3, 1, 64), (13, 3, 1, 32), (9, 3, 1, 16)])
def test_random_extensions_parity(gpu_ctx, oracle, with_cells, v, k, lam, group):
    print(f"case random {(v, k, lam)} cells={with_cells} group={group}", flush=True)
    b = ise_amd.design_params(v, k, lam)[1]
    rows, meta = ise_amd.random_extensions(v, k, lam, 1 << 13, 4, v, 0x15031 + group)
    cells = None
    if with_cells:
        cells = np.zeros_like(rows)
        for i in range(len(rows)):
            d = int(meta[i]) & 0xff
            cells[i, :d] = oracle.cell_masks(rows[i, :d], d, b)
    gpu_ctx.set_launch(group_lanes=group)
    try:
        g = gpu_ctx.check_batch(rows, meta, cells=cells, gen_cap=4)
    finally:
        gpu_ctx.set_launch()
    o = oracle.check_batch(rows, meta, cells=cells, gen_cap=4)
    assert_same(g, o, 4, "random")
    # the reported witness really makes the matrix larger
    for i in np.nonzero(g["verdict"] == 0)[0][:200]:
        d = int(meta[i]) & 0xff
        check_witness(unpack(rows[i, :d], b), g["perm_row"][i], int(g["fail_row"][i]))


@pytest.mark.parametrize("columns", [True, False])
def test_fixed_rows_parity(gpu_ctx, oracle, columns):
    """lenStab > 0 (rows held in place, the reference's stabiliserLengthExt):
    random orbit-form matrices, including those where the walk falls back past
    the fixed rows after a refuted prefix exhausts its tail
    (GroupOrder.h:183-186), and random extensions of 2-(15,3,1) and
    2-(13,3,1) with a random number of fixed rows; every field against the
    oracle, replayed search (columns=True) and column-set search."""
    print(f"case fixed rows columns={columns}", flush=True)
    _, r1, m1 = random_orbit_form_batch(0x51DE, 3600, stride=32)
    rng = np.random.default_rng(0xF1)
    parts = [(r1, m1)]
    for v, k, lam in ((15, 3, 1), (13, 3, 1)):
        r, m = ise_amd.random_extensions(v, k, lam, 4096, 4, v, 0xF15 + v)
        d = (m & 0xff).astype(np.int64)
        f = (rng.random(len(m)) * d).astype(np.uint32)
        parts.append((r, m | (f << 16)))
    rows = np.concatenate([p[0] for p in parts])
    meta = np.concatenate([p[1] for p in parts])
    assert ((meta >> 16) & 0xff).astype(bool).sum() > 3000
    fields = None if columns else tuple(f for f in FIELDS_CMP if f != "perm_col") + ("gens",)
    g = gpu_ctx.check_batch(rows, meta, gen_cap=4, fields=fields)
    o = oracle.check_batch(rows, meta, gen_cap=4)
    if columns:
        assert_same(g, o, 4, "fixed rows")
    else:
        for f in fields:
            if f == "gens":
                continue
            a, b = g[f], o[f]
            eq = (a == b) if a.ndim == 1 else (a == b).all(axis=tuple(range(1, a.ndim)))
            assert eq.all(), (f, int(np.nonzero(~eq)[0][0]))
        for i in range(len(meta)):
            ng = min(int(o["n_gens"][i]), 4)
            assert np.array_equal(g["gens"][i, :ng], o["gens"][i, :ng]), ("gens", i)


def test_group_too_small_is_rejected(gpu_ctx):
    rows, meta = ise_amd.random_extensions(15, 3, 1, 16, 4, 15, 3)
    gpu_ctx.set_launch(group_lanes=32)
    try:
        with pytest.raises(ise_amd.ISEError) as e:
            gpu_ctx.check_batch(rows, meta)
        assert e.value.code == ise_amd.ERR_ARG
    finally:
        gpu_ctx.set_launch()


@pytest.mark.parametrize("v,k,lam", [(15, 7, 3), (16, 6, 2), (19, 9, 4)])
def test_mixed_parameter_sets_parity(gpu_ctx, oracle, v, k, lam):
    print(f"case random {(v, k, lam)}", flush=True)
    rows, meta = ise_amd.random_extensions(v, k, lam, 4096, 5, v, 0xC3 + v)
    g = gpu_ctx.check_batch(rows, meta, gen_cap=4)
    o = oracle.check_batch(rows, meta, gen_cap=4)
    assert_same(g, o, 4, f"{(v, k, lam)}")


def test_2_25_4_1_parity(gpu_ctx, oracle):
    print("case random 2-(25,4,1) depth 8-16", flush=True)
    rows, meta = ise_amd.random_extensions(25, 4, 1, 4096, 8, 16, 0x25041)
    g = gpu_ctx.check_batch(rows, meta, gen_cap=4)
    o = oracle.check_batch(rows, meta, gen_cap=4)
    assert_same(g, o, 4, "25_4_1")


ENUM_CONFIGS = [(7, 3, 1), (13, 3, 1), (9, 3, 2), (15, 7, 3), (16, 6, 2), (15, 3, 1), (19, 9, 4)]


@pytest.mark.parametrize("v,k,lam", ENUM_CONFIGS)
def test_gpu_enumeration_design_lists(gpu_ctx, oracle, v, k, lam):
    """Full enumeration with every level checked on the GPU: published counts,
    |Aut| histogram, and the design list bit-identical to the oracle's."""
    print(f"case enumeration {(v, k, lam)}", flush=True)
    des, auts, stats, levels = gpu_ctx.enumerate_bibd(v, k, lam)
    ref = counts_for(v, k, lam)
    assert len(des) == ref["designs"]
    assert hist(auts) == ref["aut_hist"]
    odes, oauts, _, _ = oracle.enumerate(v, k, lam)
    assert sorted(zip(map(design_key, des), map(int, auts))) == sorted(zip(map(design_key, odes), map(int, oauts)))
    assert stats["candidates"] == int(levels.sum())


def test_printed_matrices_on_gpu(gpu_ctx):
    print("case printed matrices", flush=True)
    rows, meta, auts = [], [], []
    for block in golden("printed_bibds.json"):
        for m in block["matrices"]:
            mat = np.array([[int(c) for c in r] for r in m["rows"]], np.uint8)
            w = np.zeros(32, np.uint64)
            w[:mat.shape[0]] = ise_amd.pack_rows(mat)
            rows.append(w)
            meta.append(ise_amd.pack_meta(mat.shape[0], mat.shape[1], 0, 1 | 4))
            auts.append(m["aut"])
    r = gpu_ctx.check_batch(np.array(rows), np.array(meta, np.uint32))
    assert (r["verdict"] == 1).all()
    assert (r["group_order"] == np.array(auts, np.uint64)).all()


def test_edge_shapes(gpu_ctx, oracle):
    print("case edge shapes", flush=True)
    rng = np.random.default_rng(5)
    rows, meta = [], []
    for d, b in [(1, 1), (1, 64), (32, 1), (2, 64), (32, 64), (20, 64), (32, 40), (16, 16), (17, 16)]:
        for _ in range(8):
            mat = (rng.random((d, b)) < 0.5).astype(np.uint8)
            mat = col_sorted(mat)
            w = np.zeros(32, np.uint64)
            w[:d] = ise_amd.pack_rows(mat)
            rows.append(w)
            meta.append(ise_amd.pack_meta(d, b, 0, int(rng.integers(0, 2)) | 4))
    rows, meta = np.array(rows), np.array(meta, np.uint32)
    g = gpu_ctx.check_batch(rows, meta, gen_cap=4)
    o = oracle.check_batch(rows, meta, gen_cap=4)
    assert_same(g, o, 4, "edges")


def test_invalid_records_rejected(gpu_ctx):
    rows = np.zeros((1, 32), np.uint64)
    for bad in [ise_amd.pack_meta(0, 5), ise_amd.pack_meta(33, 5), ise_amd.pack_meta(3, 65),
                ise_amd.pack_meta(3, 0), ise_amd.pack_meta(3, 5, 3)]:
        with pytest.raises(ise_amd.ISEError) as e:
            gpu_ctx.check_batch(rows, np.array([bad], np.uint32))
        assert e.value.code == ise_amd.ERR_ARG


def test_reference_shaped_facades(oracle):
    print("case facades", flush=True)
    des, auts, _, _ = oracle.enumerate(13, 3, 1)
    chk = ise_amd.CanonicityChecker(13, 26)
    for d, a in zip(des, auts):
        mat = unpack(d, 26)
        assert chk.TestCanonicity(13, mat, ise_amd.SAVE_ROW_PERMUTATIONS)
        assert chk.groupOrder() == int(a)
        # another row order, columns re-sorted into orbit form
        other = col_sorted(mat[::-1])
        same = bool((other == mat).all())
        assert chk.TestCanonicity(13, other, ise_amd.SAVE_ROW_TO_CHANGE) == same
        if not same:
            assert chk.rowToChange() < 13
    gpu = ise_amd.GPUCanonChecker(13, 26, batch=128)
    for d in des:
        gpu.assignChecker(unpack(d, 26))
    assert gpu.numAssigned() == len(des)
    r = gpu.LaunchCanonicityTesting()
    assert (r["verdict"] == 1).all()
    assert sorted(int(x) for x in r["group_order"]) == sorted(int(x) for x in auts)


def test_device_resident_entry_point(gpu_ctx, oracle):
    """ise_check_batch_device / _timed on HBM-resident buffers."""
    print("case device entry point", flush=True)
    rows, meta = ise_amd.random_extensions(15, 3, 1, 4096, 4, 15, 77, stride=16)
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    fields = ("verdict", "fail_row", "perm_row", "group_order", "counters")
    dev, s = gpu_ctx.device_results(n, fields)
    gpu_ctx.check_batch_device(n, 16, d_rows, d_meta, s, d_head)
    gpu_ctx.sync()
    rows32 = np.zeros((n, 32), np.uint64)
    rows32[:, :16] = rows
    o = oracle.check_batch(rows32, meta)
    for f in fields:
        assert np.array_equal(dev[f].to_host(), o[f]), f
    ms = gpu_ctx.time_batch_device(n, 16, d_rows, d_meta, s, d_head, 3)
    assert (ms > 0).all()
    for f in fields:
        assert np.array_equal(dev[f].to_host(), o[f]), f


def test_full_size_stress_batch_properties(gpu_ctx, oracle):
    """The 2^20 stress batch of 2-(15,3,1): deterministic, oracle-exact on a
    seeded sample, and every reported witness makes its matrix larger."""
    print("case 2^20 stress batch", flush=True)
    n = 1 << 20
    rows, meta = ise_amd.random_extensions(15, 3, 1, n, 4, 15, 0x15031)
    fields = ("verdict", "fail_row", "perm_row", "group_order")
    a = gpu_ctx.check_batch(rows, meta, fields=fields)
    b = gpu_ctx.check_batch(rows, meta, fields=fields)
    for f in fields:
        assert np.array_equal(a[f], b[f])
    assert (a["verdict"] <= 1).all()
    idx = np.random.default_rng(1).choice(n, 4096, replace=False)
    o = oracle.check_batch(rows[idx], meta[idx])
    for f in fields:
        assert np.array_equal(a[f][idx], o[f]), f
    for i in idx[:300]:
        if a["verdict"][i] == 0:
            d = int(meta[i]) & 0xff
            check_witness(unpack(rows[i, :d], 35), a["perm_row"][i], int(a["fail_row"][i]))
