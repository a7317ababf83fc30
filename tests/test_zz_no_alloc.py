"""The allocation-free entry point (ise_check_batch_device_noalloc) against
ise_check_batch_device, through the C ABI: two calls in a row on one context
(the second with a smaller batch, so the shared device scratch keeps the
first call's contents), both search forms, and the size limit."""
import ctypes

import numpy as np
import pytest

import ise_amd

pytestmark = pytest.mark.gpu

FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "orbits",
          "n_gens", "gens", "counters")


def batch(seed, n_rand):
    r1, m1 = ise_amd.random_extensions(15, 3, 1, n_rand, 4, 15, seed)
    r2, m2 = ise_amd.random_extensions(16, 6, 2, n_rand // 2, 5, 16, seed + 1)
    rows = np.concatenate([r1, r2])
    meta = np.concatenate([m1, m2])
    order = np.random.default_rng(seed).permutation(len(meta))
    return np.ascontiguousarray(rows[order]), np.ascontiguousarray(meta[order])


def run(ctx, rows, meta, fields, noalloc):
    n = len(meta)
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    dev, s = ctx.device_results(n, fields, gen_cap=4)
    if noalloc:
        ctx.check_batch_device_noalloc(n, rows.shape[1], d_rows, d_meta, s)
    else:
        d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
        ctx.check_batch_device(n, rows.shape[1], d_rows, d_meta, s, d_head)
        d_head.free()
    ctx.sync()
    out = {k: dev[k].to_host() for k in dev}
    for b in list(dev.values()) + [d_rows, d_meta]:
        b.free()
    return out


def same(a, b, fields, what):
    for f in fields:
        if f == "gens":
            for i in range(len(a["verdict"])):
                ng = min(int(b["n_gens"][i]), 4)
                assert np.array_equal(a["gens"][i, :ng], b["gens"][i, :ng]), (what, "gens", i)
        else:
            assert np.array_equal(a[f], b[f]), (what, f)


@pytest.mark.parametrize("columns", [False, True])
def test_noalloc_matches_device_entry_point(columns):
    fields = FIELDS + (("perm_col",) if columns else ())
    print(f"case noalloc columns={columns}", flush=True)
    with ise_amd.Context(capacity=1 << 12, stride=32) as ctx:
        for call, (seed, n) in enumerate(((0xA1, 40000), (0xA2, 3000))):
            rows, meta = batch(seed, n)
            print(f"case noalloc call {call}: {len(meta)} candidates", flush=True)
            ref = run(ctx, rows, meta, fields, noalloc=False)
            got = run(ctx, rows, meta, fields, noalloc=True)
            assert (ref["verdict"] <= 1).all()
            same(got, ref, fields, f"call {call}")


def test_noalloc_rejects_batches_above_the_limit():
    print("case noalloc limit", flush=True)
    with ise_amd.Context(capacity=1 << 10, stride=32) as ctx:
        # real (small) device buffers; the size check rejects the call before
        # any kernel is launched, so they are never read
        d_rows = ctx.device_array((1, 32), np.uint64)
        d_meta = ctx.device_array((1,), np.uint32)
        n = ise_amd.NOALLOC_MAX_CANDIDATES + 1
        res = ise_amd.ResultsStruct()
        rc = ctx._L.ise_check_batch_device_noalloc(ctx.h, n, 32, d_rows.ptr, None, d_meta.ptr,
                                                   ctypes.byref(res), None)
        assert rc == ise_amd.ERR_ARG        # no verdict array
        dev, s = ctx.device_results(1, ("verdict",))
        rc = ctx._L.ise_check_batch_device_noalloc(ctx.h, n, 32, d_rows.ptr, None, d_meta.ptr,
                                                   ctypes.byref(s), None)
        assert rc == ise_amd.ERR_CAPACITY
        assert b"exceed the maximum" in ctx._L.ise_last_error(ctx.h)
        for b in (dev["verdict"], d_rows, d_meta):
            b.free()
