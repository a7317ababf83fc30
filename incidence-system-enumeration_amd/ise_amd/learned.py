"""The learned row order: witness logs of the plain search -> pair features
-> training of the row-order network on the GPU -> per-candidate order
tables for the witness hunt (ISE_ORDER_LEARNED).

Everything runs through the HIP library: the plain checker writes the
witness log (verdict, fail_row, perm_row), ise_mlp_witness_labels turns it
into per-pair labels, ise_mlp_features computes the pair features on the
device, and ise_mlp_fit trains with the library's forward, backward and
Adam kernels.  Verdicts never depend on the network: the hunt only looks for
a witness, and the plain search decides every candidate it leaves open.
"""
import os

import numpy as np

from . import (MAX_ROWS, SAVE_ROW_PERMUTATIONS, SAVE_ROW_TO_CHANGE, MLP_FEATURES, MLP_PARAMS,
               RowOrderModel, random_extensions, witness_labels)

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "row_order_weights.npy")
LEAD_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lead_order_weights.npy")
# The lead order's training stream: beam-sampled 2-(25,4,1) enumerations
# with seeds of their own (the benchmark's stream uses 0x25041; candidates
# that occur in both are dropped from training).
LEAD_TRAIN_SEED = (0x7777, 0x7778)
LEAD_GATE = 16

# Training streams: the plain enumerator's candidates and seeded random
# extensions of the three enumeration parameter families.
TRAIN_SETS = (
    ("enum", (13, 3, 1), (1, 13)),
    ("enum", (15, 3, 1), (4, 15)),
    ("random", (15, 3, 1), (4, 15)),
    ("enum", (15, 7, 3), (5, 15)),
    ("enum", (16, 6, 2), (5, 16)),
    ("enum", (19, 9, 4), (5, 19)),
    ("random", (15, 7, 3), (5, 15)),
    ("random", (16, 6, 2), (5, 16)),
    ("random", (19, 9, 4), (5, 19)),
)


def plain_witness_log(ctx, rows, meta):
    """Plain-order results (verdict, fail_row, perm_row) of a host batch."""
    return ctx.check_batch(rows, meta, fields=("verdict", "fail_row", "perm_row", "counters"))


def training_stream(ctx, per_set=4000, seed=1):
    """Non-canonical candidates with their plain witnesses, from every
    training set (at most per_set each): rows, meta, verdict, fail_row,
    perm_row."""
    rng = np.random.default_rng(seed)
    parts = []
    for kind, (v, k, lam), (dmin, dmax) in TRAIN_SETS:
        if kind == "enum":
            rows, meta = ctx.enumerator_candidates(v, k, lam, dmin, dmax, cap=1 << 20)
        else:
            rows, meta = random_extensions(v, k, lam, 4 * per_set, dmin, dmax, seed=seed + v * 1000 + k,
                                           stride=ctx.stride)
        res = plain_witness_log(ctx, rows, meta)
        rej = np.nonzero(res["verdict"] == 0)[0]
        if len(rej) > per_set:
            rej = np.sort(rng.choice(rej, per_set, replace=False))
        parts.append((rows[rej], meta[rej], res["verdict"][rej], res["fail_row"][rej], res["perm_row"][rej]))
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(5))


def pair_dataset(ctx, model, rows, meta, verdict, fail_row, perm_row, od, seed=2):
    """Device features X [M][16] (bf16 bits), labels y [M] and weights w [M]
    of every labelled pair, shuffled, M a multiple of 16."""
    n = len(meta)
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_X = ctx.device_array((n * od * od, MLP_FEATURES), np.uint16)
    model.features(n, rows.shape[1], d_rows, d_meta, od, d_X)
    X = d_X.to_host()
    del d_X, d_rows, d_meta
    y, w = witness_labels(meta, verdict, fail_row, perm_row, od)
    keep = np.nonzero(w.ravel() > 0)[0]
    rng = np.random.default_rng(seed)
    keep = keep[rng.permutation(len(keep))]
    M = len(keep) // 16 * 16
    keep = keep[:M]
    X, y, w = X[keep], y.ravel()[keep], w.ravel()[keep]
    return (ctx.device_array(X.shape, np.uint16, X), ctx.device_array(y.shape, np.float32, y),
            ctx.device_array(w.shape, np.float32, w), M)


def train(ctx, od=MAX_ROWS, per_set=4000, steps=3000, batch=4096, lr=2e-3, seed=1):
    """Train a fresh model on the witness logs; returns (model, loss curve)."""
    model = RowOrderModel(ctx, max_rows=batch, seed=seed)
    stream = training_stream(ctx, per_set=per_set, seed=seed)
    d_X, d_y, d_w, M = pair_dataset(ctx, model, *stream, od=od, seed=seed + 1)
    curve = model.fit(d_X, d_y, d_w, M, batch, steps, lr)
    return model, curve


def lead_labels(meta, verdict, perm_row, counters, n_gens, gate):
    """Labels of the rows of position 0 from plain witness logs: for a
    non-canonical candidate the walk needed more than `gate` permutations for
    (and met no automorphism on), the walk tried the rows of position 0 in
    increasing order and exhausted every row before its witness's first row
    without a witness, so row p0 = perm_row[0] is labelled 1, rows j < p0 are
    labelled 0 and rows after p0 are unknown (weight 0).  Returns y, w
    [n][16]."""
    n = len(meta)
    d = (meta & 0xff).astype(np.int64)
    take = (verdict == 0) & (counters[:, 0] > gate) & (n_gens == 0) & (d >= 2) & (d <= 16)
    p0 = perm_row[:, 0].astype(np.int64)
    j = np.arange(16)[None, :]
    w = (take[:, None] & (j <= p0[:, None])).astype(np.float32)
    y = (w > 0) & (j == p0[:, None])
    return y.astype(np.float32), w


def lead_training_stream(ctx, exclude_rows=None, gate=LEAD_GATE, seed=LEAD_TRAIN_SEED, beam=64, exact=True):
    """Witness logs of beam-sampled 2-(25,4,1) enumerations (depths 8-16; one
    seed or several):
    rows and meta of the candidates lead_labels keeps (non-canonical, past
    the gate, no automorphism met), with their exact labels
    (lead_exact_labels) or, with exact=False, the witness-prefix labels."""
    seeds = seed if isinstance(seed, (list, tuple)) else (seed,)
    parts = [ctx.enumerate_sampled(25, 4, 1, beam, sd, 8, 16, stride=16, cap=1 << 21)[:2] for sd in seeds]
    rows = np.concatenate([p[0] for p in parts])
    meta = np.concatenate([p[1] for p in parts])
    # one copy of each candidate
    _, first = np.unique(rows, axis=0, return_index=True)
    first = np.sort(first)
    rows, meta = rows[first], meta[first]
    if exclude_rows is not None:
        seen = {r.tobytes() for r in exclude_rows}
        keep = np.array([r.tobytes() not in seen for r in rows], bool)
        rows, meta = rows[keep], meta[keep]
    padded = np.zeros((len(meta), ctx.stride), np.uint64)
    padded[:, :rows.shape[1]] = rows
    res = ctx.check_batch(padded, meta, fields=("verdict", "perm_row", "counters", "n_gens"))
    y, w = lead_labels(meta, res["verdict"], res["perm_row"], res["counters"], res["n_gens"], gate)
    sel = np.nonzero(w.sum(1) > 0)[0]
    rows, meta = rows[sel], meta[sel]
    if exact:
        y, w = lead_exact_labels(ctx, rows, meta, gate)
    else:
        y, w = y[sel], w[sel]
    return rows, meta, y, w


def lead_exact_labels(ctx, rows, meta, gate=LEAD_GATE):
    """Labels of every row of position 0 from the checker's own witnesses: for
    each row j the lead-order kernel runs with j first (the other rows after
    it in increasing order), and row j is labelled 1 when the witness it
    returns starts with j, i.e. when j's subtree holds a witness, else 0.
    The witness-prefix labels of plain logs (lead_labels) see only the rows
    up to the plain witness's first row, so they teach how likely a row is
    right given that every lower row was wrong, and rank the last rows too
    high; these see every row.  Returns y, w [n][16]."""
    n, stride = rows.shape
    d = (meta & 0xff).astype(np.int64)
    model = RowOrderModel(ctx, max_rows=16, seed=1)
    model.set_gate(gate)
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_head = ctx.device_array((256,), np.uint32)
    dev, res = ctx.device_results(n, ("verdict", "perm_row"))
    y = np.zeros((n, 16), np.float32)
    for j in range(16):
        lead = np.tile(np.arange(16, dtype=np.uint8), (n, 1))
        lead[:, 0] = j
        lead[:, 1:] = [[x for x in range(16) if x != j]] * n
        d_in = ctx.device_array(lead.shape, np.uint8, lead)
        model.check_lead(n, stride, d_rows, d_meta, res, d_head, d_in)
        v, pr = dev["verdict"].to_host(), dev["perm_row"].to_host()
        y[:, j] = ((v == 0) & (pr[:, 0] == j) & (j < d)).astype(np.float32)
        d_in.free()
    model.close()
    w = (np.arange(16)[None, :] < d[:, None]).astype(np.float32)
    return y, w


def train_lead(ctx, rows, meta, y, w, steps=3000, batch=4096, lr=1e-3, seed=1):
    """Train a fresh network on the labelled rows of position 0 with the
    library's MFMA forward, backward and Adam kernels; returns (model, loss
    curve, labelled rows)."""
    model = RowOrderModel(ctx, max_rows=batch, seed=seed)
    n, stride = rows.shape
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_X = ctx.device_array((n, 16, 16), np.uint16)
    model.lead_features(n, stride, d_rows, d_meta, d_X)
    X = d_X.to_host().reshape(n * 16, 16)
    del d_X, d_rows, d_meta
    keep = np.nonzero(w.ravel() > 0)[0]
    keep = keep[np.random.default_rng(seed).permutation(len(keep))]
    M = len(keep) // 16 * 16
    keep = keep[:M]
    Xk, yk, wk = X[keep], y.ravel()[keep], w.ravel()[keep]
    d_X = ctx.device_array(Xk.shape, np.uint16, Xk)
    d_y = ctx.device_array(yk.shape, np.float32, yk)
    d_w = ctx.device_array(wk.shape, np.float32, wk)
    curve = model.fit(d_X, d_y, d_w, M, batch, steps, lr)
    return model, curve, M


def autotune_lead(ctx, model, n, stride, d_rows, d_meta, res_struct, d_head, waves=(1, 2, 4, 8), repeats=2):
    """Waves per workgroup of the lead-order launch: each of `waves` timed on
    the batch (after one untimed launch), the fastest kept in the context's
    launch settings.  Returns (best waves per workgroup, ms per setting)."""
    saved = ctx.get_launch()
    table = []
    for wpb in waves:
        ctx.set_launch(group_lanes=64, waves_per_block=wpb, blocks=saved["blocks"], max_perms=saved["max_perms"])
        model.time_lead(n, stride, d_rows, d_meta, res_struct, d_head, 1)
        table.append(float(np.mean(model.time_lead(n, stride, d_rows, d_meta, res_struct, d_head, repeats))))
    best = waves[int(np.argmin(table))]
    ctx.set_launch(group_lanes=saved["group_lanes"], waves_per_block=best, blocks=saved["blocks"],
                   max_perms=saved["max_perms"])
    return best, table


def compare_lead(ctx, model, rows, meta, gate=LEAD_GATE, repeats=3):
    """Plain order against the lead order on one device batch: kernel times,
    verdicts (must agree), rows compared and mean permutations before
    rejection from counted launches."""
    n, stride = rows.shape
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_head = ctx.device_array((256,), np.uint32)
    tdev, tres = ctx.device_results(n, ("verdict",))
    cdev, cres = ctx.device_results(n, ("verdict", "counters"))
    ms_plain = ctx.time_batch_device(n, stride, d_rows, d_meta, tres, d_head, repeats)
    ctx.time_batch_device(n, stride, d_rows, d_meta, cres, d_head, 1)
    verdict, counters = cdev["verdict"].to_host(), cdev["counters"].to_host()
    model.set_gate(gate)
    ms_lead = model.time_lead(n, stride, d_rows, d_meta, tres, d_head, repeats)
    lv = tdev["verdict"].to_host()
    model.check_lead(n, stride, d_rows, d_meta, cres, d_head)
    lc = cdev["counters"].to_host()
    model.set_gate()
    if not (np.array_equal(lv, verdict) and np.array_equal(cdev["verdict"].to_host(), verdict)):
        raise AssertionError("the lead order changed a verdict")
    rej = verdict == 0
    return {"n": int(n), "noncanonical": int(rej.sum()), "gate": int(gate),
            "plain_ms": float(np.min(ms_plain)), "lead_ms": float(np.min(ms_lead)),
            "plain_rows": float(counters[:, 1].sum()), "lead_rows": float(lc[:, 1].sum()),
            "plain_mean_perms_to_reject": float(counters[rej, 0].mean()),
            "lead_mean_perms_to_reject": float(lc[rej, 0].mean()),
            "plain_mean_rows_to_reject": float(counters[rej, 1].mean()),
            "lead_mean_rows_to_reject": float(lc[rej, 1].mean())}


def load(ctx, path=WEIGHTS):
    """A model with the stored trained parameters."""
    model = RowOrderModel(ctx, max_rows=4096)
    p = np.load(path, allow_pickle=False)
    if p.shape != (MLP_PARAMS,):
        raise ValueError("stored parameters have the wrong size")
    model.params = p
    return model


def c5_stream(ctx, n_random=1 << 20, beam=64, seed=0x25041, dmin=8, dmax=16):
    """The 2-(25,4,1) stream at depths 8-16: beam-sampled enumerator output
    followed by seeded random valid extensions."""
    er, em, _ = ctx.enumerate_sampled(25, 4, 1, beam, seed, dmin, dmax, stride=dmax, cap=1 << 21)
    rr, rm = random_extensions(25, 4, 1, n_random, dmin, dmax, seed, stride=dmax)
    return np.concatenate([er, rr]), np.concatenate([em, rm]), len(em)


def compare_orders(ctx, model, rows, meta, od, budgets=(8, 32, 128), repeats=3):
    """Plain vs learned order on one device batch: verdicts (must agree),
    mean permutations and rows compared before rejection (from a counting
    launch), and kernel times (from launches without counters)."""
    n, stride = rows.shape
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_head = ctx.device_array((256,), np.uint32)
    tdev, tres = ctx.device_results(n, ("verdict",))
    cdev, cres = ctx.device_results(n, ("verdict", "counters"))
    ms_plain = ctx.time_batch_device(n, stride, d_rows, d_meta, tres, d_head, repeats)
    ctx.time_batch_device(n, stride, d_rows, d_meta, cres, d_head, 1)
    verdict = cdev["verdict"].to_host()
    counters = cdev["counters"].to_host()
    if not np.array_equal(tdev["verdict"].to_host(), verdict):
        raise AssertionError("counting changed a verdict")
    rej = verdict == 0
    out = {"n": int(n), "noncanonical": int(rej.sum()), "plain_ms": float(np.min(ms_plain)),
           "plain_mean_perms_to_reject": float(counters[rej, 0].mean()),
           "plain_mean_rows_to_reject": float(counters[rej, 1].mean()),
           "learned": []}
    d_order = ctx.device_array((n, od, od), np.uint8)
    for budget in budgets:
        ms, oms = model.time_check(n, stride, d_rows, d_meta, od, d_order, budget, tres, d_head, repeats)
        model.time_check(n, stride, d_rows, d_meta, od, d_order, budget, cres, d_head, 1)
        got_v, got_c = cdev["verdict"].to_host(), cdev["counters"].to_host()
        if not (np.array_equal(got_v, verdict) and np.array_equal(tdev["verdict"].to_host(), verdict)):
            raise AssertionError("learned order changed a verdict")
        out["learned"].append({
            "budget": int(budget), "ms": float(np.min(ms)), "order_ms": float(np.min(oms)),
            "mean_perms_to_reject": float(got_c[rej, 0].mean()),
            "mean_rows_to_reject": float(got_c[rej, 1].mean()),
            "rows_total_ratio": float(counters[:, 1].sum() / max(got_c[:, 1].sum(), 1))})
    return out


def main_lead(argv=None):
    """Train the lead-order network and compare it with the plain order on the
    benchmark's 2-(25,4,1) stream."""
    import argparse
    import json
    import time
    from . import Context
    ap = argparse.ArgumentParser(description="train the lead-order network on plain witness logs of a "
                                             "2-(25,4,1) enumeration and compare it with the plain order")
    ap.add_argument("--out", default=LEAD_WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--gate", type=int, default=LEAD_GATE)
    ap.add_argument("--eval-random", type=int, default=1 << 20)
    args = ap.parse_args(argv)
    ctx = Context(capacity=1 << 18, stride=32)
    t0 = time.time()
    rows, meta, n_enum = c5_stream(ctx, n_random=args.eval_random)
    print(f"benchmark stream: {len(meta)} candidates ({n_enum} enumerator), {time.time() - t0:.1f}s", flush=True)
    t0 = time.time()
    tr = lead_training_stream(ctx, exclude_rows=rows[:n_enum], gate=args.gate)
    print(f"training candidates {len(tr[1])}, {time.time() - t0:.1f}s", flush=True)
    t1 = time.time()
    model, curve, M = train_lead(ctx, *tr, steps=args.steps, batch=args.batch, lr=args.lr)
    t2 = time.time()
    np.save(args.out, model.params)
    print(f"labelled rows {M}, loss {curve[:50].mean():.4f} -> {curve[-50:].mean():.4f}, "
          f"training {t2 - t1:.1f}s", flush=True)
    report = {"training": {"candidates": int(len(tr[1])), "labelled_rows": int(M), "steps": args.steps,
                           "batch": args.batch, "lr": args.lr, "seeds": list(LEAD_TRAIN_SEED), "gate": args.gate,
                           "loss_first50": float(curve[:50].mean()), "loss_last50": float(curve[-50:].mean()),
                           "seconds": t2 - t1}}
    report["enumerator"] = compare_lead(ctx, model, rows[:n_enum], meta[:n_enum], args.gate)
    print(json.dumps(report["enumerator"]), flush=True)
    report["whole"] = compare_lead(ctx, model, rows, meta, args.gate)
    print(json.dumps(report["whole"]), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


def main(argv=None):
    import argparse
    import json
    import time
    from . import Context
    ap = argparse.ArgumentParser(description="train the row-order network on plain witness logs and "
                                             "compare learned with plain order on 2-(25,4,1)")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--per-set", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-random", type=int, default=1 << 16)
    ap.add_argument("--beam", type=int, default=64)
    args = ap.parse_args(argv)
    ctx = Context(capacity=1 << 18, stride=32)
    t0 = time.time()
    model = RowOrderModel(ctx, max_rows=args.batch, seed=1)
    stream = training_stream(ctx, per_set=args.per_set, seed=1)
    d_X, d_y, d_w, M = pair_dataset(ctx, model, *stream, od=19, seed=2)
    t1 = time.time()
    curve = model.fit(d_X, d_y, d_w, M, args.batch, args.steps, args.lr)
    t2 = time.time()
    np.save(args.out, model.params)
    print(f"pairs {M}, loss {curve[:50].mean():.4f} -> {curve[-50:].mean():.4f}; data {t1 - t0:.1f}s, "
          f"training {t2 - t1:.1f}s", flush=True)
    rows, meta, n_enum = c5_stream(ctx, n_random=args.eval_random, beam=args.beam)
    report = {"training": {"pairs": int(M), "steps": args.steps, "batch": args.batch, "lr": args.lr,
                           "loss_first50": float(curve[:50].mean()), "loss_last50": float(curve[-50:].mean()),
                           "seconds": t2 - t1}}
    report["enumerator"] = compare_orders(ctx, model, rows[:n_enum], meta[:n_enum], 16)
    print(json.dumps(report["enumerator"]), flush=True)
    report["random"] = compare_orders(ctx, model, rows[n_enum:], meta[n_enum:], 16)
    print(json.dumps(report["random"]), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "lead":
        main_lead(sys.argv[2:])
    else:
        main()
