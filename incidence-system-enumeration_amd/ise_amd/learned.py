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


def load(ctx, path=WEIGHTS):
    """A model with the stored trained parameters."""
    model = RowOrderModel(ctx, max_rows=4096)
    p = np.load(path, allow_pickle=False)
    if p.shape != (MLP_PARAMS,):
        raise ValueError("stored parameters have the wrong size")
    model.params = p
    return model
