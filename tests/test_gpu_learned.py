"""The learned row order on the GPU: the MFMA network and its gradient
against the CPU oracle, the pair features bit-exact, the fused ordering
kernel, training, learned-order verdicts bit-exact with the plain search,
and the beam-sampled enumerator bit-exact with the oracle's."""
import numpy as np
import pytest

import ise_amd
from _helpers import check_witness, unpack

pytestmark = pytest.mark.gpu

P = ise_amd.MLP_PARAMS
IN, HID, OPAD = 16, 128, 16
W3 = IN * HID + HID + HID * HID + HID
B3 = W3 + HID * OPAD


def bf16_bits(x):
    """Round-to-nearest-even bf16 bit patterns of float32 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def random_params(rng, scale=0.25):
    p = (rng.standard_normal(P) * scale).astype(np.float32)
    p[W3:B3].reshape(HID, OPAD)[:, 1:] = 0
    p[B3 + 1:] = 0
    return p


@pytest.fixture(scope="module")
def model(gpu_ctx):
    m = ise_amd.RowOrderModel(gpu_ctx, max_rows=4096, seed=7)
    yield m
    m.close()


def mixed_batch(oracle, n_random=600):
    r1, m1 = ise_amd.random_extensions(13, 3, 1, n_random, 3, 13, 41)
    _, _, _, (r2, m2) = oracle.enumerate(9, 3, 1, capture=(3000, 32, 2, 9))
    r3, m3 = ise_amd.random_extensions(25, 4, 1, n_random, 8, 16, 42)
    return np.concatenate([r1, r2, r3]), np.concatenate([m1, m2, m3])


def test_mfma_forward_matches_oracle(gpu_ctx, oracle, model):
    print("case mlp forward 1024 rows", flush=True)
    rng = np.random.default_rng(1)
    p = random_params(rng)
    model.params = p
    M = 1024
    Xb = bf16_bits(rng.standard_normal((M, IN)))
    d_X = gpu_ctx.device_array(Xb.shape, np.uint16, Xb)
    d_z = gpu_ctx.device_array((M,), np.float32)
    model.forward(d_X, M, d_z)
    z = d_z.to_host()
    ref = oracle.mlp_forward(p, bits_to_f32(Xb), bf16=True)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(z - ref)) <= 0.02 * scale + 1e-3, np.max(np.abs(z - ref)) / scale
    assert np.corrcoef(z, ref)[0, 1] > 0.9999


def test_gradient_matches_oracle(gpu_ctx, oracle, model):
    print("case mlp gradient 512 rows", flush=True)
    rng = np.random.default_rng(2)
    p = random_params(rng)
    model.params = p
    M = 512
    Xb = bf16_bits(rng.standard_normal((M, IN)))
    y = (rng.random(M) < 0.3).astype(np.float32)
    w = (rng.random(M) + 0.2).astype(np.float32)
    d_X = gpu_ctx.device_array(Xb.shape, np.uint16, Xb)
    d_y = gpu_ctx.device_array(y.shape, np.float32, y)
    d_w = gpu_ctx.device_array(w.shape, np.float32, w)
    g, loss = model.gradient(d_X, d_y, d_w, M)
    ref, ref_loss = oracle.mlp_gradient(p, bits_to_f32(Xb), y, w, bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 + 1e-2 * abs(ref_loss)
    bounds = [0, IN * HID, IN * HID + HID, W3 - HID, W3, B3, P]
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        a, b = g[lo:hi], ref[lo:hi]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-6, (lo, hi)
    # the gradient does not move the parameters
    assert np.array_equal(model.params, p)


def test_pair_features_bit_exact(gpu_ctx, oracle, model):
    print("case pair features mixed batch od 16", flush=True)
    rows, meta = mixed_batch(oracle, 300)
    od = 16
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_X = gpu_ctx.device_array((n * od * od, IN), np.uint16)
    model.features(n, rows.shape[1], d_rows, d_meta, od, d_X)
    X = bits_to_f32(d_X.to_host()).reshape(n, od, od, IN)
    ref = oracle.mlp_features(rows, meta, od)
    bad = np.argwhere(X != ref)
    assert len(bad) == 0, f"first difference at {bad[0]}: gpu {X[tuple(bad[0])]} oracle {ref[tuple(bad[0])]}"


def test_fused_order_kernel(gpu_ctx, oracle, model):
    print("case fused order kernel mixed batch od 16", flush=True)
    rng = np.random.default_rng(3)
    p = random_params(rng)
    model.params = p
    rows, meta = mixed_batch(oracle, 200)
    od = 16
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_order = gpu_ctx.device_array((n, od, od), np.uint8)
    d_score = gpu_ctx.device_array((n, od, od), np.float32)
    model.order(n, rows.shape[1], d_rows, d_meta, od, d_order, d_score)
    order, score = d_order.to_host(), d_score.to_host()
    feats = oracle.mlp_features(rows, meta, od)
    ref = oracle.mlp_forward(p, feats.reshape(-1, IN), bf16=True).reshape(n, od, od)
    depth = meta & 0xff
    for c in range(n):
        d = int(depth[c])
        s, r = score[c, :d, :d], ref[c, :d, :d]
        assert np.max(np.abs(s - r)) <= 0.02 * max(np.max(np.abs(r)), 1.0), c
        for i in range(d):
            want = sorted(range(d), key=lambda j: (-s[i, j], j))
            assert order[c, i, :d].tolist() == want
            assert order[c, i, d:].tolist() == list(range(d, od))
        for i in range(d, od):
            assert order[c, i].tolist() == list(range(od))


def test_training_reduces_the_loss(gpu_ctx, model):
    print("case mlp training on a 2-(13,3,1) + 2-(15,3,1) witness log", flush=True)
    from ise_amd import learned
    rows, meta = gpu_ctx.enumerator_candidates(13, 3, 1, 1, 13)
    r2, m2 = ise_amd.random_extensions(15, 3, 1, 3000, 4, 15, 5)
    rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    res = learned.plain_witness_log(gpu_ctx, rows, meta)
    rej = res["verdict"] == 0
    m = ise_amd.RowOrderModel(gpu_ctx, max_rows=2048, seed=11)
    try:
        d_X, d_y, d_w, M = learned.pair_dataset(gpu_ctx, m, rows[rej], meta[rej], res["verdict"][rej],
                                                res["fail_row"][rej], res["perm_row"][rej], od=16)
        assert M >= 4096
        curve = m.fit(d_X, d_y, d_w, M, 2048, 300, 3e-3)
        first, last = curve[:20].mean(), curve[-20:].mean()
        print(f"loss {first:.4f} -> {last:.4f}", flush=True)
        assert np.all(np.isfinite(curve))
        assert last < 0.7 * first
    finally:
        m.close()


def test_learned_order_verdicts_bit_exact(gpu_ctx, oracle, model):
    print("case learned order check mixed batch od 16 budget 32", flush=True)
    rng = np.random.default_rng(4)
    model.params = random_params(rng)
    rows, meta = mixed_batch(oracle, 500)
    od, budget = 16, 32
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_order = gpu_ctx.device_array((n, od, od), np.uint8)
    model.order(n, rows.shape[1], d_rows, d_meta, od, d_order)
    order = d_order.to_host()
    gpu_ctx.set_order(ise_amd.ORDER_LEARNED, budget, d_order, od)
    try:
        g = gpu_ctx.check_batch(rows, meta)
    finally:
        gpu_ctx.set_order(ise_amd.ORDER_PLAIN)
    hunt = oracle.check_batch_hunt(rows, meta, order=order, order_dim=od, budget=budget)
    for f in ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "orbits",
              "counters"):
        assert np.array_equal(g[f], hunt[f]), f
    plain = oracle.check_batch(rows, meta)
    assert np.array_equal(g["verdict"], plain["verdict"])
    canon = plain["verdict"] == 1
    assert np.array_equal(g["group_order"][canon], plain["group_order"][canon])
    for c in np.nonzero(~canon)[0][:300]:
        d, b = int(meta[c]) & 0xff, (int(meta[c]) >> 8) & 0xff
        check_witness(unpack(rows[c, :d], b), g["perm_row"][c], int(g["fail_row"][c]))


@pytest.mark.parametrize("v,k,lam,beam,seed,dmin,dmax", [(15, 3, 1, 8, 5, 4, 15), (25, 4, 1, 2, 3, 8, 16)])
def test_sampled_enumerator_matches_oracle(gpu_ctx, oracle, v, k, lam, beam, seed, dmin, dmax):
    print(f"case sampled enumeration {(v, k, lam)} beam {beam}", flush=True)
    rows, meta, lc = gpu_ctx.enumerate_sampled(v, k, lam, beam, seed, dmin, dmax, cap=1 << 18)
    r2, m2, lc2 = oracle.enumerate_sampled(v, k, lam, beam, seed, dmin, dmax, cap=1 << 18)
    assert len(meta) > 100
    assert np.array_equal(meta, m2)
    assert np.array_equal(rows, r2)
    assert np.array_equal(lc[1:], lc2[1:v + 1])


def _hunt_success(oracle, rows, meta, tables, budget, idx):
    """Fraction of candidates idx whose witness the hunt finds within budget
    (tables None = increasing row order)."""
    found = 0
    for c in idx:
        d, b = int(meta[c]) & 0xff, (int(meta[c]) >> 8) & 0xff
        mat = np.ascontiguousarray(unpack(rows[c, :d], b))
        o = None if tables is None else np.ascontiguousarray(tables[c, :d, :d])
        perm = np.zeros(32, np.uint8)
        fr = np.zeros(1, np.uint8)
        nodes = np.zeros(1, np.uint32)
        found += oracle.L.orc_hunt(mat.ctypes.data, d, b, None if o is None else o.ctypes.data, budget,
                                   perm.ctypes.data, fr.ctypes.data, nodes.ctypes.data)
    return found / len(idx)


def test_trained_network_gain(gpu_ctx, oracle):
    """What the trained network gains: against an untrained network of the
    same shape, and against the method without it (plain order; increasing
    row order for the hunt), on 2-(25,4,1) candidates it never saw."""
    print("case trained row-order network vs untrained / plain on 2-(25,4,1)", flush=True)
    from ise_amd import learned
    rows, meta, _ = gpu_ctx.enumerate_sampled(25, 4, 1, 16, 5, 8, 16, stride=16, cap=1 << 16)
    r2, m2 = ise_amd.random_extensions(25, 4, 1, 2000, 8, 16, 9, stride=16)
    rows, meta = np.concatenate([rows, r2]), np.concatenate([meta, m2])
    n, od, budget = len(meta), 16, 8
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    dev, res = gpu_ctx.device_results(n, ("verdict", "counters"))
    gpu_ctx.check_batch_device(n, 16, d_rows, d_meta, res, d_head)
    verdict, plain = dev["verdict"].to_host(), dev["counters"].to_host()
    rej = np.nonzero(verdict == 0)[0]
    assert len(rej) > 10000
    out = {}
    for name, model in (("trained", learned.load(gpu_ctx)), ("untrained", ise_amd.RowOrderModel(gpu_ctx, 4096, 3))):
        d_order = gpu_ctx.device_array((n, od, od), np.uint8)
        model.set_gate(0)            # order every candidate: the whole stream is scored
        model.time_check(n, 16, d_rows, d_meta, od, d_order, budget, res, d_head, 1)
        assert np.array_equal(dev["verdict"].to_host(), verdict), name
        out[name] = (dev["counters"].to_host(), d_order.to_host())
        model.close()
    sample = rej[:: max(1, len(rej) // 3000)]
    succ = {name: _hunt_success(oracle, rows, meta, out[name][1], budget, sample) for name in out}
    succ["increasing"] = _hunt_success(oracle, rows, meta, None, budget, sample)
    perms = {"plain": plain[rej, 0].mean(), **{k: v[0][rej, 0].mean() for k, v in out.items()}}
    print(f"hunt success (budget {budget}): " + ", ".join(f"{k} {v:.3f}" for k, v in succ.items()), flush=True)
    print("mean permutations before rejection: " + ", ".join(f"{k} {v:.2f}" for k, v in perms.items()), flush=True)
    # training teaches the network where witnesses are ...
    assert succ["trained"] > succ["untrained"] + 0.3
    assert succ["trained"] >= succ["increasing"] - 0.02
    # ... but the plain order already rejects those candidates as fast: the
    # learned order is no faster than plain here, and must not be much slower
    assert perms["trained"] <= 1.1 * perms["plain"]


@pytest.mark.parametrize("gate", [1, 16])
def test_gated_learned_batch(gpu_ctx, oracle, model, gate):
    """Learned-order batches with a gate: candidates the plain search decides
    within `gate` permutations keep the plain results; the others get the
    oracle's hunt-then-plain results in the order the network proposed for
    them.  Verdicts equal the plain search's throughout."""
    print(f"case gated learned batch gate {gate}", flush=True)
    rng = np.random.default_rng(9)
    model.params = random_params(rng)
    r1, m1 = mixed_batch(oracle, 300)
    r2, m2 = oracle.enumerate_sampled(25, 4, 1, 2, 0x61, 8, 16, stride=32, cap=1 << 14)[:2]
    rows, meta = np.concatenate([r1, r2]), np.concatenate([m1, m2])
    od, budget = 32, 24
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    d_order = gpu_ctx.device_array((n, od, od), np.uint8)
    fields = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row", "orbits", "counters")
    dev, res = gpu_ctx.device_results(n, fields)
    model.set_gate(gate)
    try:
        model.time_check(n, rows.shape[1], d_rows, d_meta, od, d_order, budget, res, d_head, 1)
    finally:
        model.set_gate()
    g = {f: dev[f].to_host() for f in fields}
    order = d_order.to_host()
    plain = oracle.check_batch(rows, meta)
    assert np.array_equal(g["verdict"], plain["verdict"])
    gated = plain["counters"][:, 0] > gate          # left open after `gate` permutations
    assert gated.any() and (~gated).any()
    for f in fields:
        assert np.array_equal(g[f][~gated], plain[f][~gated]), f
    idx = np.nonzero(gated)[0]
    hunt = oracle.check_batch_hunt(rows[idx], meta[idx], order=order[idx], order_dim=od, budget=budget)
    for f in fields:
        assert np.array_equal(g[f][idx], hunt[f]), f
    for b in list(dev.values()) + [d_rows, d_meta, d_head, d_order]:
        b.free()
