"""The lead order on the GPU (ise_mlp_check_lead_device): one launch per
batch, field by field against the oracle's lead-order search
(orc_check_batch_lead) on the same first-row orders -- given tables, and the
network's own order as the kernel exports it -- counters included; the
feature kernel bit-exact; scores against the oracle's network; verdicts of
the uncounted launch (refutation filter on) equal to the counted one's and
to the plain search's."""
import numpy as np
import pytest

import ise_amd

pytestmark = pytest.mark.gpu

FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "group_order_hi", "perm_row",
          "orbits", "n_gens", "counters")
GEN_CAP = 4


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


@pytest.fixture(scope="module")
def model(gpu_ctx):
    m = ise_amd.RowOrderModel(gpu_ctx, max_rows=4096, seed=11)
    yield m
    m.close()


@pytest.fixture(scope="module")
def cases(oracle):
    out = []
    _, _, _, (r, m) = oracle.enumerate(13, 3, 1, capture=(4096, 32, 2, 13))
    out.append(("2-(13,3,1) enumerator", r, m))
    _, _, _, (r, m) = oracle.enumerate(15, 3, 1, capture=(12000, 32, 4, 15))
    out.append(("2-(15,3,1) enumerator", r, m))
    _, _, _, (r, m) = oracle.enumerate(16, 6, 2, capture=(8192, 32, 2, 16))
    out.append(("2-(16,6,2) enumerator", r, m))
    r, m, _ = oracle.enumerate_sampled(25, 4, 1, 8, 0x5151, 8, 16, stride=32, cap=1 << 16)
    out.append(("2-(25,4,1) sampled", r[:8000], m[:8000]))
    r1, m1 = ise_amd.random_extensions(15, 3, 1, 3000, 4, 15, 78, stride=32)
    r2, m2 = ise_amd.random_extensions(19, 9, 4, 1000, 5, 19, 79, stride=32)   # depths past 16: plain search
    out.append(("random 2-(15,3,1) + 2-(19,9,4)", np.concatenate([r1, r2]), np.concatenate([m1, m2])))
    return out


def run_lead(gpu_ctx, model, rows, meta, lead_in=None, counted=True):
    n, stride = rows.shape
    fields = [f for f in FIELDS if counted or f != "counters"] + ["gens"]
    dev, s = gpu_ctx.device_results(n, fields, gen_cap=GEN_CAP)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((256,), np.uint32)
    d_in = None if lead_in is None else gpu_ctx.device_array(lead_in.shape, np.uint8, lead_in)
    d_out = gpu_ctx.device_array((n, 16), np.uint8, np.full((n, 16), 0xee, np.uint8))
    d_sc = gpu_ctx.device_array((n, 16), np.float32)
    model.check_lead(n, stride, d_rows, d_meta, s, d_head, d_in, d_out, d_sc)
    res = {k: v.to_host() for k, v in dev.items()}
    return res, d_out.to_host(), d_sc.to_host()


def assert_same(name, g, o, mask=None):
    for f in FIELDS:
        if f not in g:
            continue
        a, b = g[f], o[f]
        if mask is not None:
            a, b = a[mask], b[mask]
        bad = np.nonzero((a != b).reshape(len(a), -1).any(1))[0]
        assert len(bad) == 0, (name, f, bad[:8])
    ng = np.minimum(o["n_gens"], GEN_CAP)
    keep = np.arange(GEN_CAP)[None, :] < ng[:, None]
    if mask is not None:
        keep &= mask[:, None]
    assert np.array_equal(g["gens"][keep], o["gens"][keep]), (name, "gens")


@pytest.mark.parametrize("gate", [0, 16])
def test_given_first_rows_match_oracle(gpu_ctx, oracle, model, cases, gate):
    model.set_gate(gate)
    rng = np.random.default_rng(100 + gate)
    for name, rows, meta in cases:
        for kind in ("ascending", "random"):
            print(f"case lead {name} gate {gate} {kind}", flush=True)
            n = len(meta)
            if kind == "ascending":
                lead = np.tile(np.arange(16, dtype=np.uint8), (n, 1))
            else:
                lead = np.stack([rng.permutation(16) for _ in range(n)]).astype(np.uint8)
            g, _, _ = run_lead(gpu_ctx, model, rows, meta, lead)
            o, mode = oracle.check_batch_lead(rows, meta, lead, gate, gen_cap=GEN_CAP)
            assert_same(name, g, o)
    model.set_gate()


def test_network_first_rows_match_oracle(gpu_ctx, oracle, model, cases):
    """The kernel's own order (exported) replayed by the oracle gives every
    field; the order is the scores' stable descending sort; the scores are
    the oracle network's on the oracle's features."""
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(ise_amd.MLP_PARAMS) * 0.2).astype(np.float32)
    model.params = p
    model.set_gate(8)
    for name, rows, meta in cases:
        print(f"case lead network {name}", flush=True)
        g, lead, sc = run_lead(gpu_ctx, model, rows, meta)
        o, mode = oracle.check_batch_lead(rows, meta, lead, 8, gen_cap=GEN_CAP)
        assert_same(name, g, o)
        ran = np.nonzero(lead[:, 0] != 0xee)[0]
        assert len(ran) > 0, name
        assert np.array_equal(np.sort(np.nonzero(mode >= 2)[0]), np.sort(ran)) or (mode[ran] >= 2).all(), name
        d = (meta & 0xff).astype(int)
        X = oracle.lead_features(rows[ran], meta[ran])
        ref = oracle.mlp_forward(p, X.reshape(-1, 16), bf16=True).reshape(len(ran), 16)
        for k, c in enumerate(ran[:2000]):
            dc = d[c]
            assert sorted(lead[c, :dc]) == list(range(dc)), (name, c)
            assert (lead[c, dc:] == 0xff).all()
            s = sc[c, :dc]
            assert list(lead[c, :dc]) == sorted(range(dc), key=lambda j: (-s[j], j)), (name, c)
            scale = np.max(np.abs(ref[k, :dc])) + 1e-3
            assert np.max(np.abs(s - ref[k, :dc])) <= 0.02 * scale + 1e-3, (name, c)
    model.set_gate()


def test_lead_features_bit_exact(gpu_ctx, oracle, model, cases):
    for name, rows, meta in cases:
        print(f"case lead features {name}", flush=True)
        n, stride = rows.shape
        d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
        d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
        d_X = gpu_ctx.device_array((n, 16, 16), np.uint16)
        model.lead_features(n, stride, d_rows, d_meta, d_X)
        X = d_X.to_host()
        ref = oracle.lead_features(rows, meta)
        d = meta & 0xff
        ref[(d < 2) | (d > 16)] = 0
        assert np.array_equal(X, bf16_bits(ref)), name


def test_uncounted_launch_and_plain_verdicts(gpu_ctx, oracle, model, cases):
    """With the refutation filter (no counters) the launch gives the counted
    one's fields; verdicts equal the plain search's; canonical candidates get
    every plain field."""
    rng = np.random.default_rng(9)
    model.params = (rng.standard_normal(ise_amd.MLP_PARAMS) * 0.2).astype(np.float32)
    model.set_gate(16)
    for name, rows, meta in cases:
        print(f"case lead uncounted {name}", flush=True)
        g1, lead, _ = run_lead(gpu_ctx, model, rows, meta, counted=True)
        g2, lead2, _ = run_lead(gpu_ctx, model, rows, meta, counted=False)
        assert np.array_equal(lead, lead2), name
        for f in FIELDS[:-1]:
            assert np.array_equal(g1[f], g2[f]), (name, f)
        plain = oracle.check_batch(rows, meta, gen_cap=GEN_CAP)
        assert np.array_equal(g2["verdict"], plain["verdict"]), name
        canon = plain["verdict"] == 1
        assert_same(name, {k: v for k, v in g2.items() if k != "counters"},
                    {k: v for k, v in plain.items() if k != "counters"}, canon)
    model.set_gate()


def test_trained_lead_network_gain(gpu_ctx, oracle):
    """On 2-(25,4,1) candidates of a seed used neither in training nor in the
    benchmark, the trained network's first-row order compares fewer rows than
    the increasing order (the plain walk's) and than an untrained network,
    with the same verdicts."""
    from ise_amd import learned
    rows, meta, _ = oracle.enumerate_sampled(25, 4, 1, 32, 0x2929, 8, 16, stride=32, cap=1 << 20)
    n = len(meta)
    print(f"case lead gain: {n} candidates", flush=True)
    plain = oracle.check_batch(rows, meta)
    totals = {}
    for name in ("trained", "untrained", "ascending"):
        m = learned.load(gpu_ctx, learned.LEAD_WEIGHTS) if name == "trained" else \
            ise_amd.RowOrderModel(gpu_ctx, max_rows=4096, seed=3)
        m.set_gate(learned.LEAD_GATE)
        lead = np.tile(np.arange(16, dtype=np.uint8), (n, 1)) if name == "ascending" else None
        g, _, _ = run_lead(gpu_ctx, m, rows, meta, lead)
        assert np.array_equal(g["verdict"], plain["verdict"]), name
        totals[name] = int(g["counters"][:, 1].astype(np.int64).sum())
        m.close()
    print(f"rows compared: {totals}, plain {int(plain['counters'][:, 1].astype(np.int64).sum())}", flush=True)
    assert totals["trained"] < 0.95 * totals["ascending"]
    assert totals["trained"] < totals["untrained"]


def test_exact_labels_match_oracle(gpu_ctx, oracle):
    """The training labels: row j is 1 exactly when the lead search with j
    first returns a witness starting with j -- on the GPU as in the oracle."""
    from ise_amd import learned
    rows, meta, _ = oracle.enumerate_sampled(25, 4, 1, 8, 0x3131, 8, 16, stride=16, cap=1 << 16)
    plain = oracle.check_batch(rows, meta)
    d = (meta & 0xff).astype(int)
    hard = np.nonzero((plain["verdict"] == 0) & (plain["counters"][:, 0] > learned.LEAD_GATE) &
                      (plain["n_gens"] == 0) & (d >= 2) & (d <= 16))[0][:400]
    assert len(hard) > 50
    r, m = rows[hard], meta[hard]
    print(f"case exact labels: {len(hard)} candidates", flush=True)
    y, w = learned.lead_exact_labels(gpu_ctx, r, m)
    for j in range(16):
        lead = np.array([[j] + [x for x in range(16) if x != j]] * len(hard), np.uint8)
        o, _ = oracle.check_batch_lead(r, m, lead, learned.LEAD_GATE)
        ref = (o["verdict"] == 0) & (o["perm_row"][:, 0] == j) & (j < d[hard])
        assert np.array_equal(y[:, j] > 0, ref), j
    assert np.array_equal(w > 0, np.arange(16)[None, :] < d[hard][:, None])
    # a non-canonical candidate has a right first row (unless every search that would show it
    # meets an automorphism first and hands over to the plain search)
    assert ((y * w).sum(1) >= 1).mean() > 0.9


def test_lead_edge_shapes_and_fixed_rows(gpu_ctx, oracle, model):
    """Shapes the lead order takes and those it hands to the plain search:
    depths 1..32, 1..64 columns (more than 56: the replayed search), small
    orbit-form matrices with lenStab > 0, row permutations saved (|Aut| and
    generators as row permutations); given and network first-row orders;
    every field and counter against the oracle."""
    from _helpers import col_sorted, random_orbit_form_batch
    rng = np.random.default_rng(17)
    rows, meta = [], []
    for d, b in [(1, 1), (1, 64), (2, 2), (2, 64), (16, 56), (16, 57), (17, 16), (32, 64), (20, 40), (8, 12),
                 (16, 16), (5, 50)]:
        for _ in range(12):
            mat = col_sorted((rng.random((d, b)) < 0.5).astype(np.uint8))
            w = np.zeros(32, np.uint64)
            w[:d] = ise_amd.pack_rows(mat)
            rows.append(w)
            meta.append(ise_amd.pack_meta(d, b, 0, int(rng.integers(0, 2)) | 4))
    _, r2, m2 = random_orbit_form_batch(0x1EAD, 3000, stride=32)
    rows = np.concatenate([np.array(rows), r2])
    meta = np.concatenate([np.array(meta, np.uint32), m2])
    assert ((meta >> 16) & 0xff).astype(bool).sum() > 500
    n = len(meta)
    rng = np.random.default_rng(18)
    model.params = (rng.standard_normal(ise_amd.MLP_PARAMS) * 0.2).astype(np.float32)
    for gate in (0, 4):
        model.set_gate(gate)
        lead = np.stack([rng.permutation(16) for _ in range(n)]).astype(np.uint8)
        for given in (lead, None):
            print(f"case lead edges gate {gate} given {given is not None}", flush=True)
            g, lo, _ = run_lead(gpu_ctx, model, rows, meta, given)
            use = given if given is not None else lo
            o, mode = oracle.check_batch_lead(rows, meta, use, gate, gen_cap=GEN_CAP)
            assert_same("edges", g, o)
    model.set_gate()


def test_lead_invalid_device_records(gpu_ctx, model):
    """Records that fail validation get ISE_VERDICT_INVALID from the lead
    kernel too (the device path takes unchecked records)."""
    rows = np.zeros((4, 16), np.uint64)
    rows[:, 0] = 1
    meta = np.array([ise_amd.pack_meta(0, 5), ise_amd.pack_meta(17, 5), ise_amd.pack_meta(3, 65),
                     ise_amd.pack_meta(3, 5, 3)], np.uint32)
    g, _, _ = run_lead(gpu_ctx, model, rows, meta)
    assert (g["verdict"] == ise_amd.VERDICT_INVALID).all()


@pytest.mark.parametrize("v,k,lam", [(13, 3, 1), (15, 7, 3), (16, 6, 2), (15, 3, 1)])
def test_lead_mode_enumerations(oracle, v, k, lam):
    """The enumerators with the context in the lead order (ise_mlp_use_lead):
    host-solver, graph-replayed and device-resident enumerations give the
    plain order's design lists, |Aut| and candidate counts."""
    from ise_amd import learned
    print(f"case lead-order enumeration {(v, k, lam)}", flush=True)
    with ise_amd.Context(capacity=1 << 16, stride=32) as ctx:
        des, auts, st, lv = ctx.enumerate_bibd(v, k, lam)
        m = learned.load(ctx, learned.LEAD_WEIGHTS)
        m.set_gate(learned.LEAD_GATE)
        m.use_lead(True)
        des2, auts2, st2, lv2 = ctx.enumerate_bibd(v, k, lam)
        ctx.use_graphs(True)
        des3, auts3, _, _ = ctx.enumerate_bibd(v, k, lam)
        ctx.use_graphs(False)
        des4, auts4, _, _ = ctx.enumerate_bibd_device(v, k, lam)
        m.close()
    key = lambda d, a: sorted(zip(map(tuple, d.tolist()), map(int, a)))
    ref = key(des, auts)
    assert key(des2, auts2) == ref and key(des3, auts3) == ref and key(des4, auts4) == ref
    assert st2["candidates"] == st["candidates"] and np.array_equal(lv2, lv)


def test_lead_mode_submit_collect_and_noalloc(gpu_ctx, oracle):
    """ise_submit/ise_collect and the allocation-free device entry in the lead
    order give the lead device entry point's results; with column
    permutations asked for, the plain search's; after the network is
    destroyed, the plain order again."""
    from ise_amd import learned
    rows, meta, _ = oracle.enumerate_sampled(25, 4, 1, 8, 0x4141, 8, 16, stride=32, cap=1 << 16)
    r2, m2 = ise_amd.random_extensions(15, 3, 1, 2000, 4, 15, 81, stride=32)
    rows, meta = np.concatenate([rows[:6000], r2]), np.concatenate([meta[:6000], m2])
    n = len(meta)
    print(f"case lead mode submit/collect {n}", flush=True)
    fields = tuple(f for f in FIELDS if f != "counters")
    m = learned.load(gpu_ctx, learned.LEAD_WEIGHTS)
    m.set_gate(learned.LEAD_GATE)
    ref, _, _ = run_lead(gpu_ctx, m, rows, meta, counted=False)
    plain = gpu_ctx.check_batch(rows, meta, gen_cap=GEN_CAP)
    try:
        m.use_lead(True)
        got = gpu_ctx.check_batch(rows, meta, gen_cap=GEN_CAP, fields=fields + ("gens",))
        assert_same("submit/collect", got, {k: v for k, v in ref.items() if k != "counters"})
        # the allocation-free entry point
        dev, s = gpu_ctx.device_results(n, fields + ("gens",), gen_cap=GEN_CAP)
        d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
        d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
        gpu_ctx.check_batch_device_noalloc(n, 32, d_rows, d_meta, s)
        gpu_ctx.sync()
        assert_same("noalloc", {k: v.to_host() for k, v in dev.items()}, ref)
        # column permutations asked for: the plain search, every field
        cols = gpu_ctx.check_batch(rows, meta, gen_cap=GEN_CAP)
        for f in plain:
            assert np.array_equal(cols[f], plain[f]), f
    finally:
        m.close()
    after = gpu_ctx.check_batch(rows, meta, gen_cap=GEN_CAP, fields=fields + ("gens",))
    for f in after:
        assert np.array_equal(after[f], plain[f]), f
    assert (ref["verdict"] == plain["verdict"]).all()
