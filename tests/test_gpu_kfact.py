"""GPU parity of the 1-factorization canonicity kernel against the oracle:
verdict, first witness in plain order, automorphisms counted and relabellings
tried, on random partial factorizations of K_12 (3-10 rounds), the reference's
printed canonical factorizations and their prefixes, and brute-forced K_8."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden

pytestmark = pytest.mark.gpu


def oracle_batch(oracle, rows, meta, n):
    out = {k: [] for k in ise_amd.KF_FIELDS}
    for i in range(len(meta)):
        nr = int(meta[i]) & 0xff
        v, w, a, t = oracle.kf_check(ise_amd.kf_unpack(rows[i, :nr], n))
        for k, x in zip(ise_amd.KF_FIELDS, (v, w, a, t)):
            out[k].append(x)
    return {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in out.items()}


def assert_same(g, o, what):
    for k in ise_amd.KF_FIELDS:
        if not np.array_equal(g[k], o[k]):
            i = int(np.nonzero(g[k] != o[k])[0][0])
            raise AssertionError(f"{what}: {k} differs at {i}: gpu={g[k][i]} oracle={o[k][i]}")


def test_random_k12_partials(gpu_ctx, oracle):
    print("case kfact random K12 3-10 rounds", flush=True)
    rows, meta = ise_amd.kf_random(12, 3, 10, 4096, 0x120311)
    g = gpu_ctx.kf_check(rows, meta)
    o = oracle_batch(oracle, rows, meta, 12)
    assert_same(g, o, "K12 random")


def test_printed_canonical_factorizations_and_prefixes(gpu_ctx, oracle):
    print("case kfact printed K10 + U12 + prefixes", flush=True)
    kf = golden("kfact.json")
    mats = [(m["rounds"], m["aut"]) for m in kf["k10"]["matrices"]]
    rows, meta = ise_amd.kf_pack([m for m, _ in mats])
    g = gpu_ctx.kf_check(rows, meta)
    assert (g["verdict"] == 1).all()
    assert (g["n_auts"] == np.array([a for _, a in mats], np.uint64)).all()
    assert (g["n_tried"] == 9 * 120 * 32).all()
    u12 = np.array(kf["u12"]["matrices"][0]["rounds"], np.uint8)
    prefixes = [u12[:k] for k in range(1, 12)] + [np.array(m, np.uint8)[:k] for m, _ in mats[:40] for k in (3, 5, 7)]
    rows, meta = ise_amd.kf_pack(prefixes)
    g = gpu_ctx.kf_check(rows, meta)
    o = {k: [] for k in ise_amd.KF_FIELDS}
    for p in prefixes:
        for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check(p)):
            o[k].append(x)
    o = {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in o.items()}
    assert_same(g, o, "prefixes")
    assert int(g["n_auts"][10]) == 660 and g["verdict"][10] == 1


def test_k8_against_bruteforce(gpu_ctx, oracle):
    print("case kfact K8 brute force", flush=True)
    rows, meta = ise_amd.kf_random(8, 1, 7, 300, 88)
    g = gpu_ctx.kf_check(rows, meta)
    for i in range(300):
        nr = int(meta[i]) & 0xff
        mat = ise_amd.kf_unpack(rows[i, :nr], 8)
        bv, bauts = oracle.kf_bruteforce(mat)
        assert g["verdict"][i] == bv
        if bv:
            assert int(g["n_auts"][i]) == bauts
    o = oracle_batch(oracle, rows, meta, 8)
    assert_same(g, o, "K8")


def test_mixed_sizes_and_invalid(gpu_ctx, oracle):
    print("case kfact mixed sizes", flush=True)
    parts = [ise_amd.kf_random(n, 1, min(n - 1, 6), 200, n) for n in (4, 6, 10, 14, 16)]
    rows = np.concatenate([p[0] for p in parts])
    meta = np.concatenate([p[1] for p in parts])
    g = gpu_ctx.kf_check(rows, meta)
    o = {k: [] for k in ise_amd.KF_FIELDS}
    for i in range(len(meta)):
        nr, n = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
        for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check(ise_amd.kf_unpack(rows[i, :nr], n))):
            o[k].append(x)
    o = {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in o.items()}
    assert_same(g, o, "mixed")
    bad = rows[:1].copy()
    bad[0, 1] = bad[0, 0]                       # repeats the edges of round 0
    with pytest.raises(ise_amd.ISEError):
        gpu_ctx.kf_check(bad, meta[:1])


def test_device_timed_path(gpu_ctx, oracle):
    print("case kfact device timed", flush=True)
    rows, meta = ise_amd.kf_random(12, 3, 10, 2048, 5)
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    dev, s = gpu_ctx.kf_device_results(n)
    ms = gpu_ctx.kf_time_batch_device(n, 16, d_rows, d_meta, s, d_head, 2)
    assert (ms > 0).all()
    g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
    assert_same(g, oracle_batch(oracle, rows, meta, 12), "device")


@pytest.mark.parametrize("first,ranges", [(ise_amd.DEFAULT_KF_FIRST_CHUNKS, ise_amd.DEFAULT_KF_MAX_RANGES), (0, 1)])
def test_device_records_that_are_no_factorization(gpu_ctx, oracle, first, ranges):
    """Device records are not checked on the host: a round repeated (pairs in
    two rounds) or a player twice in a round gives ISE_VERDICT_INVALID from
    the kernels (the host entry point rejects them), the others their
    results."""
    print(f"case kfact device invalid records split={first}/{ranges}", flush=True)
    rows, meta = ise_amd.kf_random(12, 3, 8, 512, 0x1D)
    good = np.ones(len(meta), bool)
    rows[5, 2] = rows[5, 0]                      # round 2 = round 0
    twice = ise_amd.kf_unpack(rows[9, :int(meta[9]) & 0xff], 12)
    twice[0] = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 9]   # player 9 twice, 11 missing
    packed, pmeta = ise_amd.kf_pack([twice])
    assert pmeta[0] == meta[9]
    rows[9] = packed[0]
    good[[5, 9]] = False
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    dev, s = gpu_ctx.kf_device_results(n)
    gpu_ctx.kf_set_split(first, ranges)
    try:
        gpu_ctx.kf_time_batch_device(n, 16, d_rows, d_meta, s, d_head, 1)
    finally:
        gpu_ctx.kf_set_split()
    g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
    for b in list(dev.values()) + [d_rows, d_meta, d_head]:
        b.free()
    assert (g["verdict"][~good] == 3).all()
    o = oracle_batch(oracle, rows[good], meta[good], 12)
    assert_same({k: v[good] for k, v in g.items()}, o, "device, valid records")
    for i in (5, 9):
        with pytest.raises(ise_amd.ISEError):
            gpu_ctx.kf_check(rows[i:i + 1], meta[i:i + 1])


_SPLIT_CASE = {}


def split_case(oracle):
    """Random K_12 partials (a few run tens of thousands of relabellings before
    their witness), the printed K_10 factorizations (all canonical), the
    uniform K_12 factorization and its prefixes, and small K_14/K_16 partials."""
    if not _SPLIT_CASE:
        kf = golden("kfact.json")
        r1, m1 = ise_amd.kf_random(12, 3, 10, 1 << 14, 0x12)
        r2, m2 = ise_amd.kf_pack([m["rounds"] for m in kf["k10"]["matrices"][:64]])
        u12 = np.array(kf["u12"]["matrices"][0]["rounds"], np.uint8)
        r3, m3 = ise_amd.kf_pack([u12[:k] for k in range(1, 12)])
        r4, m4 = ise_amd.kf_random(14, 1, 3, 64, 0x14)
        r5, m5 = ise_amd.kf_random(16, 1, 2, 16, 0x16)
        rows = np.concatenate([r1, r2, r3, r4, r5])
        meta = np.concatenate([m1, m2, m3, m4, m5])
        o = {k: [] for k in ise_amd.KF_FIELDS}
        for i in range(len(meta)):
            nr, n = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
            for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check(ise_amd.kf_unpack(rows[i, :nr], n))):
                o[k].append(x)
        o = {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in o.items()}
        _SPLIT_CASE.update(rows=rows, meta=meta, o=o)
    return _SPLIT_CASE["rows"], _SPLIT_CASE["meta"], _SPLIT_CASE["o"]


@pytest.mark.parametrize("first,ranges", [(0, 1), (1, 1), (1, 3), (2, 256), (8, 256), (1, 65536)])
def test_split_launch_same_results(gpu_ctx, oracle, first, ranges):
    """First pass + ranges + combine give the single scan's results: the
    first witness, the automorphisms before it and the relabellings tried,
    with one range per candidate, a few, and more ranges than chunks."""
    print(f"case kfact split first={first} ranges={ranges}", flush=True)
    rows, meta, o = split_case(oracle)
    # candidates that reach the last window, which runs from 2^20 to the end
    assert (o["n_tried"] > (1 << 20)).sum() >= 2
    gpu_ctx.kf_set_split(first, ranges)
    try:
        g = gpu_ctx.kf_check(rows, meta)
        g1 = gpu_ctx.kf_check(rows[-40:], meta[-40:])
    finally:
        gpu_ctx.kf_set_split()
    assert_same(g, o, f"split {first}/{ranges}")
    assert_same(g1, {k: v[-40:] for k, v in o.items()}, f"split {first}/{ranges} small batch")


@pytest.mark.parametrize("variant", ["rounds_permuted", "pairs_shuffled", "pairs_flipped"])
def test_rounds_not_in_normal_form(gpu_ctx, oracle, variant):
    """The test compares each relabelled round in normal form with the round
    as given, byte for byte.  Given rounds out of order (round 0 no longer
    (0 1)(2 3)...), with pairs out of order, or with pairs written larger
    player first, must give the oracle's results too."""
    print(f"case kfact not normal form {variant}", flush=True)
    rows, meta = ise_amd.kf_random(12, 2, 8, 3000, 0x5EED)
    rng = np.random.default_rng(len(variant))
    mats = []
    for i in range(len(meta)):
        nr = int(meta[i]) & 0xff
        mat = ise_amd.kf_unpack(rows[i, :nr], 12).copy()
        if variant == "rounds_permuted":
            mat = mat[rng.permutation(nr)]
        else:
            for r in range(nr):
                pairs = mat[r].reshape(-1, 2)
                if variant == "pairs_shuffled":
                    pairs = pairs[rng.permutation(len(pairs))]
                else:
                    flip = rng.random(len(pairs)) < 0.3
                    pairs[flip] = pairs[flip][:, ::-1]
                mat[r] = pairs.reshape(-1)
        mats.append(mat)
    rows, meta = ise_amd.kf_pack(mats)
    g = gpu_ctx.kf_check(rows, meta)
    assert_same(g, oracle_batch(oracle, rows, meta, 12), variant)


def test_split_buffers_grow_past_capacity(oracle):
    """A device batch larger than the context capacity grows the split launch's
    buffers on its first call and gives the oracle's results."""
    print("case kfact split buffers grow", flush=True)
    rows, meta = ise_amd.kf_random(12, 3, 10, 5000, 0x6B)
    with ise_amd.Context(capacity=1024, stride=16) as ctx:
        n = len(meta)
        d_rows = ctx.device_array(rows.shape, np.uint64, rows)
        d_meta = ctx.device_array(meta.shape, np.uint32, meta)
        d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
        dev, s = ctx.kf_device_results(n)
        ctx.kf_time_batch_device(n, 16, d_rows, d_meta, s, d_head, 1)
        g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
        for b in list(dev.values()) + [d_rows, d_meta, d_head]:
            b.free()
    assert_same(g, oracle_batch(oracle, rows, meta, 12), "grown buffers")


_LONG_CASE = {}


def long_records_case(oracle):
    """Records of 12-15 rounds: 200 relabelled prefixes (12-13 rounds) of the
    printed perfect 1-factorizations of K_14 and 200 random partial
    1-factorizations of K_16 with 12-15 rounds, with the oracle's results."""
    if not _LONG_CASE:
        kf = golden("kfact.json")
        src, smeta = ise_amd.kf_pack([m["rounds"] for m in kf["p14"]["matrices"]])
        r1, m1 = ise_amd.kf_relabel_p1f(src, smeta, 200, 12, 0x1412)
        r2, m2 = ise_amd.kf_random(16, 12, 15, 200, 0x1612)
        rows = np.concatenate([r1, r2])
        meta = np.concatenate([m1, m2])
        assert ((meta & 0xff) >= 12).all() and set(np.unique(m2 & 0xff)) == {12, 13, 14, 15}
        o = {k: [] for k in ise_amd.KF_FIELDS}
        for i in range(len(meta)):
            nr, n = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
            for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check(ise_amd.kf_unpack(rows[i, :nr], n))):
                o[k].append(x)
        o = {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in o.items()}
        _LONG_CASE.update(rows=rows, meta=meta, o=o)
    return _LONG_CASE["rows"], _LONG_CASE["meta"], _LONG_CASE["o"]


@pytest.mark.parametrize("first,ranges", [(ise_amd.DEFAULT_KF_FIRST_CHUNKS, ise_amd.DEFAULT_KF_MAX_RANGES), (0, 1)])
def test_full_k14_perfect_factorizations(gpu_ctx, first, ranges):
    """The 23 printed perfect 1-factorizations of K_14 (13 rounds, the full
    matrices) through the general canonizer: canonical, the printed |Aut(M)|,
    and every relabelling tried (7! pair-slot arrangements x 2^7 orientations
    x 13 rounds)."""
    print(f"case kfact full K14 P1F general split={first}/{ranges}", flush=True)
    kf = golden("kfact.json")
    mats = kf["p14"]["matrices"]
    rows, meta = ise_amd.kf_pack([m["rounds"] for m in mats])
    assert ((meta & 0xff) == 13).all()
    gpu_ctx.kf_set_split(first, ranges)
    try:
        g = gpu_ctx.kf_check(rows, meta)
    finally:
        gpu_ctx.kf_set_split()
    assert (g["verdict"] == 1).all()
    assert (g["n_auts"] == np.array([m["aut"] for m in mats], np.uint64)).all()
    assert (g["n_tried"] == 5040 * 128 * 13).all()
    assert (g["witness"] == np.uint64(2**64 - 1)).all()


@pytest.mark.parametrize("first,ranges", [(ise_amd.DEFAULT_KF_FIRST_CHUNKS, ise_amd.DEFAULT_KF_MAX_RANGES), (0, 1)])
def test_records_of_12_to_15_rounds(gpu_ctx, oracle, first, ranges):
    """Rounds 12-14 are staged in the last slot of the first look's four-lane
    groups; relabelled K_14 prefixes and K_16 partials of 12-15 rounds against
    the oracle, split launch and one wave per candidate."""
    print(f"case kfact 12-15 rounds split={first}/{ranges}", flush=True)
    rows, meta, o = long_records_case(oracle)
    gpu_ctx.kf_set_split(first, ranges)
    try:
        g = gpu_ctx.kf_check(rows, meta)
    finally:
        gpu_ctx.kf_set_split()
    assert_same(g, o, f"12-15 rounds split {first}/{ranges}")
