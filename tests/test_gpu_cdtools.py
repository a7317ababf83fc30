"""The k-Sys binary canonizer on the GPU (ise_cd_*: createCanonizer /
runCanonizer / releaseCanonizer, Source/CDTools.cpp:5-51) against the oracle
(orc_cd_run_canonizer), byte for byte: full-day k-systems (every day order
canonized) and partial ones (day order kept), groups of 2, 3 and 4, batches,
the device entry point and the argument checks.  Full-day K_10 runs (9! day
orders) compare with the oracle's answers cached in tests/golden/cd_canon.json
and with size-independent properties.  The canonical-form loop under a
column partition (ise_canonize_batch_device_part) is compared with the
oracle's orc_canonize_part field by field."""
import json
import os

import numpy as np
import pytest

import ise_amd
from test_oracle_cdtools import affine_plane_3, kf_fixture, relabel, round_robin

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def canonizer(v, k):
    return ise_amd.KSysCanonizer(v, k)


def resolvable_16_4():
    """AG(2,4): 16 players in 5 days of four groups of 4 (lines of the affine
    plane over GF(4), one parallel class per slope)."""
    mul = [[0, 0, 0, 0], [0, 1, 2, 3], [0, 2, 3, 1], [0, 3, 1, 2]]   # GF(4), elements 0, 1, a, a+1
    pts = [(x, y) for x in range(4) for y in range(4)]
    idx = {p: i for i, p in enumerate(pts)}
    days = []
    for slope in (None, 0, 1, 2, 3):
        groups = []
        for c in range(4):
            if slope is None:
                line = [(c, y) for y in range(4)]
            else:
                line = [(x, mul[slope][x] ^ c) for x in range(4)]
            groups.extend(idx[p] for p in line)
        days.append(groups)
    return np.array(days, np.uint8)


@pytest.mark.parametrize("v", [4, 6])
def test_full_days_small_k2(oracle, v):
    rng = np.random.default_rng(60 + v)
    with canonizer(v, 2) as cz:
        for t in range(4):
            m = relabel(round_robin(v), v, 2, rng, shuffle_days=True)
            print(f"K_{v} relabelling {t}", flush=True)
            assert np.array_equal(cz.run(m), oracle.cd_run_canonizer(m, v, 2))


def test_full_days_k8_fixture(oracle):
    mat = kf_fixture("u8")
    rng = np.random.default_rng(8)
    ref = oracle.cd_run_canonizer(mat, 8, 2)
    with canonizer(8, 2) as cz:
        assert np.array_equal(cz.run(mat), ref)
        batch = np.stack([relabel(mat, 8, 2, rng, shuffle_days=True) for _ in range(3)])
        out = cz.run_batch(batch)
        for o in out:
            assert np.array_equal(o, ref)


@pytest.mark.parametrize("name", ["k10", "p10", "u10"])
def test_full_days_k10_cached_oracle(name):
    fx = json.load(open(os.path.join(HERE, "golden", "cd_canon.json")))
    case = next(c for c in fx["cases"] if c["fixture"] == name)
    want = np.array([[int(ch) for ch in r] for r in case["canon"]], np.uint8)
    mat = kf_fixture(name, case["index"])
    rng = np.random.default_rng(10)
    with canonizer(10, 2) as cz:
        print(f"{name}: 9! day orders", flush=True)
        assert np.array_equal(cz.run(mat), want)
        # a relabelled copy with the days reordered gives the same matrix
        assert np.array_equal(cz.run(relabel(mat, 10, 2, rng, shuffle_days=True)), want)


def test_full_days_k10_properties(oracle):
    """Size-independent checks of a 9-day run: a canonical matrix under the
    day partition, at least as large as the canonical form of sampled day
    orders, a 1-factorization's incidence, invariant under relabelling."""
    mat = kf_fixture("k10", 3)
    with canonizer(10, 2) as cz:
        got = cz.run(mat)
    part = oracle.cd_day_starts(9, 5)
    again, _, _ = oracle.canonize_part(got, part)
    assert np.array_equal(again, got)
    assert (got.sum(0) == 2).all() and (got.sum(1) == 9).all()
    rng = np.random.default_rng(3)
    key = tuple(got.reshape(-1))
    for _ in range(5):
        m = mat[rng.permutation(9)]
        b = np.zeros((10, 45), np.uint8)
        for j in range(9):
            for g in range(5):
                b[m[j, 2 * g], 5 * j + g] = 1
                b[m[j, 2 * g + 1], 5 * j + g] = 1
        fixed, _, _ = oracle.canonize_part(b, part)
        assert tuple(fixed.reshape(-1)) <= key


@pytest.mark.parametrize("name,days", [("k10", 3), ("p10", 5), ("u10", 8), ("u12", 6), ("u12", 10),
                                       ("p14", 4)])
def test_partial_days_fixtures(oracle, name, days):
    mat = kf_fixture(name)[:days]
    v = mat.shape[1]
    rng = np.random.default_rng(days)
    with canonizer(v, 2) as cz:
        ref = oracle.cd_run_canonizer(mat, v, 2, days)
        assert np.array_equal(cz.run(mat, day=days), ref)
        batch = np.stack([relabel(mat, v, 2, rng, shuffle_days=False) for _ in range(5)] + [mat])
        for o in cz.run_batch(batch, day=days):
            assert np.array_equal(o, ref)


def test_groups_of_three(oracle):
    mat = affine_plane_3()
    rng = np.random.default_rng(9)
    with canonizer(9, 3) as cz:
        for t in range(3):
            m = relabel(mat, 9, 3, rng, shuffle_days=True)
            assert np.array_equal(cz.run(m), oracle.cd_run_canonizer(m, 9, 3))
            for day in (1, 2, 3):
                assert np.array_equal(cz.run(m[:day], day=day), oracle.cd_run_canonizer(m[:day], 9, 3, day))


def test_groups_of_four(oracle):
    mat = resolvable_16_4()
    rng = np.random.default_rng(16)
    m = relabel(mat, 16, 4, rng, shuffle_days=True)
    with canonizer(16, 4) as cz:
        assert np.array_equal(cz.run(m), oracle.cd_run_canonizer(m, 16, 4))
        assert np.array_equal(cz.run(m[:3], day=3), oracle.cd_run_canonizer(m[:3], 16, 4, 3))


def test_device_entry_point(oracle, gpu_ctx):
    mat = kf_fixture("u8")
    rng = np.random.default_rng(81)
    batch = np.stack([relabel(mat, 8, 2, rng, shuffle_days=True)[:5] for _ in range(4)])
    with canonizer(8, 2) as cz:
        d_in = ise_amd.DeviceArray(gpu_ctx, batch.shape, np.uint8, batch)
        d_out = ise_amd.DeviceArray(gpu_ctx, (4, 8), np.uint64)
        cz.run_device(4, d_in.ptr, 5, d_out.ptr)
        packed = d_out.to_host()
    for i in range(4):
        ref = oracle.cd_run_canonizer(batch[i], 8, 2, 5)
        bits = ((packed[i][:, None] >> np.arange(20, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
        assert np.array_equal(bits, ref)


def test_rejects_invalid_input():
    with pytest.raises(ise_amd.ISEError):
        ise_amd.KSysCanonizer(10, 3)                       # 3 does not divide 10
    with pytest.raises(ise_amd.ISEError):
        ise_amd.KSysCanonizer(34, 2)                       # more than 32 players
    with canonizer(12, 2) as cz:
        with pytest.raises(ise_amd.ISEError):
            cz.run(kf_fixture("u12"))                      # 11 days: 66 columns
        ok = cz.run(kf_fixture("u12")[:10], day=10)        # 60 columns
        assert ok.shape == (12, 60)
    with canonizer(8, 2) as cz:
        bad = kf_fixture("u8").copy()
        bad[2, 3] = 8
        with pytest.raises(ise_amd.ISEError):
            cz.run(bad)
        with pytest.raises(ise_amd.ISEError):
            cz.run(kf_fixture("u8"), k=4)                  # not the canonizer's group size
        with pytest.raises(ise_amd.ISEError):
            cz.run(kf_fixture("u8"), day=9)                # more days than K_8 has
        assert cz.run(kf_fixture("u8")).shape == (8, 28)   # the canonizer still works


def test_canonize_part_matches_oracle(oracle, gpu_ctx):
    """The canonical-form loop under a column partition, field by field:
    random 0/1 matrices and random partitions (one block = k = 0), and a
    wider case past the column-set search's packed width (57-64 columns: the
    replayed search under the partition)."""
    rng = np.random.default_rng(77)
    n, stride = 300, gpu_ctx.stride
    for part_kind in range(4):
        rows = np.zeros((n, stride), np.uint64)
        meta = np.zeros(n, np.uint32)
        mats, parts = [], []
        b = int(rng.integers(8, 40)) if part_kind < 3 else int(rng.integers(57, 65))
        if part_kind == 0:
            part = 1
        elif part_kind == 1:
            part = int(sum(1 << q for q in range(0, b, 5)))
        else:
            part = 1 | int(sum(1 << int(q) for q in rng.choice(np.arange(1, b), 6, replace=False)))
        for i in range(n):
            d = int(rng.integers(2, 12))
            m = (rng.random((d, b)) < rng.uniform(0.2, 0.6)).astype(np.uint8)
            mats.append(m)
            for r in range(d):
                rows[i, r] = int(sum(int(m[r, q]) << q for q in range(b)))
            meta[i] = d | (b << 8)
        got = gpu_ctx.canonize_batch_part(rows, meta, part)
        for i, m in enumerate(mats):
            want, aut, iters = oracle.canonize_part(m, part)
            d = m.shape[0]
            bits = ((got["canon"][i][:d, None] >> np.arange(b, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
            assert got["status"][i] == 1
            assert np.array_equal(bits, want), (part_kind, i)
            assert int(got["aut"][i]) == aut and int(got["iters"][i]) == iters, (part_kind, i)


def test_kirkman_solutions(oracle):
    """The seven printed 15-player k-systems in groups of three (the Kirkman
    schoolgirl solutions), all 7 days (5040 day orders each), in one batch:
    the cached oracle forms, seven different matrices, the same forms from
    relabelled copies, and the two printed in P_OrderedMatrices.txt land on
    two of the seven."""
    fx = json.load(open(os.path.join(HERE, "golden", "cd_canon.json")))
    want = {c["index"]: np.array([[int(ch) for ch in r] for r in c["canon"]], np.uint8)
            for c in fx["cases"] if c["fixture"] == "k15x3"}
    assert sorted(want) == list(range(7))
    kf = json.load(open(os.path.join(HERE, "golden", "kfact.json")))
    mats = np.array([m["rounds"] for m in kf["k15x3"]["matrices"]], np.uint8)
    rng = np.random.default_rng(153)
    with canonizer(15, 3) as cz:
        got = cz.run_batch(mats)
        again = cz.run_batch(np.stack([relabel(m, 15, 3, rng, shuffle_days=True) for m in mats]))
        p = cz.run_batch(np.array([m["rounds"] for m in kf["p15x3"]["matrices"]], np.uint8))
    for i in range(7):
        assert np.array_equal(got[i], want[i]), i
        assert np.array_equal(again[i], want[i]), i
    forms = {g.tobytes() for g in got}
    assert len(forms) == 7
    assert all(x.tobytes() in forms for x in p)


def test_group_orders_give_the_printed_aut(oracle):
    """Day-permutation group order x fixed-day |Aut| of the GPU's result equals
    the |Aut(M)| the reference prints, for the full k-systems of the fixtures
    (24 of the 396 K_10 factorizations, the uniform and perfect K_10 ones, the
    uniform K_8 one, the Kirkman solutions); the two factors equal the
    oracle's walk on K_8 and two Kirkman solutions; partial days give one day
    order and the oracle's fixed-day |Aut|; run() leaves them for
    group_order()."""
    kf = json.load(open(os.path.join(HERE, "golden", "kfact.json")))
    cases = [("k10", list(range(0, 396, 17))), ("u10", [0]), ("p10", [0]), ("u8", [0]),
             ("k15x3", list(range(7))), ("p15x3", [0, 1])]
    for name, idx in cases:
        k = 3 if name.endswith("x3") else 2
        mats = np.array([kf[name]["matrices"][i]["rounds"] for i in idx], np.uint8)
        v = mats.shape[2]
        print(f"{name}: {len(idx)} k-systems", flush=True)
        with canonizer(v, k) as cz:
            _, dord, ford = cz.run_batch(mats, orders=True)
            for j, i in enumerate(idx):
                assert int(dord[j]) * int(ford[j]) == kf[name]["matrices"][i]["aut"], (name, i)
            if name in ("u8", "k15x3"):
                for j in range(min(2, len(idx))):
                    _, od, of = oracle.cd_run_canonizer_orders(mats[j], v, k)
                    assert (int(dord[j]), int(ford[j])) == (od, of), (name, j)
                cz.run(mats[0])
                assert cz.group_order() == (int(dord[0]), int(ford[0]))
                part = np.ascontiguousarray(mats[:, :3])
                _, pd, pf = cz.run_batch(part, day=3, orders=True)
                for j in range(len(part)):
                    _, od, of = oracle.cd_run_canonizer_orders(part[j], v, k, 3)
                    assert od == 1 and (int(pd[j]), int(pf[j])) == (od, of), (name, j)


def test_all_printed_k10_factorizations():
    """All 396 printed 1-factorizations of K_10 (the complete list of their
    isomorphism classes), all 9 days: 396 different canonical matrices, and
    day-group order x fixed-day |Aut| equal to every printed |Aut(M)|."""
    kf = json.load(open(os.path.join(HERE, "golden", "kfact.json")))
    ms = kf["k10"]["matrices"]
    mats = np.array([m["rounds"] for m in ms], np.uint8)
    print(f"k10: {len(ms)} k-systems, 9! day orders each", flush=True)
    with canonizer(10, 2) as cz:
        forms, dord, ford = [], [], []
        for i in range(0, len(mats), 33):       # a line of progress per part
            f, d, o = cz.run_batch(mats[i:i + 33], orders=True)
            forms.append(f)
            dord.append(d)
            ford.append(o)
            print(f"k10: {i + len(f)} done", flush=True)
    forms = np.concatenate(forms)
    aut = np.concatenate(dord).astype(np.uint64) * np.concatenate(ford).astype(np.uint64)
    assert len({f.tobytes() for f in forms}) == len(ms)
    assert [int(a) for a in aut] == [int(m["aut"]) for m in ms]


@pytest.mark.parametrize("v,k,day,mat", [(2, 2, 0, [[0, 1]]), (3, 3, 0, [[2, 0, 1]]), (4, 2, 1, [[0, 3, 1, 2]]),
                                         (6, 2, 2, [[5, 0, 1, 4, 2, 3], [5, 1, 2, 0, 3, 4]])])
def test_smallest_k_systems(oracle, v, k, day, mat):
    """One group a day (v = k: one day, one column), single days, and a
    two-day prefix of K_6: forms and both group orders against the oracle."""
    m = np.array(mat, np.uint8)
    want, d_order, f_order = oracle.cd_run_canonizer_orders(m, v, k, day)
    with canonizer(v, k) as cz:
        got = cz.run(m, day=day)
        assert np.array_equal(got, want)
        assert cz.group_order() == (d_order, f_order)
