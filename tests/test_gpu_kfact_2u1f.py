"""GPU parity of the two-row canonizer of uniform 1-factorizations
(cnvCheck2U1F) against the oracle: verdict, witness index in the two-row
order, automorphisms counted and relabellings tried, on the printed uniform
1-factorizations of K_8, K_10, K_12 (and the perfect ones of K_10, K_14, one
cycle per pair), their prefixes, relabelled prefixes, and general partial
factorizations whose pairs of rounds have other cycle types; agreement with
the general canonizer; device path and rejected records."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden, p1f_candidates, u1f_candidates

pytestmark = pytest.mark.gpu


def oracle_2u1f(oracle, mats):
    out = {k: [] for k in ise_amd.KF_FIELDS}
    for m in mats:
        for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check_2u1f(m)):
            out[k].append(x)
    return {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in out.items()}


def assert_same(g, o, what):
    for k in ise_amd.KF_FIELDS:
        if not np.array_equal(g[k], o[k]):
            i = int(np.nonzero(g[k] != o[k])[0][0])
            raise AssertionError(f"{what}: {k} differs at {i}: gpu={g[k][i]} oracle={o[k][i]}")


def run(gpu_ctx, oracle, mats, what):
    rows, meta = ise_amd.kf_pack(mats)
    g = gpu_ctx.kf_check(rows, meta, two_row_u1f=True)
    assert_same(g, oracle_2u1f(oracle, mats), what)
    return g


def test_printed_factorizations_and_prefixes(gpu_ctx, oracle):
    print("case 2u1f printed U8/U10/U12/P10/P14 + prefixes", flush=True)
    kf = golden("kfact.json")
    keys = ("u8", "u10", "u12", "p10", "p14")
    mats = [np.array(m["rounds"], np.uint8) for k in keys for m in kf[k]["matrices"]]
    auts = [m["aut"] for k in keys for m in kf[k]["matrices"]]
    g = run(gpu_ctx, oracle, mats, "printed")
    assert (g["verdict"] == 1).all() and (g["n_auts"] == np.array(auts, np.uint64)).all()
    run(gpu_ctx, oracle, [m[:r] for m in mats for r in range(1, len(m) + 1)], "prefixes")


@pytest.mark.parametrize("key,count", [("u8", 400), ("u10", 1500), ("u12", 1500)])
def test_relabelled_prefixes(gpu_ctx, oracle, key, count):
    print(f"case 2u1f relabelled {key}", flush=True)
    mats = u1f_candidates(golden("kfact.json")[key], count, 0x2C + count, rmin=1)
    g = run(gpu_ctx, oracle, mats, key)
    assert (g["verdict"] == 1).any()


def test_other_cycle_types(gpu_ctx, oracle):
    """General partial factorizations: pairs of rounds with cycle lengths other
    than rounds 0, 1's end the test (not canonical, no witness)."""
    print("case 2u1f mixed cycle types", flush=True)
    rows, meta = ise_amd.kf_random(12, 2, 9, 1500, 0x2D)
    mats = [ise_amd.kf_unpack(rows[i, :int(meta[i]) & 0xff], 12) for i in range(len(meta))]
    k10 = [np.array(m["rounds"], np.uint8)[:r] for m in golden("kfact.json")["k10"]["matrices"][:60]
           for r in (3, 5, 9)]
    g = run(gpu_ctx, oracle, mats + k10, "mixed")
    assert ((g["verdict"] == 0) & (g["witness"] == np.uint64(2 ** 64 - 1))).any()


def test_agrees_with_general_canonizer(gpu_ctx):
    print("case 2u1f vs general U12/P14", flush=True)
    kf = golden("kfact.json")
    mats = u1f_candidates(kf["u12"], 1500, 0x2E) + p1f_candidates(kf["p14"], 1000, 0x2F)
    rows, meta = ise_amd.kf_pack(mats)
    g2 = gpu_ctx.kf_check(rows, meta, two_row_u1f=True)
    g = gpu_ctx.kf_check(rows, meta)
    assert np.array_equal(g2["verdict"], g["verdict"])
    c = g["verdict"] == 1
    assert c.any() and np.array_equal(g2["n_auts"][c], g["n_auts"][c])


def test_device_path_and_invalid_records(gpu_ctx, oracle):
    print("case 2u1f device + invalid", flush=True)
    mats = u1f_candidates(golden("kfact.json")["u10"], 300, 0x30, rmin=1)
    flipped = mats[0].copy()
    flipped[0, [0, 1]] = flipped[0, [1, 0]]
    # in normal form but no factorization: round 2 repeats round 0's pairs;
    # a player twice in round 0 (and one missing)
    full = np.array(golden("kfact.json")["u10"]["matrices"][0]["rounds"], np.uint8)
    repeated = np.concatenate([full[:2], full[:1]])
    twice = full[:3].copy()
    twice[0] = [0, 1, 2, 3, 4, 5, 6, 9, 8, 9]
    bad = [flipped, repeated, twice]
    rows, meta = ise_amd.kf_pack(mats + bad)
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    dev, res = gpu_ctx.kf_device_results(n)
    gpu_ctx.kf_check_2u1f_device(n, 16, d_rows, d_meta, res)
    gpu_ctx.sync()
    g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
    assert (g["verdict"][-len(bad):] == 3).all()
    assert_same({k: v[:-len(bad)] for k, v in g.items()}, oracle_2u1f(oracle, mats), "device")
    for b in list(dev.values()) + [d_rows, d_meta]:
        b.free()
    for mat in bad:
        r, m = ise_amd.kf_pack([mat])
        with pytest.raises(ise_amd.ISEError):
            gpu_ctx.kf_check(r, m, two_row_u1f=True)
