"""GPU parity of the two-row canonizer of perfect 1-factorizations
(cnvCheck2P1F) against the oracle: verdict, witness index in the two-row
order, automorphisms counted and relabellings tried, on the reference's printed
perfect 1-factorizations of K_10 and K_14, their prefixes, and relabelled
prefixes (rounds 0, 1 the standard pair); the verdicts and |Aut| agree with
the general canonizer; records that are not perfect or not in normal form are
rejected."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden, p1f_candidates

pytestmark = pytest.mark.gpu


def oracle_2p1f(oracle, mats):
    out = {k: [] for k in ise_amd.KF_FIELDS}
    for m in mats:
        for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check_2p1f(m)):
            out[k].append(x)
    return {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in out.items()}


def assert_same(g, o, what):
    for k in ise_amd.KF_FIELDS:
        if not np.array_equal(g[k], o[k]):
            i = int(np.nonzero(g[k] != o[k])[0][0])
            raise AssertionError(f"{what}: {k} differs at {i}: gpu={g[k][i]} oracle={o[k][i]}")


def test_printed_perfect_factorizations_and_prefixes(gpu_ctx, oracle):
    print("case 2p1f printed K10/K14 + prefixes", flush=True)
    kf = golden("kfact.json")
    mats = [np.array(m["rounds"], np.uint8) for key in ("p10", "p14") for m in kf[key]["matrices"]]
    auts = [m["aut"] for key in ("p10", "p14") for m in kf[key]["matrices"]]
    rows, meta = ise_amd.kf_pack(mats)
    g = gpu_ctx.kf_check(rows, meta, two_row_p1f=True)
    assert (g["verdict"] == 1).all()
    assert (g["n_auts"] == np.array(auts, np.uint64)).all()
    assert (g["n_tried"] == np.array([len(m) * (len(m) - 1) * m.shape[1] for m in mats], np.uint64)).all()
    prefixes = [m[:k] for m in mats for k in range(1, len(m) + 1)]
    rows, meta = ise_amd.kf_pack(prefixes)
    assert_same(gpu_ctx.kf_check(rows, meta, two_row_p1f=True), oracle_2p1f(oracle, prefixes), "prefixes")


@pytest.mark.parametrize("key,count", [("p10", 1500), ("p14", 4000)])
def test_relabelled_prefixes(gpu_ctx, oracle, key, count):
    print(f"case 2p1f relabelled {key}", flush=True)
    mats = p1f_candidates(golden("kfact.json")[key], count, 0x2F + count, rmin=1)
    rows, meta = ise_amd.kf_pack(mats)
    g = gpu_ctx.kf_check(rows, meta, two_row_p1f=True)
    o = oracle_2p1f(oracle, mats)
    assert_same(g, o, key)
    assert 0 < int((g["verdict"] == 1).sum()) < len(mats)


def test_agrees_with_general_canonizer(gpu_ctx):
    """Same verdicts as the general canonizer (all relabellings) and the same
    |Aut| for the canonical ones, on relabelled K_14 prefixes."""
    print("case 2p1f vs general K14", flush=True)
    mats = p1f_candidates(golden("kfact.json")["p14"], 3000, 0x14E, rmin=3)
    rows, meta = ise_amd.kf_pack(mats)
    g2 = gpu_ctx.kf_check(rows, meta, two_row_p1f=True)
    g = gpu_ctx.kf_check(rows, meta)
    assert np.array_equal(g2["verdict"], g["verdict"])
    c = g["verdict"] == 1
    assert c.any() and np.array_equal(g2["n_auts"][c], g["n_auts"][c])


def test_device_path_and_invalid_records(gpu_ctx, oracle):
    """The device entry point gives the host path's results; a factorization
    that is not perfect, one with a pair written larger player first and one
    with a player twice in a round get ISE_VERDICT_INVALID there and
    ISE_ERR_ARG on the host path."""
    print("case 2p1f device + invalid", flush=True)
    kf = golden("kfact.json")
    mats = p1f_candidates(kf["p14"], 500, 0xDE, rmin=1)
    not_p1f = np.array(kf["k10"]["matrices"][0]["rounds"], np.uint8)
    assert oracle.kf_check(not_p1f)[0] == 1
    flipped = mats[0].copy()
    flipped[0, [0, 1]] = flipped[0, [1, 0]]
    twice = np.array(kf["p14"]["matrices"][0]["rounds"], np.uint8)[:3].copy()
    twice[0] = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 13, 12, 13]   # in normal form, player 13 twice
    rows, meta = ise_amd.kf_pack(mats + [not_p1f, flipped, twice])
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    dev, res = gpu_ctx.kf_device_results(n)
    gpu_ctx.kf_check_2p1f_device(n, 16, d_rows, d_meta, res)
    gpu_ctx.sync()
    g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
    assert (g["verdict"][-3:] == 3).all()
    assert_same({k: v[:-3] for k, v in g.items()}, oracle_2p1f(oracle, mats), "device")
    ms = gpu_ctx.kf_time_2p1f_device(n, 16, d_rows, d_meta, res, 2)
    assert (ms > 0).all()
    for b in list(dev.values()) + [d_rows, d_meta]:
        b.free()
    for bad in (not_p1f, flipped, twice):
        r, m = ise_amd.kf_pack([bad])
        with pytest.raises(ise_amd.ISEError):
            gpu_ctx.kf_check(r, m, two_row_p1f=True)
