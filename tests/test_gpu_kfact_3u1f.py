"""GPU parity of cnvCheck3U1F for groups of 2 (the two-row U1F kernel with each
ordered pair's relabellings tried 32 times over and the run's cycle types)
against the oracle: verdict, witness index, automorphisms and relabellings
tried, on the printed uniform and perfect 1-factorizations and their
prefixes, relabelled prefixes, general partial factorizations with other
cycle types, rounds 0, 1 of a type the run does not allow; device path,
invalid records and malformed type lists."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden, u1f_candidates

pytestmark = pytest.mark.gpu


def oracle_3u1f(oracle, mats, types):
    out = {k: [] for k in ise_amd.KF_FIELDS}
    for m in mats:
        for k, x in zip(ise_amd.KF_FIELDS, oracle.kf_check_3u1f(m, types)):
            out[k].append(x)
    return {k: np.array(v, np.uint64 if k != "verdict" else np.uint8) for k, v in out.items()}


def assert_same(g, o, what):
    for k in ise_amd.KF_FIELDS:
        if not np.array_equal(g[k], o[k]):
            i = int(np.nonzero(g[k] != o[k])[0][0])
            raise AssertionError(f"{what}: {k} differs at {i}: gpu={g[k][i]} oracle={o[k][i]}")


def run(gpu_ctx, oracle, mats, types, what):
    rows, meta = ise_amd.kf_pack(mats)
    g = gpu_ctx.kf_check_3u1f(rows, meta, types)
    assert_same(g, oracle_3u1f(oracle, mats, types), what)
    return g


@pytest.mark.parametrize("key,cycles", [("u8", (4, 4)), ("u10", (4, 6)), ("u12", (6, 6)), ("p10", (10,)),
                                        ("p14", (14,))])
def test_printed_factorizations_and_prefixes(gpu_ctx, oracle, key, cycles):
    print(f"case 3u1f printed {key} + prefixes", flush=True)
    block = golden("kfact.json")[key]
    mats = [np.array(m["rounds"], np.uint8) for m in block["matrices"]]
    g = run(gpu_ctx, oracle, mats, [cycles], f"printed {key}")
    assert (g["verdict"] == 1).all()
    assert (g["n_auts"] == np.array([m["aut"] for m in block["matrices"]], np.uint64)).all()
    run(gpu_ctx, oracle, [m[:r] for m in mats for r in range(1, len(m) + 1)], [cycles], f"prefixes {key}")
    # no type given: one cycle through every player
    g = run(gpu_ctx, oracle, mats, None, f"default types {key}")
    assert ((g["verdict"] == 1) == (key[0] == "p")).all()


@pytest.mark.parametrize("key,cycles,count", [("u8", (4, 4), 400), ("u10", (4, 6), 1500), ("u12", (6, 6), 1500)])
def test_relabelled_prefixes(gpu_ctx, oracle, key, cycles, count):
    print(f"case 3u1f relabelled {key}", flush=True)
    mats = u1f_candidates(golden("kfact.json")[key], count, 0x3D + count, rmin=1)
    g = run(gpu_ctx, oracle, mats, [(4, 4, 4), cycles], key)
    assert (g["verdict"] == 1).any()


def test_other_cycle_types(gpu_ctx, oracle):
    print("case 3u1f mixed cycle types", flush=True)
    rows, meta = ise_amd.kf_random(12, 2, 9, 1500, 0x3E)
    mats = [ise_amd.kf_unpack(rows[i, :int(meta[i]) & 0xff], 12) for i in range(len(meta))]
    types = [(4, 8), (12,)]   # rounds 0, 1 of types 6:6 and 4:4:4 are not allowed
    g = run(gpu_ctx, oracle, mats, types, "mixed")
    assert ((g["verdict"] == 0) & (g["witness"] == np.uint64(2 ** 64 - 1)) & (g["n_tried"] == 0)).any()
    assert ((g["verdict"] == 0) & (g["witness"] == np.uint64(2 ** 64 - 1)) & (g["n_tried"] > 0)).any()


def test_device_path_invalid_records_and_types(gpu_ctx, oracle):
    print("case 3u1f device + invalid", flush=True)
    mats = u1f_candidates(golden("kfact.json")["u10"], 300, 0x3F, rmin=1)
    full = np.array(golden("kfact.json")["u10"]["matrices"][0]["rounds"], np.uint8)
    repeated = np.concatenate([full[:2], full[:1]])
    rows, meta = ise_amd.kf_pack(mats + [repeated])
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    dev, res = gpu_ctx.kf_device_results(n)
    gpu_ctx.kf_check_3u1f_device(n, 16, d_rows, d_meta, res, [(4, 6)])
    gpu_ctx.sync()
    g = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
    o = oracle_3u1f(oracle, mats, [(4, 6)])
    assert_same({k: v[:-1] for k, v in g.items()}, o, "device")
    assert g["verdict"][-1] == ise_amd.VERDICT_INVALID
    for b in list(dev.values()) + [d_rows, d_meta]:
        b.free()
    # malformed type lists are refused
    r1, m1 = ise_amd.kf_pack(mats[:4])
    for bad in ([(4, 5)], [(2, 8)], [(4, 4)] * 5, [(0, 4)], [(4, 4, 4, 4, 4)]):
        with pytest.raises((ise_amd.ISEError, ValueError)):
            gpu_ctx.kf_check_3u1f(r1, m1, bad)
