"""GPU k-Sys backjump position (ise_kf_results.player_index,
alldata::setPlayerIndex, kmSupport.cpp:303-356) against the oracle's, for the
general canonizer (split launch and one wave per candidate, host and device
entry points) and the two-row P1F and U1F canonizers."""
import numpy as np
import pytest

import ise_amd
from _helpers import golden

pytestmark = pytest.mark.gpu


def enumerator_shaped(oracle, rng, count):
    """Relabelled prefixes of the printed K_10 factorizations (round i pairs
    player 0 with i + 1, as the k-Sys enumerator builds them)."""
    k10 = [np.array(m["rounds"], np.uint8) for m in golden("kfact.json")["k10"]["matrices"]]
    mats = []
    for _ in range(count):
        full = oracle.kf_normalize(rng.permutation(10).astype(np.uint8)[k10[int(rng.integers(len(k10)))]])
        mats.append(full[:int(rng.integers(2, 10))])
    return ise_amd.kf_pack(mats)


def oracle_positions(oracle, rows, meta, mode):
    out = []
    for i in range(len(meta)):
        nr, n = int(meta[i]) & 0xff, (int(meta[i]) >> 8) & 0xff
        out.append(oracle.kf_player_index(ise_amd.kf_unpack(rows[i, :nr], n), mode)[0])
    return np.array(out, np.uint16)


@pytest.mark.parametrize("first,ranges", [(ise_amd.DEFAULT_KF_FIRST_CHUNKS, ise_amd.DEFAULT_KF_MAX_RANGES), (0, 1)])
def test_general_canonizer_backjump(gpu_ctx, oracle, first, ranges):
    print(f"case kfact backjump general split={first}/{ranges}", flush=True)
    rng = np.random.default_rng(0xB1)
    r1, m1 = enumerator_shaped(oracle, rng, 3000)
    r2, m2 = ise_amd.kf_random(12, 3, 10, 2000, 0xB12)
    r3, m3 = ise_amd.kf_random(16, 12, 15, 100, 0xB16)
    rows, meta = np.concatenate([r1, r2, r3]), np.concatenate([m1, m2, m3])
    gpu_ctx.kf_set_split(first, ranges)
    try:
        g = gpu_ctx.kf_check(rows, meta, backjump=True)
    finally:
        gpu_ctx.kf_set_split()
    o = oracle_positions(oracle, rows, meta, 0)
    bad = np.nonzero(g["player_index"] != o)[0]
    assert len(bad) == 0, (int(bad[0]), int(g["player_index"][bad[0]]), int(o[bad[0]]))
    start = ((meta & 0xff) * ((meta >> 8) & 0xff) - 3).astype(np.uint16)
    assert (g["player_index"][:3000] < start[:3000]).sum() > 2000   # the position moves for these


def test_two_row_canonizers_backjump(gpu_ctx, oracle):
    print("case kfact backjump two-row", flush=True)
    kf = golden("kfact.json")
    src, smeta = ise_amd.kf_pack([m["rounds"] for m in kf["p14"]["matrices"]])
    rows, meta = ise_amd.kf_relabel_p1f(src, smeta, 3000, 3, 0xB14)
    g = gpu_ctx.kf_check(rows, meta, two_row_p1f=True, backjump=True)
    assert np.array_equal(g["player_index"], oracle_positions(oracle, rows, meta, 1))
    src, smeta = ise_amd.kf_pack([kf["u12"]["matrices"][0]["rounds"]])
    rows, meta = ise_amd.kf_relabel_p1f(src, smeta, 3000, 3, 0xB12, uniform=True)
    g = gpu_ctx.kf_check(rows, meta, two_row_u1f=True, backjump=True)
    assert np.array_equal(g["player_index"], oracle_positions(oracle, rows, meta, 2))


def test_device_path_backjump(gpu_ctx, oracle):
    print("case kfact backjump device", flush=True)
    rng = np.random.default_rng(0xB2)
    rows, meta = enumerator_shaped(oracle, rng, 2048)
    n = len(meta)
    d_rows = gpu_ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = gpu_ctx.device_array(meta.shape, np.uint32, meta)
    d_head = gpu_ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    dev, _ = gpu_ctx.kf_device_results(n)
    dev["player_index"] = gpu_ctx.device_array((n,), np.uint16)
    s = ise_amd.kf_struct(dev, device=True)
    gpu_ctx.kf_time_batch_device(n, 16, d_rows, d_meta, s, d_head, 1)
    got = dev["player_index"].to_host()
    for b in list(dev.values()) + [d_rows, d_meta, d_head]:
        b.free()
    assert np.array_equal(got, oracle_positions(oracle, rows, meta, 0))
