"""Throughput of the batched canonicity test on one MI355X, plain and learned
row order.

Workload (one "step" = one pass of the hot path over one batch already
resident in HBM): the 2-(25,4,1) candidate stream at row depths 8..16 --
every candidate a beam-sampled run of the row-by-row enumerator tests there
(beam 64, seed 0x25041) followed by 2^20 seeded random valid extensions
(seed 0x25041) -- rows bit-packed as uint64 words (50 columns).  Each
candidate gets its verdict, witness row permutation, failing row,
rowToChange, |Aut| and the automorphism generators found (up to 16, each the
permutation it induces on the next row's column orbits).  The launch shape
comes from the autotuner, run on the same batch before timing.

  * plain order: one launch of the canonicity kernel per step (column-set
    search, candidates started deepest first);
  * learned order: per step, one launch of the lead-order kernel: the plain
    search on every candidate with a budget of 16 permutations; on the
    candidates it leaves open, the 16x128x128x1 bf16 MFMA network scores
    every row as the row of position 0 and a depth-first search tries them
    in that order (meeting an automorphism sends a candidate back to the
    plain search).  Its launch shape comes from its own autotuning run on the
    batch.  Verdicts, and every field of the canonical candidates, are
    checked equal to the plain order's.

`value` is the plain-order rate (the library's default order); the
learned-order rate and learned/plain ratio, both mean
permutations-to-rejection and rows compared, the 1-factorization kernels and
a full enumeration's designs/s are reported beside it.

Prints one JSON line.
  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "incidence-system-enumeration_amd"))

METRIC = "candidates canon-checked/sec on 1 MI355X, learned-order vs plain; bit-exact"
HBM_PEAK_GBS = 8000.0
# LDS: 128 bytes per clock per CU x 256 CUs x 2.4 GHz
LDS_PEAK_GBS = 128 * 256 * 2.4
GEN_CAP = 16           # automorphism generators returned per candidate in the timed launch
STRIDE = 16            # depth <= 16 rows per candidate
SEED = 0x25041
BEAM = 64
DMIN, DMAX = 8, 16
RESULT_BYTES = 1 + 1 + 1 + 8 + 32   # verdict, fail_row, row_to_change, |Aut|, witness row perm
PROFILE = os.path.join(ROOT, "profiles", "check_kernel64_2-25-4-1_stream_traffic.txt")
# the sources of the profiled kernel: PMC figures are reported only when the
# profile was taken of these exact sources
KERNEL_SOURCES = [os.path.join(ROOT, "incidence-system-enumeration_amd", "csrc", f)
                  for f in ("ise_wave.h", "ise_kernels.hip", "ise_internal.h")]


def kernel_sources_sha():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def build_batch(n_random, path):
    """Beam-sampled enumerator candidates (each level checked on the GPU)
    plus random extensions, written to `path`.  Runs in a child process so
    that the measured process launches nothing but the timed batches."""
    import ise_amd
    with ise_amd.Context(capacity=1 << 16, stride=32) as ctx:
        rows_e, meta_e, _ = ctx.enumerate_sampled(25, 4, 1, BEAM, SEED, DMIN, DMAX, stride=STRIDE, cap=1 << 21)
    rows_r, meta_r = ise_amd.random_extensions(25, 4, 1, n_random, DMIN, DMAX, SEED, stride=STRIDE)
    np.savez(path, rows=np.concatenate([rows_e, rows_r]), meta=np.concatenate([meta_e, meta_r]),
             n_enum=np.array([len(meta_e)]))


def load_batch(n_random):
    import subprocess
    import tempfile
    path = os.path.join(tempfile.gettempdir(), f"ise_bench_25_4_1_{n_random}_{SEED:x}_b{BEAM}.npz")
    if not os.path.exists(path):
        subprocess.run([sys.executable, os.path.abspath(__file__), "--build-batch", path,
                        "--random", str(n_random)], check=True)
    z = np.load(path)
    return z["rows"], z["meta"], int(z["n_enum"][0])


def cpu_baseline(rows, meta, gpu, seconds=15.0):
    """The oracle (a scalar port of the reference's checker), one thread, on a
    bounded sample of the same batch in seeded random order.  Its results on
    the sample are also compared with the GPU's (`gpu`: host copies of the
    timed launch's result fields)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    orc = _oracle.load()
    step = 2048
    done, t0 = 0, time.perf_counter()
    idx = np.random.default_rng(0).permutation(len(meta))
    while done < len(idx) and time.perf_counter() - t0 < seconds:
        sel = idx[done:done + step]
        o = orc.check_batch(np.ascontiguousarray(rows[sel]), np.ascontiguousarray(meta[sel]), gen_cap=GEN_CAP)
        for f, a in gpu.items():
            if f == "gens":
                ng = np.minimum(o["n_gens"], GEN_CAP)
                keep = np.arange(GEN_CAP)[None, :] < ng[:, None]
                if not np.array_equal(a[sel][keep], o[f][keep]):
                    raise RuntimeError("GPU generators differ from the oracle's on the CPU sample")
            elif not np.array_equal(a[sel], o[f]):
                raise RuntimeError(f"GPU field {f} differs from the oracle on the CPU sample")
        done += len(sel)
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="candidates/s", cores=1, kind="port",
                sample=f"{done} candidates of the same batch (seeded random order), oracle single "
                       f"thread, {dt:.1f} s; its results equal the GPU's on the sample")


def host_cores():
    """Cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_all_cores(rows, meta, seconds=15.0):
    """The same oracle on every host core (one thread per core, chunks of the
    same seeded random order; the oracle releases the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    orc = _oracle.load()
    cores = host_cores()
    step = 1024
    idx = np.random.default_rng(0).permutation(len(meta))
    chunks = [idx[i:i + step] for i in range(0, len(idx), step)]
    t0 = time.perf_counter()
    done = 0

    def work(sel):
        if time.perf_counter() - t0 > seconds:
            return 0
        orc.check_batch(np.ascontiguousarray(rows[sel]), np.ascontiguousarray(meta[sel]))
        return len(sel)

    with ThreadPoolExecutor(cores) as ex:
        for k in ex.map(work, chunks):
            done += k
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="candidates/s", cores=cores, kind="port",
                sample=f"{done} candidates of the same batch, {cores} threads, {dt:.1f} s")


def host_buffer_rate(rows, meta):
    """Whole batch through the host-buffer entry points (ise_submit /
    ise_collect: copies to and from the device included)."""
    import ise_amd
    with ise_amd.Context(capacity=1 << 18, stride=STRIDE) as ctx:
        ctx.check_batch(rows[:4096], meta[:4096])
        t0 = time.perf_counter()
        ctx.check_batch(rows, meta, fields=("verdict", "fail_row", "row_to_change", "group_order", "perm_row"))
        dt = time.perf_counter() - t0
    return {"candidates_per_s": len(meta) / dt, "seconds": dt,
            "note": "host arrays in and out, chunks of 2^18 candidates, PCIe copies included"}


def profiled_launch_figures():
    """HBM bytes and instruction counts of one plain launch of this workload,
    as measured by rocprofv3 PMC runs and committed under profiles/; empty
    unless the profile was taken of the current kernel sources."""
    out = {}
    if os.path.exists(PROFILE):
        for line in open(PROFILE):
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                k, v = k.strip(), v.strip()
                out[k] = v if k in ("sources_sha", "commit") else float(v)
    if out.get("sources_sha") != kernel_sources_sha():
        return {}
    return out


def kfact_rate(ctx, steps, warmup):
    """1-factorizations of K_12: 2^18 seeded partial factorizations of 3-10
    rounds, plain relabelling order, one split launch per step (a first look
    at 4 relabellings, a first pass of 256, then up to 1024 ranges per
    undecided candidate); the
    one-wave-per-candidate launch is timed beside it."""
    import ise_amd
    rows, meta = ise_amd.kf_random(12, 3, 10, 1 << 18, 0x12, stride=16)
    n = len(meta)
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    timed, results = {}, {}
    for name, split in (("one_pass", (0, 1)), ("split", (ise_amd.DEFAULT_KF_FIRST_CHUNKS,
                                                         ise_amd.DEFAULT_KF_MAX_RANGES))):
        dev, res = ctx.kf_device_results(n)
        ctx.kf_set_split(*split)
        if warmup:
            ctx.kf_time_batch_device(n, 16, d_rows, d_meta, res, d_head, warmup)
        timed[name] = float(np.mean(ctx.kf_time_batch_device(n, 16, d_rows, d_meta, res, d_head, steps)))
        results[name] = {k: dev[k].to_host() for k in ise_amd.KF_FIELDS}
        for b in dev.values():
            b.free()
    ctx.kf_set_split()
    for b in (d_rows, d_meta, d_head):
        b.free()
    ms = timed["split"]
    g = results["split"]
    if (g["verdict"] > 1).any():
        raise RuntimeError("invalid record in the K_12 batch")
    for k in ise_amd.KF_FIELDS:
        if not np.array_equal(g[k], results["one_pass"][k]):
            raise RuntimeError(f"K_12: split launch and one-pass launch differ in {k}")
    return {"workload": "K_12 partial 1-factorizations, 3-10 rounds, 2^18 seeded (seed 0x12), plain order",
            "candidates": n, "ms_per_step": ms, "candidates_per_s": n / (ms / 1e3),
            "canonical": int((g["verdict"] == 1).sum()),
            "relabellings_per_s": float(g["n_tried"].sum()) / (ms / 1e3),
            "split": {"first_chunks": ise_amd.DEFAULT_KF_FIRST_CHUNKS, "max_ranges": ise_amd.DEFAULT_KF_MAX_RANGES},
            "one_pass_ms_per_step": timed["one_pass"],
            "one_pass_equal_to_split": True,
            "_results": (rows, meta, g)}


def kfact_cpu_baseline(rows, meta, g, seconds=10.0):
    """The oracle's 1-factorization canonizer on one thread over a seeded
    sample of the K_12 batch; its results are compared with the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    import ise_amd
    orc = _oracle.load()
    idx = np.random.default_rng(0x12).permutation(len(meta))
    done, t0 = 0, time.perf_counter()
    for i in idx:
        if done >= 256 and time.perf_counter() - t0 > seconds:
            break
        nr = int(meta[i]) & 0xff
        v, w, a, t = orc.kf_check(ise_amd.kf_unpack(rows[i, :nr], 12))
        if (v, w, a, t) != (int(g["verdict"][i]), int(g["witness"][i]), int(g["n_auts"][i]), int(g["n_tried"][i])):
            raise RuntimeError(f"K_12 candidate {i}: GPU and oracle differ")
        done += 1
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="candidates/s", cores=1, kind="port",
                sample=f"{done} candidates of the K_12 batch (seeded random order), oracle single thread, "
                       f"{dt:.1f} s; results equal to the GPU's")


def kfact_p1f_rate(ctx, steps, warmup, uniform=False):
    """Two-row canonizers, one launch per step: perfect 1-factorizations of
    K_14 (cnvCheck2P1F): 2^18 seeded relabelled prefixes (2-13 rounds) of the
    23 perfect 1-factorizations of K_14 the reference prints; uniform=True:
    the uniform one of K_12 (cnvCheck2U1F, cycles 6:6), 2^18 relabelled
    prefixes (2-11 rounds).  Sources from tests/golden/kfact.json."""
    import ise_amd
    with open(os.path.join(ROOT, "tests", "golden", "kfact.json")) as f:
        block = json.load(f)["u12" if uniform else "p14"]
    src_rows, src_meta = ise_amd.kf_pack([m["rounds"] for m in block["matrices"]])
    rows, meta = ise_amd.kf_relabel_p1f(src_rows, src_meta, 1 << 18, 2, 0x12F1 if uniform else 0x14F1, stride=16,
                                        uniform=uniform)
    n = len(meta)
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    dev, res = ctx.kf_device_results(n)
    if warmup:
        ctx.kf_time_2p1f_device(n, 16, d_rows, d_meta, res, warmup, uniform=uniform)
    ms = float(np.mean(ctx.kf_time_2p1f_device(n, 16, d_rows, d_meta, res, steps, uniform=uniform)))
    verdict = dev["verdict"].to_host()
    tried = dev["n_tried"].to_host()
    for b in list(dev.values()) + [d_rows, d_meta]:
        b.free()
    if (verdict > 1).any():
        raise RuntimeError("invalid record in the two-row canonizer batch")
    workload = ("K_12 uniform 1-factorization (cycles 6:6), two-row canonizer: 2^18 relabelled prefixes "
                "(2-11 rounds, seed 0x12F1) of the printed one" if uniform else
                "K_14 perfect 1-factorizations, two-row canonizer: 2^18 relabelled prefixes "
                "(2-13 rounds, seed 0x14F1) of the 23 printed ones")
    return {"workload": workload,
            "candidates": n, "ms_per_step": ms, "candidates_per_s": n / (ms / 1e3),
            "canonical": int((verdict == 1).sum()), "relabellings_per_s": float(tried.sum()) / (ms / 1e3)}


def stream_rates(ctx, steps, warmup):
    """The other candidate-stream configurations, plain order, HBM-resident:
    2-(15,3,1) -- every candidate the enumerator tests at depth 4-15 plus 2^20
    seeded random valid extensions (seed 0x15031) whose depth is uniform in
    [4, 15] (rejection sampling of the depth, generated on the GPU); and the
    uneven batch -- a
    seeded 1:1:1 mix of 2-(15,7,3), 2-(16,6,2) and 2-(19,9,4) random valid
    extensions (2^18 each) at depth 5..v, one persistent launch."""
    import ise_amd
    out = {}
    fields = ("verdict", "fail_row", "row_to_change", "group_order", "perm_row")
    r1, m1 = ctx.enumerator_candidates(15, 3, 1, 4, 15, stride=16)
    t_gen = time.perf_counter()
    r2, m2 = ise_amd.random_extensions(15, 3, 1, 1 << 20, 4, 15, 0x15031, stride=16,
                                       attempts=ise_amd.RANDOM_MAX_ATTEMPTS, ctx=ctx)
    t_gen = time.perf_counter() - t_gen
    hist = np.bincount((m2 & 0xff).astype(np.int64), minlength=16)[4:16]
    parts = [("2-(15,3,1) enumerator candidates (depth 4-15) + 2^20 random extensions, depth uniform in 4-15",
              np.concatenate([r1, r2]), np.concatenate([m1, m2]), 16)]
    mix_r, mix_m = [], []
    for i, (v, k, lam) in enumerate(((15, 7, 3), (16, 6, 2), (19, 9, 4))):
        r, m = ise_amd.random_extensions(v, k, lam, 1 << 18, 5, v, 0xC3 + i, stride=32)
        mix_r.append(r)
        mix_m.append(m)
    mr, mm = np.concatenate(mix_r), np.concatenate(mix_m)
    perm = np.random.default_rng(0xC3).permutation(len(mm))
    parts.append(("1:1:1 mix of 2-(15,7,3), 2-(16,6,2), 2-(19,9,4) random extensions, 2^18 each, depth 5..v",
                  mr[perm], mm[perm], 32))
    for name, rows, meta, stride in parts:
        n = len(meta)
        d_rows = ctx.device_array(rows.shape, np.uint64, rows)
        d_meta = ctx.device_array(meta.shape, np.uint32, meta)
        d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
        dev, res = ctx.device_results(n, fields)
        ctx.set_launch(group_lanes=0)
        if warmup:
            ctx.time_batch_device(n, stride, d_rows, d_meta, res, d_head, warmup)
        ms = float(np.mean(ctx.time_batch_device(n, stride, d_rows, d_meta, res, d_head, steps)))
        verdict = dev["verdict"].to_host()
        if (verdict > 1).any():
            raise RuntimeError(f"search limit hit or invalid record in the {name} batch")
        out[name] = {"candidates": n, "ms_per_step": ms, "candidates_per_s": n / (ms / 1e3),
                     "canonical": int((verdict == 1).sum())}
        for b in list(dev.values()) + [d_rows, d_meta, d_head]:
            b.free()
    first = parts[0][0]
    out[first]["random_depth_histogram_4_15"] = [int(x) for x in hist]
    out[first]["random_generation_s"] = t_gen
    ctx.set_launch(group_lanes=64)
    return out


def enumeration_rate(cpu=True):
    """Full enumeration of 2-(13,3,1) and 2-(15,3,1): host row solver + GPU
    canonicity per level, and the device-resident enumeration (GPU row
    solver, canonicity and compaction); designs per second, wall clock.  With
    cpu: the oracle's depth-first enumeration of the same parameter sets on
    one host thread beside it (design lists compared)."""
    import ise_amd
    out = {}
    orc = None
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        orc = _oracle.load()
    with ise_amd.Context(capacity=1 << 16, stride=32) as ctx:
        for v, k, lam in ((13, 3, 1), (15, 3, 1)):
            ctx.enumerate_bibd(v, k, lam)          # warm-up (graphs, allocations)
            t0 = time.perf_counter()
            des, _, st, _ = ctx.enumerate_bibd(v, k, lam)
            dt = time.perf_counter() - t0
            ctx.enumerate_bibd_device(v, k, lam)
            t0 = time.perf_counter()
            des2, _, _, _ = ctx.enumerate_bibd_device(v, k, lam)
            dt2 = time.perf_counter() - t0
            if not np.array_equal(des, des2):
                raise RuntimeError("device enumeration differs from the host-solver enumeration")
            # the same host-solver enumeration with each level's launch replayed from a
            # captured graph (updated in place across levels and tail chunks)
            ctx.use_graphs(True)
            ctx.enumerate_bibd(v, k, lam)
            g0 = ctx.graph_stats()
            t0 = time.perf_counter()
            des3, _, _, _ = ctx.enumerate_bibd(v, k, lam)
            dt3 = time.perf_counter() - t0
            g1 = ctx.graph_stats()
            ctx.use_graphs(False)
            if not np.array_equal(des, des3):
                raise RuntimeError("the enumeration with graphs differs")
            # the same in the learned order (ise_mlp_use_lead): the same design list
            from ise_amd import learned
            lm = learned.load(ctx, learned.LEAD_WEIGHTS)
            lm.set_gate(learned.LEAD_GATE)
            lm.use_lead(True)
            ctx.enumerate_bibd(v, k, lam)
            t0 = time.perf_counter()
            des5, _, _, _ = ctx.enumerate_bibd(v, k, lam)
            dt5 = time.perf_counter() - t0
            lm.close()
            if sorted(map(tuple, des5.tolist())) != sorted(map(tuple, des.tolist())):
                raise RuntimeError("the enumeration in the learned order differs")
            rec = {"designs": len(des), "seconds": dt, "designs_per_s": len(des) / dt,
                   "candidates_per_s": st["candidates"] / dt,
                   "device_resident_seconds": dt2,
                   "device_resident_designs_per_s": len(des) / dt2,
                   "graphs_seconds": dt3, "graphs_designs_per_s": len(des) / dt3,
                   "learned_order_seconds": dt5, "learned_order_designs_per_s": len(des) / dt5,
                   "graph_replays": int(g1[1] - g0[1]), "graphs_cached": int(g1[0])}
            if orc is not None:
                t0 = time.perf_counter()
                odes, _, ost, _ = orc.enumerate(v, k, lam)
                dtc = time.perf_counter() - t0
                if sorted(map(tuple, odes.tolist())) != sorted(map(tuple, des.tolist())):
                    raise RuntimeError(f"2-({v},{k},{lam}): oracle and GPU design lists differ")
                rec["cpu_baseline"] = {"seconds": dtc, "designs_per_s": len(odes) / dtc,
                                       "candidates": int(ost[0]), "cores": 1, "kind": "port",
                                       "note": "oracle depth-first enumeration, one thread; same design list"}
            out[f"2-({v},{k},{lam})"] = rec
    return out


def cd_canonizer_rate(cpu=True, reps=3):
    """runCanonizer (ise_cd_*): the binary canonical form of k-systems, on the
    396 canonical 1-factorizations of K_10 the reference prints
    (tests/golden/kfact.json).  partial: their first 5 days (one canonization
    each, day order kept -- the call the k-Sys result check makes for every
    result); full: 8 of them with all 9 days (9! = 362880 day orders each,
    the largest form kept).  Host entry points (batch in, canonical bytes
    out); the oracle's partial forms are compared with the GPU's, and the full
    runs' |Aut(M)| (day-group order x fixed-day |Aut|) with the printed
    values.  CPU: the oracle on one thread -- partial on the whole batch, full
    as day orders canonized per second over 20000 seeded random orders of the
    first k-system (a whole 9-day walk takes about 40 s), divided by 9!."""
    import ise_amd
    with open(os.path.join(ROOT, "tests", "golden", "kfact.json")) as f:
        kf = json.load(f)
    mats = np.array([m["rounds"] for m in kf["k10"]["matrices"]], np.uint8)     # [396][9][10]
    out = {}
    with ise_amd.KSysCanonizer(10, 2) as cz:
        part = np.ascontiguousarray(mats[:, :5])
        cz.run_batch(part, day=5)
        t0 = time.perf_counter()
        for _ in range(reps):
            got_p = cz.run_batch(part, day=5)
        dt_p = (time.perf_counter() - t0) / reps
        # one call per k-system, as the reference calls runCanonizer: launch latency
        t0 = time.perf_counter()
        for i in range(32):
            cz.run(part[i], day=5)
        dt_single = (time.perf_counter() - t0) / 32
        full = np.ascontiguousarray(mats[:8])
        cz.run_batch(full[:3])          # warm-up at the full batch size (2^20 day orders)
        t0 = time.perf_counter()
        got_f, day_ord, fixed_ord = cz.run_batch(full, orders=True)
        dt_f = time.perf_counter() - t0
    aut = [int(d) * int(f) for d, f in zip(day_ord, fixed_ord)]
    printed = [int(m["aut"]) for m in kf["k10"]["matrices"][:len(full)]]
    if aut != printed:
        raise RuntimeError(f"runCanonizer: |Aut(M)| {aut} differs from the printed {printed}")
    out["partial"] = {"workload": "first 5 days of the 396 printed K_10 1-factorizations, day order kept",
                      "k_systems": len(part), "seconds": dt_p, "k_systems_per_s": len(part) / dt_p,
                      "single_call_ms": dt_single * 1e3}
    out["full"] = {"workload": "8 printed K_10 1-factorizations, all 9 days (362880 day orders each)",
                   "k_systems": len(full), "seconds": dt_f, "k_systems_per_s": len(full) / dt_f,
                   "day_orders_per_s": len(full) * 362880 / dt_f,
                   "aut_equal_to_printed": True}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        orc = _oracle.load()
        t0 = time.perf_counter()
        for i in range(len(part)):
            if not np.array_equal(orc.cd_run_canonizer(part[i], 10, 2, 5), got_p[i]):
                raise RuntimeError(f"runCanonizer: GPU and oracle differ on K_10 factorization {i}")
        dt = time.perf_counter() - t0
        out["partial"]["cpu_baseline"] = dict(value=len(part) / dt, unit="k-systems/s", cores=1, kind="port",
                                              sample=f"all {len(part)} k-systems, oracle single thread, "
                                                     f"{dt:.2f} s; results equal to the GPU's")
        rng = np.random.default_rng(0xcd)
        orders = np.array([rng.permutation(9) for _ in range(20000)], np.uint8)
        t0 = time.perf_counter()
        sample_best = orc.cd_canonize_orders(full[0], 10, 2, orders)
        dt = time.perf_counter() - t0
        if tuple(sample_best.reshape(-1)) > tuple(got_f[0].reshape(-1)):
            raise RuntimeError("runCanonizer: a day order's form exceeds the GPU's largest")
        out["full"]["cpu_baseline"] = dict(value=len(orders) / dt / 362880, unit="k-systems/s", cores=1,
                                           kind="port",
                                           sample=f"{len(orders)} seeded random day orders of the first "
                                                  f"k-system canonized by the oracle in {dt:.1f} s (the "
                                                  f"largest at most the GPU's form), rate / 9!")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--random", type=int, default=1 << 20)
    ap.add_argument("--group-lanes", type=int, default=0)
    ap.add_argument("--waves-per-block", type=int, default=4)
    ap.add_argument("--learned-gate", type=int, default=16,
                    help="plain permutations after which the network orders a candidate's first row (0: all)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="use --group-lanes/--waves-per-block instead of the autotuner's choice")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the K_12, enumeration and runCanonizer figures")
    ap.add_argument("--only-kfact", action="store_true", help="run only the K_12 figure (profiling)")
    ap.add_argument("--only-cd", action="store_true", help="run only the runCanonizer figure (profiling)")
    ap.add_argument("--build-batch", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run one device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU sharding is a design "
                                   "note (candidates are independent), not a measured path"}))
        sys.exit(2)

    if args.build_batch:
        build_batch(args.random, args.build_batch)
        return
    if args.only_kfact:
        import ise_amd
        with ise_amd.Context(capacity=1 << 18, stride=16) as kctx:
            print(json.dumps({"kfact_k12": kfact_rate(kctx, args.steps, args.warmup),
                              "kfact_p1f_k14": kfact_p1f_rate(kctx, args.steps, args.warmup),
                              "kfact_u1f_k12": kfact_p1f_rate(kctx, args.steps, args.warmup, uniform=True)}),
                  flush=True)
        return
    if args.only_cd:
        print(json.dumps({"cd_canonizer_k10": cd_canonizer_rate(cpu=not args.no_cpu)}), flush=True)
        return

    t_setup = time.time()
    rows, meta, n_enum = load_batch(args.random)
    import ise_amd
    from ise_amd import learned

    ctx = ise_amd.Context(capacity=1 << 10, stride=32)
    n = len(meta)
    depth = (meta & 0xff).astype(np.int64)
    print(f"bench: batch of {n} candidates ({n_enum} enumerator, {n - n_enum} random) built in "
          f"{time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)

    # inputs resident in HBM before the timed region
    d_rows = ctx.device_array(rows.shape, np.uint64, rows)
    d_meta = ctx.device_array(meta.shape, np.uint32, meta)
    d_head = ctx.device_array((ise_amd.QUEUE_WORDS,), np.uint32)
    fields = ("verdict", "fail_row", "row_to_change", "group_order", "perm_row", "n_gens", "gens")
    out, res = ctx.device_results(n, fields, gen_cap=GEN_CAP)
    tuned = None
    if args.autotune:
        # launch shape from the autotuner, on this batch (2-(25,4,1) needs 64 lanes: 50 columns)
        best, table = ctx.autotune(n, STRIDE, d_rows, d_meta, res, d_head, max_depth=DMAX, max_cols=50,
                                   repeats=2)
        tuned = {"best": best, "ms_table_64_lanes_1_2_4_8_waves": [float(x) for x in table[2]]}
        print(f"bench: autotuned {best}", file=sys.stderr, flush=True)
    else:
        ctx.set_launch(group_lanes=args.group_lanes, waves_per_block=args.waves_per_block)

    # ---- plain order: the timed steps (every result field of the search,
    # automorphism generators included)
    if args.warmup:
        ctx.time_batch_device(n, STRIDE, d_rows, d_meta, res, d_head, args.warmup)
    t0 = time.perf_counter()
    per_step_ms = ctx.time_batch_device(n, STRIDE, d_rows, d_meta, res, d_head, args.steps)
    wall = time.perf_counter() - t0
    ms = float(np.mean(per_step_ms))
    value = n / (ms / 1e3)
    verdict = out["verdict"].to_host()
    gpu_host = {f: out[f].to_host() for f in fields}
    print(f"bench: plain {ms:.2f} ms/step", file=sys.stderr, flush=True)
    # the same without generators
    nout, nres = ctx.device_results(n, fields[:5])
    if args.warmup:
        ctx.time_batch_device(n, STRIDE, d_rows, d_meta, nres, d_head, 1)
    ms_nogens = float(np.mean(ctx.time_batch_device(n, STRIDE, d_rows, d_meta, nres, d_head,
                                                    max(3, args.steps // 2))))
    for f in fields[:5]:
        if not np.array_equal(nout[f].to_host(), gpu_host[f]):
            raise RuntimeError(f"{f} differs between the launches with and without generators")

    # statistics from one extra, untimed launch with the counters compiled in
    cout, cres = ctx.device_results(n, ("verdict", "counters"))
    ctx.check_batch_device(n, STRIDE, d_rows, d_meta, cres, d_head)
    counters = cout["counters"].to_host()
    if not np.array_equal(cout["verdict"].to_host(), verdict):
        raise RuntimeError("verdicts differ between the counted and the timed kernel")
    if (verdict > 1).any():
        raise RuntimeError("search limit hit or invalid record in the bench batch")
    rej = verdict == 0

    # ---- learned order: one launch per step of the lead-order kernel (the
    # plain search up to the gate, the network's order of the first row on
    # the candidates it leaves open, every result field)
    learned_line = None
    if not args.no_learned:
        model = learned.load(ctx, learned.LEAD_WEIGHTS)
        model.set_gate(args.learned_gate)
        plain_launch = ctx.get_launch()
        lead_tuned = None
        if args.autotune:
            # the lead-order launch's own shape, on this batch (one candidate per wave)
            best_wpb, ltable = learned.autotune_lead(ctx, model, n, STRIDE, d_rows, d_meta, res, d_head)
            lead_tuned = {"best_waves_per_block": best_wpb, "ms_table_1_2_4_8_waves": ltable}
            print(f"bench: lead order autotuned to {best_wpb} waves per workgroup", file=sys.stderr, flush=True)
        if args.warmup:
            model.time_lead(n, STRIDE, d_rows, d_meta, res, d_head, args.warmup)
        lms = model.time_lead(n, STRIDE, d_rows, d_meta, res, d_head, args.steps)
        lhost = {f: out[f].to_host() for f in fields}
        if not np.array_equal(lhost["verdict"], verdict):
            raise RuntimeError("the learned order changed a verdict")
        canon = verdict == 1
        for f in fields:
            a, b = lhost[f], gpu_host[f]
            if f == "gens":
                keep = (np.arange(GEN_CAP)[None, :] < np.minimum(gpu_host["n_gens"], GEN_CAP)[:, None]) & \
                    canon[:, None]
                same = np.array_equal(a[keep], b[keep])
            else:
                same = np.array_equal(a[canon], b[canon])
            if not same:
                raise RuntimeError(f"the learned order changed {f} of a canonical candidate")
        model.check_lead(n, STRIDE, d_rows, d_meta, cres, d_head)
        lcount = cout["counters"].to_host()
        if not np.array_equal(cout["verdict"].to_host(), verdict):
            raise RuntimeError("the learned order changed a verdict (counted launch)")
        lmean = float(np.mean(lms))
        learned_line = {
            "value": n / (lmean / 1e3), "ms_per_step": lmean,
            "order": "learned order of the first row (lead order), one launch per step",
            "gate_perms": args.learned_gate,
            "weights": "incidence-system-enumeration_amd/ise_amd/lead_order_weights.npy",
            "mean_perms_to_reject": float(lcount[rej, 0].mean()),
            "mean_rows_to_reject": float(lcount[rej, 1].mean()),
            "rows_compared": float(lcount[:, 1].astype(np.float64).sum()),
            "verdicts_equal_to_plain": True,
            "canonical_fields_equal_to_plain": True,
            "autotune": lead_tuned,
        }
        print(f"bench: learned {lmean:.2f} ms/step", file=sys.stderr, flush=True)
        model.close()
        ctx.set_launch(**plain_launch)

    # algorithmic HBM bytes of one plain launch: candidate rows + meta in, result
    # record out (generators: 64 bytes per stored automorphism)
    ngens = gpu_host["n_gens"].astype(np.int64)
    stored = np.minimum(ngens, GEN_CAP)
    in_bytes = float((depth * 8 + 4).sum())
    hbm_bytes = in_bytes + n * (RESULT_BYTES + 2) + 64.0 * float(stored.sum())
    achieved_gbs = hbm_bytes / (ms / 1e3) / 1e9
    # algorithmic LDS bytes of the search: per row compared the 8-byte packed
    # row, per matrix cell (row, column) examined up to the decision 8 bytes,
    # plus the candidate stream
    rows_cmp = float(counters[:, 1].astype(np.float64).sum())
    cells_cmp = float(counters[:, 3].astype(np.float64).sum())
    orbits_cmp = float(counters[:, 2].astype(np.float64).sum())
    lds_bytes = 8.0 * (rows_cmp + cells_cmp) + in_bytes
    lds_gbs = lds_bytes / (ms / 1e3) / 1e9
    prof = profiled_launch_figures()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded beam-sampled enumerator candidates + random valid extensions)",
        "config": {
            "workload": "2-(25,4,1) candidate stream at depth 8-16: beam-sampled enumerator candidates "
                        "(beam 64) + 2^20 seeded random valid extensions; value = plain order",
            "candidates": n, "enumerator_candidates": n_enum, "random_candidates": n - n_enum,
            "seed": SEED, "order": "plain", "group_lanes": ctx.get_launch()["group_lanes"] or 64,
            "waves_per_block": ctx.get_launch()["waves_per_block"], "gen_cap": GEN_CAP,
        },
        # the search is integer work on LDS and registers: its roofline is LDS
        # bandwidth over the algorithmic LDS bytes; the HBM fraction is kept beside it
        "roofline": {
            "bound": "lds", "achieved": lds_gbs, "peak": LDS_PEAK_GBS, "unit": "GB/s",
            "frac": lds_gbs / LDS_PEAK_GBS,
            "traffic": (prof["fetch_bytes"] + prof["write_bytes"]) if prof else None,
            "lds": {"algorithmic_bytes": lds_bytes, "rows_compared": rows_cmp, "cells_examined": cells_cmp,
                    "column_orbits_compared": orbits_cmp,
                    "achieved_gbs": lds_gbs, "peak_gbs": LDS_PEAK_GBS, "frac": lds_gbs / LDS_PEAK_GBS,
                    "peak_note": "128 B/clk/CU x 256 CUs x 2.4 GHz"},
            "hbm": {"algorithmic_bytes": hbm_bytes, "achieved_gbs": achieved_gbs, "peak_gbs": HBM_PEAK_GBS,
                    "frac": achieved_gbs / HBM_PEAK_GBS},
            "kernel": "check_kernel<64,false,false,true> (column-set search)",
            "binding": "instruction issue (scalar and vector, near even) and the dependency latency of "
                       "each search step",
        },
        # what binds: instruction issue (rocprofv3 SQ counters of one plain launch of the
        # same kernel sources; null when the committed profile is of other sources)
        "issue": ({"valu_frac": prof["valu_insts"] / prof["cu_cycles"],
                   "salu_frac": prof["salu_insts"] / prof["cu_cycles"],
                   "peak": "1 wave64 VALU + 1 SALU instruction per CU per cycle",
                   "sources_sha": prof["sources_sha"],
                   "source": "profiles/check_kernel64_2-25-4-1_stream.md"} if prof else None),
        "kernel_sources_sha": kernel_sources_sha(),
        "generators": {"gen_cap": GEN_CAP, "candidates_with_generators": int((ngens > 0).sum()),
                       "generators_found": int(ngens.sum()), "candidates_over_cap": int((ngens > GEN_CAP).sum()),
                       "ms_per_step_without_generators": ms_nogens},
        "autotune": tuned,
        "canonical": int((verdict == 1).sum()),
        "noncanonical": int(rej.sum()),
        "mean_perms_to_reject_plain": float(counters[rej, 0].mean()) if rej.any() else 0.0,
        "mean_rows_to_reject_plain": float(counters[rej, 1].mean()) if rej.any() else 0.0,
        "learned": learned_line,
        "learned_over_plain": (learned_line["value"] / value) if learned_line else None,
        "rows_compared_per_s": rows_cmp / (ms / 1e3),
        "rows_compared_plain": rows_cmp,
        "rows_compared_learned": learned_line["rows_compared"] if learned_line else None,
        "wall_s_timed": wall,
        "ms_per_step_min": float(np.min(per_step_ms)),
    }
    if args.copy_bytes:
        src = ctx.device_array((args.copy_bytes,), np.uint8)
        dst = ctx.device_array((args.copy_bytes,), np.uint8)
        ctx._check(ctx._L.ise_copy_on_device(ctx.h, dst.ptr, src.ptr, args.copy_bytes))
        src.free()
        dst.free()
    for buf in list(out.values()) + list(nout.values()) + list(cout.values()) + [d_rows, d_meta, d_head]:
        buf.free()
    kf_sample = None
    if not args.no_extra:
        line["kfact_k12"] = kfact_rate(ctx, max(3, args.steps // 4), 1)
        kf_sample = line["kfact_k12"].pop("_results")
        line["kfact_p1f_k14"] = kfact_p1f_rate(ctx, max(3, args.steps // 4), 1)
        line["kfact_u1f_k12"] = kfact_p1f_rate(ctx, max(3, args.steps // 4), 1, uniform=True)
        line["streams"] = stream_rates(ctx, max(3, args.steps // 4), 1)
        line["enumeration"] = enumeration_rate(cpu=not args.no_cpu)
        line["cd_canonizer_k10"] = cd_canonizer_rate(cpu=not args.no_cpu)
        line["host_buffers"] = host_buffer_rate(rows, meta)
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(rows, meta, gpu_host)
        line["cpu_all_cores"] = cpu_all_cores(rows, meta)
        if kf_sample is not None:
            line["kfact_k12"]["cpu_baseline"] = kfact_cpu_baseline(*kf_sample)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
