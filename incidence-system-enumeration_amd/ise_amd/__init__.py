"""Python view of the MI355X batched canonicity test (libise_amd.so).

The classes mirror the reference's host interface:

* `CanonicityChecker.TestCanonicity` follows
  `CCanonicityChecker::TestCanonicity(nRowMax, pCanonParam, outInfo, ...)`
  (Source/CanonicityChecker.h:58) and exposes `permRow()`, `permCol()`,
  `groupOrder()`, `orbits()` like the reference accessors (:61-70, :115).
* `GPUCanonChecker.assignChecker` / `LaunchCanonicityTesting` follow
  `CGPU_CanonChecker` (Source/MatrixCanonChecker.h:248-329).
* `Context.check_batch` is the batched form used by the enumerator and the
  benchmark.

All work runs in the HIP library; importing this module without the built
library, or using it without a gfx950 device, raises instead of falling back
to any CPU path.
"""
import ctypes
import os

import numpy as np

from ._abi import (MAX_COLS, MAX_ROWS, SAVE_COL_PERMUTATIONS, SAVE_ROW_PERMUTATIONS,
                   SAVE_ROW_TO_CHANGE, VERDICT_CANONICAL, VERDICT_INVALID, VERDICT_LIMIT,
                   VERDICT_NONCANONICAL, ResultsStruct, alloc_results, pack_meta, pack_rows,
                   results_struct, unpack_rows)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libise_amd.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NODEV = -3
ERR_CAPACITY = -4
ERR_STATE = -5
DEFAULT_DEFER_PERMS = 0
DEFAULT_DEFER_PERMS2 = 0
DEFAULT_LEARNED_GATE = 16
DEFAULT_KF_FIRST_CHUNKS = 4   # ISE_DEFAULT_KF_FIRST_CHUNKS
DEFAULT_KF_MAX_RANGES = 1024   # ISE_DEFAULT_KF_MAX_RANGES

ORDER_PLAIN = 0     # the reference's search alone
ORDER_HUNT = 1      # budgeted witness hunt in increasing row order, then the plain search
ORDER_LEARNED = 2   # budgeted witness hunt in a per-candidate order table, then the plain search
ORDER_LEAD = 3      # the learned order of the first row (RowOrderModel.use_lead)

QUEUE_WORDS = 256   # ISE_QUEUE_WORDS: device words behind a work-queue argument
NOALLOC_MAX_CANDIDATES = 1 << 20   # ISE_NOALLOC_MAX_CANDIDATES
CANONIZE_MAX_ITERS = 4096   # ISE_CANONIZE_MAX_ITERS

_lib = None


class ISEError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"ise status {code}: {msg}")
        self.code = code


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ISEError(ERR_NODEV, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P, U32, U64, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.ise_open.argtypes = [ctypes.POINTER(P), I, U32, U32]
        L.ise_close.argtypes = [P]
        L.ise_set_launch.argtypes = [P, P]
        L.ise_get_launch.argtypes = [P, P]
        L.ise_assign.argtypes = [P, P, U32, U32, U32, U32]
        L.ise_num_assigned.argtypes = [P, ctypes.POINTER(U32)]
        L.ise_submit.argtypes = [P, U32, U32, P, P, P, U32]
        L.ise_collect.argtypes = [P, ctypes.POINTER(ResultsStruct)]
        L.ise_test_canonicity.argtypes = [P, P, U32, U32, U32, U32, U32, ctypes.POINTER(ResultsStruct)]
        L.ise_check_batch_device.argtypes = [P, U32, U32, P, P, P, ctypes.POINTER(ResultsStruct), P, P]
        L.ise_check_batch_device_timed.argtypes = [P, U32, U32, P, P, P, ctypes.POINTER(ResultsStruct), P,
                                                   U32, P]
        L.ise_check_batch_device_noalloc.argtypes = [P, U32, U32, P, P, P, ctypes.POINTER(ResultsStruct), P]
        L.ise_canonize_batch.argtypes = [P, U32, U32, P, P, P, P, P, P, U32]
        L.ise_canonize_batch_device.argtypes = [P, U32, U32, P, P, P, P, P, P, U32]
        L.ise_canonize_batch_device_part.argtypes = [P, U32, U32, P, P, U64, P, P, P, P, U32]
        L.ise_cd_create_canonizer.argtypes = [ctypes.POINTER(P), I, U32, U32]
        L.ise_cd_release_canonizer.argtypes = [P]
        L.ise_cd_run_canonizer.argtypes = [P, P, U32, I, ctypes.POINTER(P), ctypes.POINTER(U32)]
        L.ise_cd_run_canonizer_batch.argtypes = [P, U32, P, I, P, P, P]
        L.ise_cd_run_canonizer_device.argtypes = [P, U32, P, I, P, P, P]
        L.ise_cd_group_order.argtypes = [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]
        L.ise_cd_error.argtypes = [P]
        L.ise_cd_error.restype = ctypes.c_char_p
        L.ise_improved_solution.argtypes = [P, U32, U32, P, P, P, U32, P, ctypes.POINTER(U32)]
        L.ise_canonize_matrix.argtypes = [P, P, U32, U32, U32, P, ctypes.POINTER(U64), ctypes.POINTER(U32)]
        L.ise_device_alloc.argtypes = [P, ctypes.POINTER(P), U64]
        L.ise_device_free.argtypes = [P, P]
        L.ise_copy_to_device.argtypes = [P, P, P, U64]
        L.ise_copy_to_host.argtypes = [P, P, P, U64]
        L.ise_copy_on_device.argtypes = [P, P, P, U64]
        L.ise_set_order.argtypes = [P, I, U32, P, U32]
        L.ise_use_graphs.argtypes = [P, I]
        L.ise_report_columns.argtypes = [P, I]
        L.ise_set_defer.argtypes = [P, U32, U32]
        L.ise_set_schedule.argtypes = [P, I]
        L.ise_kf_set_split.argtypes = [P, U32, U32]
        L.ise_graph_stats.argtypes = [P, ctypes.POINTER(U32), ctypes.POINTER(U64)]
        L.ise_graph_updates.argtypes = [P, ctypes.POINTER(U64)]
        L.ise_autotune.argtypes = [P, U32, U32, P, P, P, ctypes.POINTER(ResultsStruct), P, U32, U32, U32, P, P]
        L.ise_sync.argtypes = [P]
        L.ise_last_error.argtypes = [P]
        L.ise_last_error.restype = ctypes.c_char_p
        L.ise_last_kernel_ms.argtypes = [P]
        L.ise_last_kernel_ms.restype = ctypes.c_float
        L.ise_enumerate_bibd.argtypes = [P, I, I, I, P, P, ctypes.c_long, P, P]
        L.ise_enumerate_bibd.restype = ctypes.c_long
        L.ise_enumerate_bibd_pruned.argtypes = [P, I, I, I, P, P, ctypes.c_long, P, P]
        L.ise_enumerate_bibd_pruned.restype = ctypes.c_long
        L.ise_enumerate_bibd_candidates.argtypes = [P, I, I, I, I, I, U32, P, P, ctypes.c_long,
                                                    ctypes.POINTER(ctypes.c_long)]
        L.ise_enumerate_bibd_candidates.restype = ctypes.c_long
        L.ise_enumerate_bibd_sampled.argtypes = [P, I, I, I, U32, U64, I, I, U32, P, P, ctypes.c_long,
                                                 ctypes.POINTER(ctypes.c_long), P]
        L.ise_enumerate_bibd_sampled.restype = ctypes.c_long
        L.ise_random_extensions.argtypes = [I, I, I, U64, I, I, U64, U32, P, P]
        L.ise_random_extensions_ex.argtypes = [P, I, I, I, U64, I, I, U64, U32, U32, P, P]
        F = ctypes.c_float
        L.ise_mlp_create.argtypes = [P, ctypes.POINTER(P), U32, U64]
        L.ise_mlp_destroy.argtypes = [P]
        L.ise_mlp_get_params.argtypes = [P, P]
        L.ise_mlp_set_params.argtypes = [P, P]
        L.ise_mlp_features.argtypes = [P, U32, U32, P, P, U32, P]
        L.ise_mlp_forward.argtypes = [P, P, U32, P]
        L.ise_mlp_gradient.argtypes = [P, P, P, P, U32, P, ctypes.POINTER(F)]
        L.ise_mlp_train_step.argtypes = [P, P, P, P, U32, F, ctypes.POINTER(F)]
        L.ise_mlp_fit.argtypes = [P, P, P, P, U64, U32, U32, F, P]
        L.ise_mlp_order.argtypes = [P, U32, U32, P, P, U32, P, P]
        L.ise_mlp_check_batch_device_timed.argtypes = [P, U32, U32, P, P, U32, P, U32,
                                                       ctypes.POINTER(ResultsStruct), P, U32, P, P]
        L.ise_mlp_witness_labels.argtypes = [U32, U32, P, P, P, P, P, P]
        L.ise_mlp_set_gate.argtypes = [P, U32]
        L.ise_mlp_lead_features.argtypes = [P, U32, U32, P, P, P]
        L.ise_mlp_use_lead.argtypes = [P, I]
        L.ise_mlp_check_lead_device.argtypes = [P, U32, U32, P, P, ctypes.POINTER(ResultsStruct), P, P, P, P]
        L.ise_mlp_check_lead_device_timed.argtypes = [P, U32, U32, P, P, ctypes.POINTER(ResultsStruct), P, U32, P]
        L.ise_group_closure.argtypes = [P, U32, U32, P, U64, ctypes.POINTER(U64)]
        L.ise_adjust_generators.argtypes = [P, U32, U32, P, U32, P]
        L.ise_solution_orbits.argtypes = [P, P, U64, U32, P, U32, P, P]
        L.ise_solve_rows_device.argtypes = [P, I, I, I, U32, U32, P, I, P, P, P, U64, ctypes.POINTER(U64)]
        L.ise_enumerate_bibd_device.argtypes = [P, I, I, I, P, P, ctypes.c_long, P, P]
        L.ise_enumerate_bibd_device.restype = ctypes.c_long
        L.ise_format_matrix.argtypes = [P, I, I, U64, U64, I, U64, P, U64, ctypes.POINTER(U64)]
        L.ise_kf_format_matrix.argtypes = [P, I, I, U64, U64, P, U64, ctypes.POINTER(U64)]
        L.ise_format_aut_table.argtypes = [I, I, I, I, I, U32, P, P, P, P, U64, ctypes.POINTER(U64)]
        L.ise_kf_check.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct)]
        L.ise_kf_check_batch_device.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct), P, P]
        L.ise_kf_check_batch_device_timed.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct), P,
                                                      U32, P]
        L.ise_kf_random.argtypes = [I, I, I, U64, U64, U32, P, P]
        L.ise_kf_check_2p1f.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct)]
        L.ise_kf_relabel_p1f.argtypes = [P, P, U32, U32, U64, I, U64, U32, P, P]
        L.ise_kf_relabel_u1f.argtypes = [P, P, U32, U32, U64, I, U64, U32, P, P]
        L.ise_kf_check_2u1f.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct)]
        L.ise_kf_check_3u1f.argtypes = [P, U32, U32, P, P, P, U32, ctypes.POINTER(KfResultsStruct)]
        L.ise_kf_check_3u1f_batch_device.argtypes = [P, U32, U32, P, P, P, U32, ctypes.POINTER(KfResultsStruct), P]
        L.ise_kf_check_2u1f_batch_device.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct), P]
        L.ise_kf_check_2u1f_batch_device_timed.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct),
                                                           U32, P]
        L.ise_kf_check_2p1f_batch_device.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct), P]
        L.ise_kf_check_2p1f_batch_device_timed.argtypes = [P, U32, U32, P, P, ctypes.POINTER(KfResultsStruct),
                                                           U32, P]
        _lib = L
    return _lib


class KfResultsStruct(ctypes.Structure):
    _fields_ = [("verdict", ctypes.c_void_p), ("witness", ctypes.c_void_p), ("n_auts", ctypes.c_void_p),
                ("n_tried", ctypes.c_void_p), ("player_index", ctypes.c_void_p)]


KF_FIELDS = ("verdict", "witness", "n_auts", "n_tried")


def kf_alloc(n, backjump=False):
    out = {"verdict": np.zeros(n, np.uint8), "witness": np.zeros(n, np.uint64),
           "n_auts": np.zeros(n, np.uint64), "n_tried": np.zeros(n, np.uint64)}
    if backjump:
        out["player_index"] = np.zeros(n, np.uint16)
    return out


def kf_struct(arrays, device=False):
    s = KfResultsStruct()
    for k in KF_FIELDS + ("player_index",):
        a = arrays.get(k)
        setattr(s, k, None if a is None else (a.ptr if device else a.ctypes.data))
    return s


def kf_pack(rounds_list, stride=16):
    """List of matrices (nrounds x n players) -> (rounds [m][stride] uint64, meta [m])."""
    m = len(rounds_list)
    rows = np.zeros((m, stride), np.uint64)
    meta = np.zeros(m, np.uint32)
    for i, mat in enumerate(rounds_list):
        mat = np.asarray(mat, dtype=np.uint64)
        nr, n = mat.shape
        shifts = (np.arange(n, dtype=np.uint64) * np.uint64(4))
        rows[i, :nr] = np.bitwise_or.reduce(mat << shifts, axis=1)
        meta[i] = nr | (n << 8)
    return rows, meta


def kf_unpack(words, n):
    words = np.asarray(words, dtype=np.uint64)
    return ((words[:, None] >> (np.arange(n, dtype=np.uint64) * np.uint64(4))) & np.uint64(15)).astype(np.uint8)


class LaunchCfg(ctypes.Structure):
    _fields_ = [("group_lanes", ctypes.c_uint32), ("waves_per_block", ctypes.c_uint32),
                ("blocks", ctypes.c_uint32), ("max_perms", ctypes.c_uint32)]


def _ptr(a):
    return None if a is None else a.ctypes.data


def design_params(v, k, lam):
    """(v, b, r, k, lambda) of a 2-design; raises if not admissible."""
    if (lam * (v - 1)) % (k - 1):
        raise ValueError("r is not an integer")
    r = lam * (v - 1) // (k - 1)
    if (v * r) % k:
        raise ValueError("b is not an integer")
    return v, v * r // k, r, k, lam


class Context:
    """One device context (stream + buffers) for batches of candidates."""

    def __init__(self, capacity=1 << 16, stride=32, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.ise_open(ctypes.byref(h), device, capacity, stride)
        if rc != OK:
            raise ISEError(rc, "ise_open failed (no gfx950 device?)")
        self.h = h
        self.capacity = capacity
        self.stride = stride

    def close(self):
        if self.h:
            self._L.ise_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise ISEError(rc, self._L.ise_last_error(self.h).decode(errors="replace"))
        return rc

    def set_launch(self, group_lanes=0, waves_per_block=4, blocks=0, max_perms=1 << 26):
        cfg = LaunchCfg(group_lanes, waves_per_block, blocks, max_perms)
        self._check(self._L.ise_set_launch(self.h, ctypes.byref(cfg)))

    def set_defer(self, first_pass_perms=DEFAULT_DEFER_PERMS, second_pass_perms=DEFAULT_DEFER_PERMS2):
        """Launch in passes that defer searches longer than first_pass_perms,
        then second_pass_perms (0: one launch in queue order)."""
        self._check(self._L.ise_set_defer(self.h, first_pass_perms, second_pass_perms))

    def set_schedule(self, deepest_first=True):
        """Start the candidates of a batch deepest first (default) or in index order."""
        self._check(self._L.ise_set_schedule(self.h, 1 if deepest_first else 0))

    def get_launch(self):
        cfg = LaunchCfg()
        self._check(self._L.ise_get_launch(self.h, ctypes.byref(cfg)))
        return dict(group_lanes=cfg.group_lanes, waves_per_block=cfg.waves_per_block,
                    blocks=cfg.blocks, max_perms=cfg.max_perms)

    def check_batch(self, rows, meta, cells=None, gen_cap=0, fields=None):
        """Canonicity of a batch: rows [n][stride] uint64, meta [n] uint32.
        Returns the result arrays (see include/ise.h)."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n = len(meta)
        if rows.shape != (n, self.stride):
            raise ValueError(f"rows must be [n][{self.stride}]")
        if cells is not None:
            cells = np.ascontiguousarray(cells, dtype=np.uint64)
            if cells.shape != rows.shape:
                raise ValueError("cells must have the shape of rows")
        # without a column permutation to report the kernel runs the column-set search
        self._check(self._L.ise_report_columns(self.h, int(fields is None or "perm_col" in fields)))
        res_all = []
        for off in range(0, n, self.capacity):
            cnt = min(self.capacity, n - off)
            self._check(self._L.ise_submit(self.h, cnt, self.stride, _ptr(rows[off:off + cnt]),
                                           _ptr(None if cells is None else cells[off:off + cnt]),
                                           _ptr(meta[off:off + cnt]), gen_cap))
            r = alloc_results(cnt, gen_cap, fields)
            s = results_struct(r, gen_cap)
            self._check(self._L.ise_collect(self.h, ctypes.byref(s)))
            res_all.append(r)
        if not res_all:
            return alloc_results(0, gen_cap, fields)
        out = {k: np.concatenate([r[k] for r in res_all]) for k in res_all[0]}
        return out

    def canonize_batch(self, rows, meta, max_iters=CANONIZE_MAX_ITERS):
        """Canonical forms of a batch (CanonizeMatrix with k = 0): dict of
        canon [n][stride] rows, status (1 reached, 2 iteration limit, 3 invalid),
        aut (|Aut|, row action) and iters (tests run)."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rows.shape
        out = {"canon": np.zeros_like(rows), "status": np.zeros(n, np.uint8), "aut": np.zeros(n, np.uint64),
               "iters": np.zeros(n, np.uint32)}
        for off in range(0, n, self.capacity):
            cnt = min(self.capacity, n - off)
            part = {k: v[off:off + cnt] for k, v in out.items()}
            tmp = {k: np.ascontiguousarray(v) for k, v in part.items()}
            self._check(self._L.ise_canonize_batch(self.h, cnt, stride, rows[off:off + cnt].ctypes.data,
                                                   meta[off:off + cnt].ctypes.data, tmp["canon"].ctypes.data,
                                                   tmp["status"].ctypes.data, tmp["aut"].ctypes.data,
                                                   tmp["iters"].ctypes.data, max_iters))
            for k in out:
                out[k][off:off + cnt] = tmp[k]
        return out

    def canonize_batch_part(self, rows, meta, part, max_iters=CANONIZE_MAX_ITERS):
        """canonize_batch under an initial column partition `part` (bit c set:
        column c starts a block; CanonizeMatrix with k > 0), through the
        device entry point: dict of canon, status, aut, iters."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rows.shape
        if n > self.capacity:
            raise ValueError("more matrices than the context holds")
        d_rows = DeviceArray(self, rows.shape, np.uint64, rows)
        d_meta = DeviceArray(self, meta.shape, np.uint32, meta)
        d_canon = DeviceArray(self, rows.shape, np.uint64)
        d_status = DeviceArray(self, n, np.uint8)
        d_aut = DeviceArray(self, n, np.uint64)
        d_iters = DeviceArray(self, n, np.uint32)
        self._check(self._L.ise_canonize_batch_device_part(self.h, n, stride, d_rows.ptr, d_meta.ptr, part,
                                                           d_canon.ptr, d_status.ptr, d_aut.ptr, d_iters.ptr,
                                                           max_iters))
        return {"canon": d_canon.to_host(), "status": d_status.to_host(), "aut": d_aut.to_host(),
                "iters": d_iters.to_host()}

    def canonize_matrix(self, mat, len_stab=0):
        """CanonizeMatrix of one uint8 0/1 matrix: (canonical matrix, |Aut|, tests run)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros_like(mat)
        aut, it = ctypes.c_uint64(0), ctypes.c_uint32(0)
        self._check(self._L.ise_canonize_matrix(self.h, mat.ctypes.data, mat.shape[0], mat.shape[1], len_stab,
                                                out.ctypes.data, ctypes.byref(aut), ctypes.byref(it)))
        return out, aut.value, it.value

    def solution_orbits(self, group, sols):
        """Canonical (largest) image of every solution under the group, on the
        GPU, and the index of the first listed solution equal to it
        (CPermutStorage::findSolutionIndex)."""
        group = np.ascontiguousarray(group, dtype=np.uint8)
        sols = np.ascontiguousarray(sols, dtype=np.uint8)
        canon = np.zeros_like(sols)
        rep = np.zeros(len(sols), np.uint32)
        self._check(self._L.ise_solution_orbits(self.h, _ptr(group), len(group), group.shape[1], _ptr(sols),
                                                len(sols), _ptr(canon), _ptr(rep)))
        return canon, rep

    # ---- device-resident batches ------------------------------------------
    def device_array(self, shape, dtype, host=None):
        """Zero-filled device buffer (optionally filled from a host array)."""
        return DeviceArray(self, shape, dtype, host)

    def device_results(self, n, fields, gen_cap=0):
        """Device result arrays for n candidates; returns (dict, ResultsStruct)."""
        host_like = alloc_results(n, gen_cap, fields)
        dev = {k: self.device_array(a.shape, a.dtype) for k, a in host_like.items()}
        s = ResultsStruct()
        for k in ("verdict", "fail_row", "fail_col", "row_to_change", "group_order", "perm_row",
                  "perm_col", "orbits", "n_gens", "gens", "counters", "group_order_hi"):
            setattr(s, k, dev[k].ptr if k in dev else None)
        s.gen_cap = gen_cap if "gens" in dev else 0
        return dev, s

    def check_batch_device(self, n, stride, d_rows, d_meta, res_struct, d_head, d_cells=None):
        self._check(self._L.ise_check_batch_device(self.h, n, stride, d_rows.ptr,
                                                   d_cells.ptr if d_cells is not None else None,
                                                   d_meta.ptr, ctypes.byref(res_struct), d_head.ptr, None))

    def check_batch_device_noalloc(self, n, stride, d_rows, d_meta, res_struct, d_cells=None):
        """ise_check_batch_device_noalloc: the same batch with no allocation and
        no queue buffer (the library's device scratch array holds them)."""
        self._check(self._L.ise_check_batch_device_noalloc(self.h, n, stride, d_rows.ptr,
                                                           d_cells.ptr if d_cells is not None else None,
                                                           d_meta.ptr, ctypes.byref(res_struct), None))

    def time_batch_device(self, n, stride, d_rows, d_meta, res_struct, d_head, repeats, d_cells=None):
        """Launch `repeats` times; per-launch milliseconds from HIP events."""
        ms = np.zeros(repeats, np.float32)
        self._check(self._L.ise_check_batch_device_timed(
            self.h, n, stride, d_rows.ptr, d_cells.ptr if d_cells is not None else None, d_meta.ptr,
            ctypes.byref(res_struct), d_head.ptr, repeats, ms.ctypes.data))
        return ms

    def set_order(self, mode=ORDER_PLAIN, hunt_budget=0, d_order=None, order_dim=0):
        """Search order of the following launches (verdicts never depend on it)."""
        ptr = None if d_order is None else d_order.ptr
        self._check(self._L.ise_set_order(self.h, mode, hunt_budget, ptr, order_dim))

    def use_graphs(self, on=True):
        self._check(self._L.ise_use_graphs(self.h, 1 if on else 0))

    def graph_stats(self):
        g, r = ctypes.c_uint32(0), ctypes.c_uint64(0)
        self._check(self._L.ise_graph_stats(self.h, ctypes.byref(g), ctypes.byref(r)))
        return g.value, r.value

    def graph_updates(self):
        """In-place updates of cached graphs for batches of another size."""
        u = ctypes.c_uint64(0)
        self._check(self._L.ise_graph_updates(self.h, ctypes.byref(u)))
        return u.value

    def autotune(self, n, stride, d_rows, d_meta, res_struct, d_head, max_depth, max_cols, repeats=2,
                 d_cells=None):
        """Sweep lane-group size x waves per workgroup; keeps the fastest.
        Returns (best launch config, table[16/32/64][1/2/4/8 waves] in ms)."""
        table = np.zeros((3, 4), np.float32)
        best = LaunchCfg()
        self._check(self._L.ise_autotune(self.h, n, stride, d_rows.ptr, d_cells.ptr if d_cells else None,
                                         d_meta.ptr, ctypes.byref(res_struct), d_head.ptr, max_depth, max_cols,
                                         repeats, table.ctypes.data, ctypes.byref(best)))
        return dict(group_lanes=best.group_lanes, waves_per_block=best.waves_per_block), table

    def sync(self):
        self._check(self._L.ise_sync(self.h))

    # ---- 1-factorizations -------------------------------------------------
    def kf_set_split(self, first_chunks=DEFAULT_KF_FIRST_CHUNKS, max_ranges=DEFAULT_KF_MAX_RANGES):
        """1-factorization launches: first pass of 64*first_chunks relabellings,
        the rest of undecided candidates in up to max_ranges ranges each
        (first_chunks=0: one wave per candidate to its decision)."""
        self._check(self._L.ise_kf_set_split(self.h, first_chunks, max_ranges))

    def kf_check(self, rounds, meta, two_row_p1f=False, two_row_u1f=False, backjump=False):
        """Canonicity of partial 1-factorizations (host arrays in and out);
        two_row_p1f / two_row_u1f: the two-row canonizers of perfect / uniform
        1-factorizations; backjump: also the k-Sys backjump position
        ("player_index")."""
        fn = (self._L.ise_kf_check_2p1f if two_row_p1f else
              self._L.ise_kf_check_2u1f if two_row_u1f else self._L.ise_kf_check)
        rounds = np.ascontiguousarray(rounds, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rounds.shape
        outs = []
        for off in range(0, n, self.capacity):
            cnt = min(self.capacity, n - off)
            r = kf_alloc(cnt, backjump)
            s = kf_struct(r)
            self._check(fn(self.h, cnt, stride, rounds[off:off + cnt].ctypes.data,
                           meta[off:off + cnt].ctypes.data, ctypes.byref(s)))
            outs.append(r)
        if not outs:
            return kf_alloc(0, backjump)
        return {k: np.concatenate([o[k] for o in outs]) for k in outs[0]}

    def kf_device_results(self, n):
        dev = {"verdict": self.device_array((n,), np.uint8), "witness": self.device_array((n,), np.uint64),
               "n_auts": self.device_array((n,), np.uint64), "n_tried": self.device_array((n,), np.uint64)}
        return dev, kf_struct(dev, device=True)

    def kf_time_batch_device(self, n, stride, d_rounds, d_meta, res_struct, d_head, repeats):
        ms = np.zeros(repeats, np.float32)
        self._check(self._L.ise_kf_check_batch_device_timed(self.h, n, stride, d_rounds.ptr, d_meta.ptr,
                                                            ctypes.byref(res_struct), d_head.ptr, repeats,
                                                            ms.ctypes.data))
        return ms

    def kf_check_3u1f(self, rounds, meta, u1f_cycles=None):
        """cnvCheck3U1F (groups of 2, complete graphs) on host arrays; u1f_cycles:
        the run's cycle types, a list of ascending cycle-length lists (None: one
        cycle through every player)."""
        types = _u1f_types(u1f_cycles)
        rounds = np.ascontiguousarray(rounds, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rounds.shape
        outs = []
        for off in range(0, n, self.capacity):
            cnt = min(self.capacity, n - off)
            r = kf_alloc(cnt)
            s = kf_struct(r)
            self._check(self._L.ise_kf_check_3u1f(self.h, cnt, stride, rounds[off:off + cnt].ctypes.data,
                                                  meta[off:off + cnt].ctypes.data,
                                                  types.ctypes.data if len(types) else None, len(types),
                                                  ctypes.byref(s)))
            outs.append(r)
        if not outs:
            return kf_alloc(0)
        return {k: np.concatenate([o[k] for o in outs]) for k in outs[0]}

    def kf_check_3u1f_device(self, n, stride, d_rounds, d_meta, res_struct, u1f_cycles=None):
        types = _u1f_types(u1f_cycles)
        self._check(self._L.ise_kf_check_3u1f_batch_device(self.h, n, stride, d_rounds.ptr, d_meta.ptr,
                                                           types.ctypes.data if len(types) else None, len(types),
                                                           ctypes.byref(res_struct), None))

    def kf_check_2p1f_device(self, n, stride, d_rounds, d_meta, res_struct):
        self._check(self._L.ise_kf_check_2p1f_batch_device(self.h, n, stride, d_rounds.ptr, d_meta.ptr,
                                                           ctypes.byref(res_struct), None))

    def kf_check_2u1f_device(self, n, stride, d_rounds, d_meta, res_struct):
        self._check(self._L.ise_kf_check_2u1f_batch_device(self.h, n, stride, d_rounds.ptr, d_meta.ptr,
                                                           ctypes.byref(res_struct), None))

    def kf_time_2p1f_device(self, n, stride, d_rounds, d_meta, res_struct, repeats, uniform=False):
        """Timed launches of the two-row canonizer (uniform=True: the U1F one)."""
        ms = np.zeros(repeats, np.float32)
        fn = (self._L.ise_kf_check_2u1f_batch_device_timed if uniform
              else self._L.ise_kf_check_2p1f_batch_device_timed)
        self._check(fn(self.h, n, stride, d_rounds.ptr, d_meta.ptr, ctypes.byref(res_struct), repeats,
                       ms.ctypes.data))
        return ms

    def last_kernel_ms(self):
        return float(self._L.ise_last_kernel_ms(self.h))

    def enumerate_bibd(self, v, k, lam, max_designs=1 << 16, pruned=False):
        """Full enumeration; returns (designs [m][v] uint64, auts [m], stats, level_counts).
        pruned: solutions pruned by each parent's automorphism group
        (ise_enumerate_bibd_pruned)."""
        des = np.zeros((max_designs, v), np.uint64)
        auts = np.zeros(max_designs, np.uint64)
        stats = np.zeros(5, np.uint64)
        levels = np.zeros(v + 1, np.uint64)
        fn = self._L.ise_enumerate_bibd_pruned if pruned else self._L.ise_enumerate_bibd
        n = fn(self.h, v, k, lam, _ptr(des), _ptr(auts), max_designs, _ptr(stats), _ptr(levels))
        if n < 0:
            self._check(int(n))
        m = min(n, max_designs)
        return des[:m], auts[:m], dict(candidates=int(stats[0]), canonical_partial=int(stats[1]),
                                       fully_constructed=int(stats[2]), designs=int(stats[3]),
                                       kernel_ns=int(stats[4])), levels

    def enumerator_candidates(self, v, k, lam, dmin, dmax, stride=None, cap=1 << 20):
        """Every candidate the enumerator tests at depths dmin..dmax."""
        stride = stride or self.stride
        rows = np.zeros((cap, stride), np.uint64)
        meta = np.zeros(cap, np.uint32)
        nout = ctypes.c_long(0)
        rc = self._L.ise_enumerate_bibd_candidates(self.h, v, k, lam, dmin, dmax, stride, _ptr(rows),
                                                   _ptr(meta), cap, ctypes.byref(nout))
        if rc < 0:
            self._check(int(rc))
        return rows[:nout.value], meta[:nout.value]


    def enumerate_bibd_device(self, v, k, lam, max_designs=1 << 16):
        """Full enumeration with row extension, canonicity and compaction all on
        the device; same returns as enumerate_bibd."""
        des = np.zeros((max_designs, v), np.uint64)
        auts = np.zeros(max_designs, np.uint64)
        stats = np.zeros(5, np.uint64)
        levels = np.zeros(v + 1, np.uint64)
        n = self._L.ise_enumerate_bibd_device(self.h, v, k, lam, _ptr(des), _ptr(auts), max_designs,
                                              _ptr(stats), _ptr(levels))
        if n < 0:
            self._check(int(n))
        m = min(n, max_designs)
        return des[:m], auts[:m], dict(candidates=int(stats[0]), canonical_partial=int(stats[1]),
                                       fully_constructed=int(stats[2]), designs=int(stats[3])), levels

    def solve_rows_device(self, v, k, lam, parents, depth, cap=1 << 20):
        """Children of host parents [n][stride] by the GPU row solver:
        (children [m][stride], solutions per parent)."""
        parents = np.ascontiguousarray(parents, dtype=np.uint64)
        n, stride = parents.shape
        d_par = self.device_array(parents.shape, np.uint64, parents)
        d_cnt = self.device_array((max(n, 1),), np.uint32)
        d_off = self.device_array((max(n, 1),), np.uint32)
        d_kids = self.device_array((cap, stride), np.uint64)
        total = ctypes.c_uint64(0)
        self._check(self._L.ise_solve_rows_device(self.h, v, k, lam, n, stride, d_par.ptr, depth, d_cnt.ptr,
                                                  d_off.ptr, d_kids.ptr, cap, ctypes.byref(total)))
        kids = d_kids.to_host()[:total.value]
        return kids, d_cnt.to_host()[:n]

    def enumerate_sampled(self, v, k, lam, beam, seed, dmin, dmax, stride=None, cap=1 << 21):
        """Beam-sampled level enumeration (ise_enumerate_bibd_sampled):
        (cand rows, cand meta, candidates tested per depth)."""
        stride = stride or self.stride
        rows = np.zeros((cap, stride), np.uint64)
        meta = np.zeros(cap, np.uint32)
        lc = np.zeros(v + 1, np.uint64)
        nout = ctypes.c_long(0)
        rc = self._L.ise_enumerate_bibd_sampled(self.h, v, k, lam, beam, seed, dmin, dmax, stride, _ptr(rows),
                                                _ptr(meta), cap, ctypes.byref(nout), _ptr(lc))
        if rc < 0:
            self._check(int(rc))
        return rows[:nout.value], meta[:nout.value], lc


MLP_FEATURES = 16
MLP_PARAMS = 20752


class RowOrderModel:
    """The learned row-order scorer (include/ise.h, ise_mlp_*): a
    16 -> 128 -> 128 -> 1 ReLU network with bf16 weights and FP32
    accumulation on MFMA, trained on the GPU from plain-search witness logs."""

    def __init__(self, ctx, max_rows=1 << 16, seed=1):
        self.ctx = ctx
        self._L = ctx._L
        h = ctypes.c_void_p()
        ctx._check(self._L.ise_mlp_create(ctx.h, ctypes.byref(h), max_rows, seed))
        self.h = h.value
        self.max_rows = (max_rows + 15) // 16 * 16

    def close(self):
        if self.h:
            self._L.ise_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self.ctx._check(self._L.ise_mlp_get_params(self.h, p.ctypes.data))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        self.ctx._check(self._L.ise_mlp_set_params(self.h, p.ctypes.data))

    def features(self, n, stride, d_rows, d_meta, od, d_X):
        self.ctx._check(self._L.ise_mlp_features(self.ctx.h, n, stride, d_rows.ptr, d_meta.ptr, od, d_X.ptr))

    def forward(self, d_X, M, d_score):
        self.ctx._check(self._L.ise_mlp_forward(self.h, d_X.ptr, M, d_score.ptr))

    def gradient(self, d_X, d_y, d_w, M):
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float(0)
        self.ctx._check(self._L.ise_mlp_gradient(self.h, d_X.ptr, d_y.ptr, d_w.ptr, M, g.ctypes.data,
                                                 ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, d_X, d_y, d_w, M, lr):
        loss = ctypes.c_float(0)
        self.ctx._check(self._L.ise_mlp_train_step(self.h, d_X.ptr, d_y.ptr, d_w.ptr, M, lr, ctypes.byref(loss)))
        return loss.value

    def fit(self, d_X, d_y, d_w, M_total, batch, steps, lr):
        curve = np.zeros(steps, np.float32)
        self.ctx._check(self._L.ise_mlp_fit(self.h, d_X.ptr, d_y.ptr, d_w.ptr, M_total, batch, steps, lr,
                                            curve.ctypes.data))
        return curve

    def order(self, n, stride, d_rows, d_meta, od, d_order, d_score=None):
        self.ctx._check(self._L.ise_mlp_order(self.h, n, stride, d_rows.ptr, d_meta.ptr, od, d_order.ptr,
                                              None if d_score is None else d_score.ptr))

    def set_gate(self, gate_perms=DEFAULT_LEARNED_GATE):
        """Order only the candidates the plain search leaves open after
        gate_perms permutations (0: order every candidate)."""
        self.ctx._check(self._L.ise_mlp_set_gate(self.h, gate_perms))

    def use_lead(self, on=True):
        """The context's launches (submit/collect, device batches, the
        enumerators) run the lead order with this network, or, with on=False,
        the plain order again."""
        self.ctx._check(self._L.ise_mlp_use_lead(self.h, 1 if on else 0))

    def lead_features(self, n, stride, d_rows, d_meta, d_X):
        """Features of every row as the row of position 0: d_X [n][16][16] bf16 bits."""
        self.ctx._check(self._L.ise_mlp_lead_features(self.ctx.h, n, stride, d_rows.ptr, d_meta.ptr, d_X.ptr))

    def check_lead(self, n, stride, d_rows, d_meta, res_struct, d_head, d_lead_in=None, d_lead_out=None,
                   d_score_out=None):
        """Learned-order (lead order) test of a device batch, one launch."""
        self.ctx._check(self._L.ise_mlp_check_lead_device(
            self.h, n, stride, d_rows.ptr, d_meta.ptr, ctypes.byref(res_struct), d_head.ptr,
            None if d_lead_in is None else d_lead_in.ptr, None if d_lead_out is None else d_lead_out.ptr,
            None if d_score_out is None else d_score_out.ptr))

    def time_lead(self, n, stride, d_rows, d_meta, res_struct, d_head, repeats):
        """Milliseconds of `repeats` lead-order launches (HIP events on the context's stream)."""
        ms = np.zeros(repeats, np.float32)
        self.ctx._check(self._L.ise_mlp_check_lead_device_timed(
            self.h, n, stride, d_rows.ptr, d_meta.ptr, ctypes.byref(res_struct), d_head.ptr, repeats,
            ms.ctypes.data))
        return ms

    def time_check(self, n, stride, d_rows, d_meta, od, d_order, hunt_budget, res_struct, d_head, repeats):
        """Per repeat: gated plain pass + ordering + learned-order check;
        (total ms, ms up to the end of the ordering)."""
        ms = np.zeros(repeats, np.float32)
        oms = np.zeros(repeats, np.float32)
        self.ctx._check(self._L.ise_mlp_check_batch_device_timed(
            self.h, n, stride, d_rows.ptr, d_meta.ptr, od, d_order.ptr, hunt_budget, ctypes.byref(res_struct),
            d_head.ptr, repeats, ms.ctypes.data, oms.ctypes.data))
        return ms, oms


def group_closure(gens, cap=1 << 20):
    """All elements of the group generated by gens [n][len] (uint8
    permutations), identity first (CPermutStorage::constructGroup)."""
    gens = np.ascontiguousarray(gens, dtype=np.uint8)
    n, length = gens.shape
    out = np.zeros((cap, length), np.uint8)
    order = ctypes.c_uint64(0)
    rc = lib().ise_group_closure(_ptr(gens) if n else None, n, length, _ptr(out), cap, ctypes.byref(order))
    if rc != OK:
        raise ISEError(rc, f"ise_group_closure (order >= {order.value})")
    return out[:order.value]


def adjust_generators(gens, idx):
    """Generators restricted to the column orbits idx (CPermutStorage::adjustGenerators)."""
    gens = np.ascontiguousarray(gens, dtype=np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    n, length = gens.shape
    out = np.zeros((n, len(idx)), np.uint8)
    rc = lib().ise_adjust_generators(_ptr(gens) if n else None, n, length, _ptr(idx), len(idx),
                                     _ptr(out) if n else None)
    if rc != OK:
        raise ISEError(rc, "ise_adjust_generators")
    return out


def _format(call):
    need = ctypes.c_uint64(0)
    call(None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value + 1)
    rc = call(buf, need.value + 1, ctypes.byref(need))
    if rc != OK:
        raise ISEError(rc, "format")
    return buf.value.decode()


def format_matrix(rows, ncols, number=0, sub=0, canon=1, aut=1):
    """Text of one matrix as the reference prints it (packed rows)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    return _format(lambda b, c, n: lib().ise_format_matrix(_ptr(rows), len(rows), ncols, number, sub, canon, aut,
                                                           b, c, n))


def format_aut_table(v, k, lam, designs, auts, orbits):
    """The |Aut(D)| summary table of an enumeration (reference layout)."""
    _, b, r, _, _ = design_params(v, k, lam)
    designs = np.ascontiguousarray(designs, dtype=np.uint64)
    auts = np.ascontiguousarray(auts, dtype=np.uint64)
    orbits = np.ascontiguousarray(orbits, dtype=np.uint8)
    n = len(auts)
    return _format(lambda bf, c, ln: lib().ise_format_aut_table(v, b, r, k, lam, n, _ptr(designs), _ptr(auts),
                                                                _ptr(orbits), bf, c, ln))


def witness_labels(meta, verdict, fail_row, perm_row, od):
    """Training labels from a plain-search witness log (ise_mlp_witness_labels):
    y, w of shape [n][od][od]."""
    meta = np.ascontiguousarray(meta, dtype=np.uint32)
    verdict = np.ascontiguousarray(verdict, dtype=np.uint8)
    fail_row = np.ascontiguousarray(fail_row, dtype=np.uint8)
    perm_row = np.ascontiguousarray(perm_row, dtype=np.uint8)
    n = len(meta)
    if perm_row.shape != (n, MAX_ROWS):
        raise ValueError("perm_row must be [n][32]")
    y = np.zeros((n, od, od), np.float32)
    w = np.zeros((n, od, od), np.float32)
    rc = lib().ise_mlp_witness_labels(n, od, _ptr(meta), _ptr(verdict), _ptr(fail_row), _ptr(perm_row),
                                      _ptr(y), _ptr(w))
    if rc != OK:
        raise ISEError(rc, "ise_mlp_witness_labels")
    return y, w


class DeviceArray:
    """A device buffer of a context (freed with the object)."""

    def __init__(self, ctx, shape, dtype, host=None):
        self.ctx = ctx
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        ctx._check(ctx._L.ise_device_alloc(ctx.h, ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.ctx._check(self.ctx._L.ise_copy_to_device(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes))

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx._L.ise_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx._L.ise_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


RANDOM_MAX_ATTEMPTS = 4096
KF_MAX_U1F_TYPES = 4
KF_MAX_U1F_CYCLES = 4


def _u1f_types(u1f_cycles):
    """The run's cycle types as [n_types][4] uint8 (ascending lengths, zero after the last)."""
    types = np.zeros((len(u1f_cycles or ()), KF_MAX_U1F_CYCLES), np.uint8)
    for k, lens in enumerate(u1f_cycles or ()):
        types[k, :len(lens)] = sorted(lens)
    return types


def random_extensions(v, k, lam, n, dmin, dmax, seed, stride=32, attempts=1, ctx=None):
    """Seeded random valid partial matrices.  attempts > 1: a walk that falls
    short of its depth starts again (rejection sampling of the depth).  With a
    Context the batch is generated on its GPU (same candidates as the host)."""
    rows = np.zeros((n, stride), np.uint64)
    meta = np.zeros(n, np.uint32)
    if attempts == 1 and ctx is None:
        rc = lib().ise_random_extensions(v, k, lam, n, dmin, dmax, seed, stride, _ptr(rows), _ptr(meta))
    else:
        rc = lib().ise_random_extensions_ex(ctx.h if ctx is not None else None, v, k, lam, n, dmin, dmax, seed,
                                            attempts, stride, _ptr(rows), _ptr(meta))
    if rc != OK:
        raise ISEError(rc, "ise_random_extensions")
    return rows, meta


def improved_solution(rows, depth, ncols, perm_row, perm_col, fail_row, cells=None):
    """reconstructSolution's improved solution from a witness on the last row
    (host helper of the library); None when fail_row is not the last row."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    perm_row = np.ascontiguousarray(perm_row, dtype=np.uint8)
    perm_col = np.ascontiguousarray(perm_col, dtype=np.uint8)
    cptr = None
    if cells is not None:
        cells = np.ascontiguousarray(cells, dtype=np.uint64)
        cptr = cells.ctypes.data
    out = np.zeros(64, np.uint8)
    n = ctypes.c_uint32(0)
    rc = lib().ise_improved_solution(rows.ctypes.data, depth, ncols, cptr, perm_row.ctypes.data,
                                     perm_col.ctypes.data, int(fail_row), out.ctypes.data, ctypes.byref(n))
    if rc == ERR_ARG:
        return None
    if rc != OK:
        raise ISEError(rc, "ise_improved_solution")
    return out[:n.value].copy()


def kf_format_matrix(mat, number, aut):
    """A 1-factorization (rounds x players) as the k-Sys result files print it."""
    rows, meta = kf_pack([mat])
    nr, n = int(meta[0]) & 0xff, (int(meta[0]) >> 8) & 0xff
    need = ctypes.c_uint64(0)
    lib().ise_kf_format_matrix(rows.ctypes.data, nr, n, number, aut, None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value + 1)
    rc = lib().ise_kf_format_matrix(rows.ctypes.data, nr, n, number, aut, buf, need.value + 1, ctypes.byref(need))
    if rc != OK:
        raise ISEError(rc, "ise_kf_format_matrix")
    return buf.value.decode()


def kf_random(players, rmin, rmax, count, seed, stride=16):
    """Seeded random partial 1-factorizations of K_players in normal form."""
    rows = np.zeros((count, stride), np.uint64)
    meta = np.zeros(count, np.uint32)
    rc = lib().ise_kf_random(players, rmin, rmax, count, seed, stride, rows.ctypes.data, meta.ctypes.data)
    if rc != OK:
        raise ISEError(rc, "ise_kf_random")
    return rows, meta


def kf_relabel_p1f(src_rows, src_meta, count, rmin, seed, stride=16, uniform=False):
    """Seeded relabelled prefixes of perfect (uniform=True: uniform)
    1-factorizations (host code)."""
    src_rows = np.ascontiguousarray(src_rows, dtype=np.uint64)
    src_meta = np.ascontiguousarray(src_meta, dtype=np.uint32)
    rows = np.zeros((count, stride), np.uint64)
    meta = np.zeros(count, np.uint32)
    fn = lib().ise_kf_relabel_u1f if uniform else lib().ise_kf_relabel_p1f
    rc = fn(src_rows.ctypes.data, src_meta.ctypes.data, len(src_meta), src_rows.shape[1],
            count, rmin, seed, stride, rows.ctypes.data, meta.ctypes.data)
    if rc != OK:
        raise ISEError(rc, "ise_kf_relabel_u1f" if uniform else "ise_kf_relabel_p1f")
    return rows, meta


CD_MAX_WALK_DAYS = 10


class KSysCanonizer:
    """The k-Sys binary canonizer: createCanonizer(v, k) / runCanonizer /
    releaseCanonizer (k-Factorization/Utils/Include/CDTools.h:9-11,
    Source/CDTools.cpp:5-51) on the GPU.  run(matrix, k, day) takes `day`
    days (all when day <= 0) of v player numbers, the groups of k of each day
    in order, and returns the v x b canonical binary matrix (uint8), b =
    days * v / k; with every day present the day order is canonized too."""

    def __init__(self, v, k, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.ise_cd_create_canonizer(ctypes.byref(h), device, v, k)
        if rc != OK:
            raise ISEError(rc, "createCanonizer: invalid v, k or no gfx950 device")
        self.h, self.v, self.k = h, v, k
        self.days_max = (v - 1) // (k - 1)

    def close(self):
        if self.h:
            self._L.ise_cd_release_canonizer(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise ISEError(rc, self._L.ise_cd_error(self.h).decode(errors="replace"))

    def _days(self, day):
        return self.days_max if day <= 0 else day

    def run(self, matrix, k=None, day=0):
        """runCanonizer(matrix, k, day): the canonical v x b player rows."""
        matrix = np.ascontiguousarray(matrix, dtype=np.uint8)
        out = ctypes.c_void_p()
        nc = ctypes.c_uint32(0)
        self._check(self._L.ise_cd_run_canonizer(self.h, matrix.ctypes.data, self.k if k is None else k, day,
                                                 ctypes.byref(out), ctypes.byref(nc)))
        b = nc.value
        buf = (ctypes.c_uint8 * (self.v * b)).from_address(out.value)
        return np.frombuffer(buf, np.uint8).reshape(self.v, b).copy()

    def group_order(self):
        """(day-permutation group order, |Aut| with the days fixed) of the last
        run(): their product is |Aut(M)| of a full k-system."""
        d, f = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(self._L.ise_cd_group_order(self.h, ctypes.byref(d), ctypes.byref(f)))
        return d.value, f.value

    def run_batch(self, matrices, day=0, orders=False):
        """runCanonizer on every k-system of matrices [n][days][v]: [n][v][b]
        (with orders=True also the day-permutation group orders and the
        fixed-day |Aut| per k-system)."""
        matrices = np.ascontiguousarray(matrices, dtype=np.uint8)
        n = matrices.shape[0]
        b = self._days(day) * (self.v // self.k)
        out = np.zeros((n, self.v, b), np.uint8)
        dord = np.zeros(n, np.uint64)
        ford = np.zeros(n, np.uint64)
        self._check(self._L.ise_cd_run_canonizer_batch(self.h, n, matrices.ctypes.data, day, out.ctypes.data,
                                                       dord.ctypes.data, ford.ctypes.data))
        return (out, dord, ford) if orders else out

    def run_device(self, n, d_matrices, day, d_canon, d_day_orders=None, d_fixed_orders=None):
        """Device-resident batch: d_matrices [n][days][v] bytes, d_canon [n][v]
        packed rows, d_day_orders / d_fixed_orders [n] words or None
        (pointers of device memory)."""
        self._check(self._L.ise_cd_run_canonizer_device(self.h, n, d_matrices, day, d_canon, d_day_orders,
                                                        d_fixed_orders))


class CanonicityChecker:
    """Reference-shaped single-matrix checker on the GPU."""

    def __init__(self, nRow, nCol, rank=2, ctx=None):
        if rank != 2:
            raise ValueError("only 0/1 incidence matrices (rank 2)")
        self.nRow, self.nCol = nRow, nCol
        self._ctx = ctx or Context(capacity=64, stride=max(nRow, 1))
        self._res = None
        self.stabiliserLengthExt = 0

    def TestCanonicity(self, nRowMax, matrix, outInfo=0, gen_cap=0):
        mat = np.ascontiguousarray(np.asarray(matrix, dtype=np.uint8)[:nRowMax, :self.nCol])
        rows = np.zeros((1, self._ctx.stride), np.uint64)
        rows[0, :nRowMax] = pack_rows(mat)
        meta = np.array([pack_meta(nRowMax, self.nCol, self.stabiliserLengthExt, outInfo)], np.uint32)
        self._res = self._ctx.check_batch(rows, meta, gen_cap=gen_cap)
        self._nrow = nRowMax
        return bool(self._res["verdict"][0] == VERDICT_CANONICAL)

    def permRow(self):
        return self._res["perm_row"][0, :self._nrow].copy()

    def permCol(self):
        return self._res["perm_col"][0, :self.nCol].copy()

    def orbits(self):
        return self._res["orbits"][0, :self._nrow].copy()

    def groupOrder(self):
        """|Aut| as a Python int (the reference's 128-bit UInt)."""
        return int(self._res["group_order"][0]) | (int(self._res["group_order_hi"][0]) << 64)

    def rowToChange(self):
        return int(self._res["row_to_change"][0])


class GPUCanonChecker:
    """assignChecker / LaunchCanonicityTesting batching like CGPU_CanonChecker."""

    def __init__(self, nRow, nCol, batch=128, ctx=None):
        self.nRow, self.nCol = nRow, nCol
        self._ctx = ctx or Context(capacity=batch, stride=nRow)
        self._L = lib()

    def assignChecker(self, matrix, nRowMax=None, outInfo=SAVE_ROW_PERMUTATIONS | SAVE_ROW_TO_CHANGE):
        mat = np.asarray(matrix, dtype=np.uint8)
        d = nRowMax or mat.shape[0]
        words = np.ascontiguousarray(pack_rows(mat[:d, :self.nCol]))
        self._ctx._check(self._L.ise_assign(self._ctx.h, _ptr(words), d, self.nCol, 0, outInfo))

    def numAssigned(self):
        n = ctypes.c_uint32(0)
        self._ctx._check(self._L.ise_num_assigned(self._ctx.h, ctypes.byref(n)))
        return n.value

    def LaunchCanonicityTesting(self):
        n = self.numAssigned()
        self._ctx._check(self._L.ise_submit(self._ctx.h, 0, self._ctx.stride, None, None, None, 0))
        r = alloc_results(n)
        s = results_struct(r)
        self._ctx._check(self._L.ise_collect(self._ctx.h, ctypes.byref(s)))
        return r


__all__ = [
    "QUEUE_WORDS", "ORDER_PLAIN", "ORDER_HUNT", "ORDER_LEARNED", "kf_random", "kf_relabel_p1f", "kf_pack", "kf_unpack", "kf_alloc", "KF_FIELDS", "Context", "DeviceArray", "ResultsStruct", "CanonicityChecker", "GPUCanonChecker", "ISEError", "random_extensions", "lib",
    "design_params", "alloc_results", "pack_meta", "pack_rows", "unpack_rows", "MAX_ROWS",
    "MAX_COLS", "SAVE_ROW_PERMUTATIONS", "SAVE_COL_PERMUTATIONS", "SAVE_ROW_TO_CHANGE",
    "VERDICT_CANONICAL", "VERDICT_NONCANONICAL", "VERDICT_LIMIT", "VERDICT_INVALID",
]
