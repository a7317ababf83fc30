"""ctypes bindings of the CPU oracle (oracle/liboracle_ise.so).

The oracle is test infrastructure: only tests, the smoke check and the CPU
baseline of the benchmark load it.  It is built from oracle/ with make when
the shared library is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle_ise.so")

_P = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        U32 = ctypes.c_uint32
        L.orc_check_batch.argtypes = [U32, U32, _P, _P, _P, U32, _P]
        L.orc_test_canonicity.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, U32]
        L.orc_cell_masks.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P]
        L.orc_walk_step.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        L.orc_improved_solution.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        L.orc_canonize.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_enumerate_bibd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_long,
                                         _P, _P, _P, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_enumerate_bibd.restype = ctypes.c_long
        L.orc_enumerate_sampled.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                            ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P,
                                            ctypes.c_long, _P]
        L.orc_enumerate_sampled.restype = ctypes.c_long
        L.orc_row_solutions.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P,
                                        ctypes.c_long]
        L.orc_row_solutions.restype = ctypes.c_long
        L.orc_bruteforce_canon.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P]
        L.orc_check_batch_hunt.argtypes = [U32, U32, _P, _P, _P, _P, U32, U32, U32, _P]
        L.orc_lead_features.argtypes = [U32, U32, _P, _P, _P]
        L.orc_check_batch_lead.argtypes = [U32, U32, _P, _P, _P, _P, U32, U32, _P, _P]
        L.orc_hunt.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, U32, _P, _P, _P]
        L.orc_kf_check.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_kf_check_2p1f.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_kf_check_2u1f.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_kf_check_3u1f.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P, _P, _P]
        L.orc_kf_normalize.argtypes = [_P, ctypes.c_int, ctypes.c_int]
        L.orc_kf_bruteforce.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P]
        L.orc_kf_player_index.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        L.orc_cd_run_canonizer.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
        L.orc_cd_run_canonizer_ex.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
        L.orc_cd_canonize_orders.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_long, _P]
        L.orc_cd_day_starts.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_cd_day_starts.restype = ctypes.c_uint64
        L.orc_canonize_part.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _P, _P, _P]
        L.orc_kf_quickperm.argtypes = [ctypes.c_int, _P]
        L.orc_kf_quickperm.restype = ctypes.c_long
        L.orc_mlp_features.argtypes = [U32, U32, _P, _P, U32, _P]
        L.orc_mlp_forward.argtypes = [_P, _P, U32, ctypes.c_int, _P]
        L.orc_mlp_gradient.argtypes = [_P, _P, _P, _P, U32, ctypes.c_int, _P, _P]
        _lib = Oracle(L)
    return _lib


class Oracle:
    def __init__(self, L):
        self.L = L

    def check_batch(self, rows, meta, cells=None, gen_cap=0):
        import ise_amd._abi as abi
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rows.shape
        res = abi.alloc_results(n, gen_cap)
        s = abi.results_struct(res, gen_cap)
        cells_p = None
        if cells is not None:
            cells = np.ascontiguousarray(cells, dtype=np.uint64)
            cells_p = cells.ctypes.data
        rc = self.L.orc_check_batch(n, stride, rows.ctypes.data, cells_p, meta.ctypes.data, gen_cap,
                                    ctypes.addressof(s))
        if rc != 0:
            raise ValueError(f"oracle rejected the batch ({rc})")
        return res

    def check_batch_hunt(self, rows, meta, order=None, order_dim=0, budget=64, gen_cap=0):
        """Witness hunt first (increasing row order, or order [n][dim][dim]), then plain."""
        import ise_amd._abi as abi
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rows.shape
        res = abi.alloc_results(n, gen_cap)
        s = abi.results_struct(res, gen_cap)
        optr = None
        if order is not None:
            order = np.ascontiguousarray(order, dtype=np.uint8)
            optr = order.ctypes.data
        rc = self.L.orc_check_batch_hunt(n, stride, rows.ctypes.data, None, meta.ctypes.data, optr, order_dim,
                                         budget, gen_cap, ctypes.addressof(s))
        if rc != 0:
            raise ValueError("oracle rejected the batch")
        return res

    def lead_features(self, rows, meta):
        """Features of every row as the row of position 0: [n][16][16] float32."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        X = np.zeros((len(meta), 16, 16), np.float32)
        self.L.orc_lead_features(len(meta), rows.shape[1], rows.ctypes.data, meta.ctypes.data, X.ctypes.data)
        return X

    def check_batch_lead(self, rows, meta, lead, gate_perms, cells=None, gen_cap=0):
        """Learned-order test with the rows of position 0 in the order lead[n][16]:
        (results, mode[n]) -- see orc_check_batch_lead."""
        import ise_amd._abi as abi
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        lead = np.ascontiguousarray(lead, dtype=np.uint8)
        n, stride = rows.shape
        res = abi.alloc_results(n, gen_cap)
        s = abi.results_struct(res, gen_cap)
        mode = np.zeros(n, np.uint8)
        cells_p = None
        if cells is not None:
            cells = np.ascontiguousarray(cells, dtype=np.uint64)
            cells_p = cells.ctypes.data
        rc = self.L.orc_check_batch_lead(n, stride, rows.ctypes.data, cells_p, meta.ctypes.data, lead.ctypes.data,
                                         gate_perms, gen_cap, ctypes.addressof(s), mode.ctypes.data)
        if rc != 0:
            raise ValueError(f"oracle rejected the batch ({rc})")
        return res, mode

    def walk_step(self, perm, orbits, at, fixed, state):
        """One step of the oracle's row-permutation walk; perm (uint8 [n]) and
        state (uint16 [2]) are updated in place.  Returns the first changed
        position or 255."""
        return self.L.orc_walk_step(perm.ctypes.data, orbits.ctypes.data, len(perm), at, fixed,
                                    state.ctypes.data)

    def improved_solution(self, mat, len_stab=0):
        """reconstructSolution's vector when the test fails on the last row, else None."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros(64, np.uint8)
        n = self.L.orc_improved_solution(mat.ctypes.data, mat.shape[0], mat.shape[1], len_stab, out.ctypes.data)
        return None if n < 0 else out[:n].copy()

    def canonize(self, mat, len_stab=0, max_iters=1 << 16):
        """CanonizeMatrix (k = 0): (canonical matrix, |Aut|, tests run)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros_like(mat)
        aut = np.zeros(1, np.uint64)
        it = np.zeros(1, np.uint32)
        rc = self.L.orc_canonize(mat.ctypes.data, mat.shape[0], mat.shape[1], len_stab, max_iters,
                                 out.ctypes.data, aut.ctypes.data, it.ctypes.data)
        if rc != 1:
            raise ValueError(f"oracle canonize failed ({rc})")
        return out, int(aut[0]), int(it[0])

    def cell_masks(self, rows, nrows, ncols):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        out = np.zeros(nrows, np.uint64)
        self.L.orc_cell_masks(rows.ctypes.data, nrows, ncols, out.ctypes.data)
        return out

    def enumerate(self, v, k, lam, max_designs=1 << 16, capture=None):
        """DFS enumeration: (designs [m][v], auts [m], stats[5], (cand rows, cand meta))."""
        des = np.zeros((max_designs, v), np.uint64)
        auts = np.zeros(max_designs, np.uint64)
        st = np.zeros(5, np.uint64)
        crow = cmeta = None
        cap, stride, dmin, dmax = 0, 0, 0, 0
        if capture:
            cap, stride, dmin, dmax = capture
            crow = np.zeros((cap, stride), np.uint64)
            cmeta = np.zeros(cap, np.uint32)
        n = self.L.orc_enumerate_bibd(v, k, lam, des.ctypes.data, auts.ctypes.data, max_designs,
                                      st.ctypes.data, None if crow is None else crow.ctypes.data,
                                      None if cmeta is None else cmeta.ctypes.data, cap, stride,
                                      dmin, dmax)
        if n < 0:
            raise ValueError(f"oracle enumerate failed ({n})")
        m = min(n, max_designs)
        cands = None
        if capture:
            c = int(st[4])
            cands = (crow[:c], cmeta[:c])
        return des[:m], auts[:m], st, cands

    def enumerate_sampled(self, v, k, lam, beam, seed, dmin, dmax, stride=32, cap=1 << 20):
        """Beam-sampled level enumeration: (cand rows, cand meta, level counts)."""
        crow = np.zeros((cap, stride), np.uint64)
        cmeta = np.zeros(cap, np.uint32)
        lc = np.zeros(33, np.uint64)
        n = self.L.orc_enumerate_sampled(v, k, lam, beam, ctypes.c_uint64(seed), dmin, dmax, stride,
                                         crow.ctypes.data, cmeta.ctypes.data, cap, lc.ctypes.data)
        if n < 0:
            raise ValueError(f"oracle sampled enumeration failed ({n})")
        return crow[:n], cmeta[:n], lc[:v + 1]

    # ---- learned row-order scorer ----
    def mlp_features(self, rows, meta, od):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n, stride = rows.shape
        X = np.zeros((n, od, od, 16), np.float32)
        self.L.orc_mlp_features(n, stride, rows.ctypes.data, meta.ctypes.data, od, X.ctypes.data)
        return X

    def mlp_forward(self, params, X, bf16=True):
        params = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        z = np.zeros(len(X), np.float64)
        self.L.orc_mlp_forward(params.ctypes.data, X.ctypes.data, len(X), int(bf16), z.ctypes.data)
        return z

    def mlp_gradient(self, params, X, y, w, bf16=True):
        params = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        y = np.ascontiguousarray(y, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        g = np.zeros(len(params), np.float64)
        loss = np.zeros(1, np.float64)
        rc = self.L.orc_mlp_gradient(params.ctypes.data, X.ctypes.data, y.ctypes.data, w.ctypes.data, len(X),
                                     int(bf16), g.ctypes.data, loss.ctypes.data)
        if rc:
            raise ValueError("zero total weight")
        return g, float(loss[0])

    def row_solutions(self, v, k, lam, rows, depth, cap=1 << 16):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        out = np.zeros(cap, np.uint64)
        n = self.L.orc_row_solutions(v, k, lam, rows.ctypes.data, depth, out.ctypes.data, cap)
        if n < 0:
            raise ValueError(f"oracle row solutions failed ({n})")
        return out[:min(n, cap)]

    def test_canonicity(self, mat, len_stab=0, out_info=0, gen_cap=0):
        import ise_amd._abi as abi
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        res = abi.alloc_results(1, gen_cap)
        s = abi.results_struct(res, gen_cap)
        self.L.orc_test_canonicity(mat.ctypes.data, mat.shape[0], mat.shape[1], len_stab, out_info,
                                   ctypes.addressof(s), gen_cap)
        return res

    def kf_check(self, mat):
        """(verdict, witness index, automorphisms counted, relabellings tried)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        w, a, t = (np.zeros(1, np.uint64) for _ in range(3))
        v = self.L.orc_kf_check(mat.ctypes.data, mat.shape[1], mat.shape[0], w.ctypes.data, a.ctypes.data,
                                t.ctypes.data)
        if v < 0:
            raise ValueError("oracle rejected the factorization")
        return v, int(w[0]) if v == 0 else (1 << 64) - 1, int(a[0]), int(t[0])

    def kf_check_2p1f(self, mat):
        """Two-row P1F canonizer: (verdict, witness index, automorphisms, tried)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        w, a, t = (np.zeros(1, np.uint64) for _ in range(3))
        v = self.L.orc_kf_check_2p1f(mat.ctypes.data, mat.shape[1], mat.shape[0], w.ctypes.data, a.ctypes.data,
                                     t.ctypes.data)
        if v < 0:
            raise ValueError("oracle rejected the factorization (not a perfect 1-factorization)")
        return v, int(w[0]) if v == 0 else (1 << 64) - 1, int(a[0]), int(t[0])

    def kf_check_2u1f(self, mat):
        """Two-row U1F canonizer: (verdict, witness index, automorphisms, tried)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        w, a, t = (np.zeros(1, np.uint64) for _ in range(3))
        v = self.L.orc_kf_check_2u1f(mat.ctypes.data, mat.shape[1], mat.shape[0], w.ctypes.data, a.ctypes.data,
                                     t.ctypes.data)
        if v < 0:
            raise ValueError("oracle rejected the factorization")
        return v, int(w[0]), int(a[0]), int(t[0])

    def kf_check_3u1f(self, mat, u1f_cycles=None):
        """cnvCheck3U1F for groups of 2: (verdict, witness index, automorphisms, tried)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        types = np.zeros((len(u1f_cycles or ()), 4), np.uint8)
        for k, lens in enumerate(u1f_cycles or ()):
            types[k, :len(lens)] = sorted(lens)
        w, a, t = (np.zeros(1, np.uint64) for _ in range(3))
        v = self.L.orc_kf_check_3u1f(mat.ctypes.data, mat.shape[1], mat.shape[0],
                                     types.ctypes.data if len(types) else None, len(types), w.ctypes.data,
                                     a.ctypes.data, t.ctypes.data)
        if v < 0:
            raise ValueError("oracle rejected the factorization")
        return v, int(w[0]), int(a[0]), int(t[0])

    def kf_player_index(self, mat, mode=0):
        """(backjump position, first smaller row, its source round, set?)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        info = np.zeros(3, np.int32)
        pi = self.L.orc_kf_player_index(mat.ctypes.data, mat.shape[1], mat.shape[0], mode, info.ctypes.data)
        if pi < 0:
            raise ValueError("oracle rejected the factorization")
        return pi, int(info[0]), int(info[1]), bool(info[2])

    def cd_run_canonizer(self, mat, v, k, day=0):
        """runCanonizer: the v x b player rows of the canonical binary matrix."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        days = day if day > 0 else (v - 1) // (k - 1)
        out = np.zeros((v, 64), np.uint8)
        nc = np.zeros(1, np.int32)
        rc = self.L.orc_cd_run_canonizer(mat.ctypes.data, v, k, day, out.ctypes.data, nc.ctypes.data)
        if rc != 0:
            raise ValueError(f"oracle runCanonizer failed ({rc})")
        b = int(nc[0])
        assert b == days * (v // k)
        return out.reshape(-1)[:v * b].reshape(v, b).copy()

    def cd_run_canonizer_orders(self, mat, v, k, day=0):
        """runCanonizer with the group orders: (v x b matrix, day orders that
        reach it, its |Aut| with the days fixed)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros((v, 64), np.uint8)
        nc = np.zeros(1, np.int32)
        dord = np.zeros(1, np.uint64)
        ford = np.zeros(1, np.uint64)
        rc = self.L.orc_cd_run_canonizer_ex(mat.ctypes.data, v, k, day, out.ctypes.data, nc.ctypes.data,
                                            dord.ctypes.data, ford.ctypes.data)
        if rc != 0:
            raise ValueError(f"oracle runCanonizer failed ({rc})")
        b = int(nc[0])
        return out.reshape(-1)[:v * b].reshape(v, b).copy(), int(dord[0]), int(ford[0])

    def cd_canonize_orders(self, mat, v, k, orders):
        """Largest canonical form (v x b) over the given day orders [n][days]."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        orders = np.ascontiguousarray(orders, dtype=np.uint8)
        days = orders.shape[1]
        best = np.zeros((v, days * (v // k)), np.uint8)
        rc = self.L.orc_cd_canonize_orders(mat.ctypes.data, v, k, days, orders.ctypes.data, len(orders),
                                           best.ctypes.data)
        if rc != 0:
            raise ValueError(f"oracle canonize_orders failed ({rc})")
        return best

    def cd_day_starts(self, days, groups):
        return int(self.L.orc_cd_day_starts(days, groups))

    def canonize_part(self, mat, part, max_iters=1 << 16):
        """CanonizeMatrix under an initial column partition (bit c: column c
        starts a block): (canonical matrix, |Aut|, tests run)."""
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros_like(mat)
        aut = np.zeros(1, np.uint64)
        it = np.zeros(1, np.uint32)
        rc = self.L.orc_canonize_part(mat.ctypes.data, mat.shape[0], mat.shape[1], part, max_iters,
                                      out.ctypes.data, aut.ctypes.data, it.ctypes.data)
        if rc != 1:
            raise ValueError(f"oracle canonize failed ({rc})")
        return out, int(aut[0]), int(it[0])

    def kf_bruteforce(self, mat):
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        a = np.zeros(1, np.uint64)
        v = self.L.orc_kf_bruteforce(mat.ctypes.data, mat.shape[1], mat.shape[0], a.ctypes.data)
        return v, int(a[0])

    def kf_normalize(self, mat):
        mat = np.ascontiguousarray(mat, dtype=np.uint8).copy()
        self.L.orc_kf_normalize(mat.ctypes.data, mat.shape[1], mat.shape[0])
        return mat

    def kf_quickperm(self, k):
        import math
        out = np.zeros((math.factorial(k), k), np.uint8)
        self.L.orc_kf_quickperm(k, out.ctypes.data)
        return out

    def bruteforce(self, mat):
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        out = np.zeros_like(mat)
        auts = np.zeros(1, np.uint64)
        rc = self.L.orc_bruteforce_canon(mat.ctypes.data, mat.shape[0], mat.shape[1], out.ctypes.data,
                                         auts.ctypes.data)
        if rc != 0:
            raise ValueError("bruteforce: matrix too large")
        return out, int(auts[0])
