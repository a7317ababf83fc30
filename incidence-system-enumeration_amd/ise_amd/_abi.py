"""ctypes description of the ise_results record and numpy result buffers.

The layout mirrors `ise_results` in include/ise.h: a structure of arrays with
one entry (or one fixed-size row) per candidate.
"""
import ctypes

import numpy as np

MAX_ROWS = 32
MAX_COLS = 64

SAVE_ROW_PERMUTATIONS = 1
SAVE_COL_PERMUTATIONS = 2
SAVE_ROW_TO_CHANGE = 4

VERDICT_NONCANONICAL = 0
VERDICT_CANONICAL = 1
VERDICT_LIMIT = 2
VERDICT_INVALID = 3

_P = ctypes.c_void_p


class ResultsStruct(ctypes.Structure):
    _fields_ = [
        ("verdict", _P), ("fail_row", _P), ("fail_col", _P), ("row_to_change", _P),
        ("group_order", _P), ("perm_row", _P), ("perm_col", _P), ("orbits", _P),
        ("n_gens", _P), ("gens", _P), ("counters", _P), ("gen_cap", ctypes.c_uint32),
        ("group_order_hi", _P),
    ]


RESULT_FIELDS = ("verdict", "fail_row", "fail_col", "row_to_change", "group_order",
                 "perm_row", "perm_col", "orbits", "n_gens", "gens", "counters", "group_order_hi")


def alloc_results(n, gen_cap=0, fields=None):
    """Host result arrays for n candidates (all fields unless `fields` given)."""
    want = set(fields or RESULT_FIELDS)
    want.add("verdict")
    shapes = {
        "verdict": ((n,), np.uint8), "fail_row": ((n,), np.uint8),
        "fail_col": ((n,), np.uint8), "row_to_change": ((n,), np.uint8),
        "group_order": ((n,), np.uint64), "perm_row": ((n, MAX_ROWS), np.uint8),
        "perm_col": ((n, MAX_COLS), np.uint8), "orbits": ((n, MAX_ROWS), np.uint8),
        "n_gens": ((n,), np.uint16), "counters": ((n, 4), np.uint32),
        "group_order_hi": ((n,), np.uint64),
    }
    out = {}
    for name, (shape, dt) in shapes.items():
        if name in want:
            out[name] = np.zeros(shape, dt)
    if gen_cap and "gens" in want:
        out["gens"] = np.zeros((n, gen_cap, MAX_COLS), np.uint8)
    return out


def results_struct(arrays, gen_cap=0):
    s = ResultsStruct()
    for name in RESULT_FIELDS:
        a = arrays.get(name)
        setattr(s, name, a.ctypes.data if a is not None else None)
    s.gen_cap = gen_cap if arrays.get("gens") is not None else 0
    return s


def pack_meta(depth, ncols, len_stab=0, out_info=0):
    return (int(depth) | (int(ncols) << 8) | (int(len_stab) << 16) | (int(out_info) << 24))


def pack_rows(mat):
    """uint8 0/1 matrix (rows x cols) -> packed uint64 row words."""
    mat = np.asarray(mat, dtype=np.uint8)
    weights = (np.uint64(1) << np.arange(mat.shape[1], dtype=np.uint64))
    return (mat.astype(np.uint64) * weights).sum(axis=1).astype(np.uint64)


def unpack_rows(words, ncols):
    words = np.asarray(words, dtype=np.uint64)
    bits = (words[:, None] >> np.arange(ncols, dtype=np.uint64)) & np.uint64(1)
    return bits.astype(np.uint8)
