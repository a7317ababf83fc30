"""Text output in the reference's layout: printed matrices against verbatim
blocks of a reference result file (CPU), and the |Aut(D)| summary tables of
full GPU enumerations against the reference's result files (GPU)."""
import re

import numpy as np
import pytest

import ise_amd
from _helpers import golden


def _pack(lines):
    return np.array([sum(1 << c for c, ch in enumerate(t) if ch == "1") for t in lines], np.uint64)


def test_printed_matrix_blocks_verbatim():
    blocks = golden("format_golden.json")["matrices"]["blocks"]
    assert len(blocks) >= 10
    for block in blocks:
        m = re.match(r"Matrix #\s*(\d+)\.(\d+)\s+\|Aut\(M\)\| =\s*(\d+)", block[0])
        number, sub, aut = map(int, m.groups())
        rows = block[1:]
        text = ise_amd.format_matrix(_pack(rows), len(rows[0]), number, sub, 1, aut)
        assert text.split("\n") == [""] + block + [""]


def test_matrix_rows_only_and_flags():
    rows = _pack(["110", "011"])
    assert ise_amd.format_matrix(rows, 3) == "110\n011\n"
    t = ise_amd.format_matrix(rows, 3, number=7, canon=0, aut=1)
    assert t.splitlines()[1] == "Matrix #   7  *** Non-canonical ***    |Aut(M)| =      1"


def test_aut_table_counts_simple_and_transitive():
    # two copies of the Fano plane's blocks: not simple; its group is transitive
    fano = ["1110000", "1001100", "1000011", "0101010", "0100101", "0011001", "0010110"]
    cols = [[int(r[c]) for r in fano] for c in range(7)]
    rows = np.array([sum(cols[c][i] << c for c in range(7)) for i in range(7)], np.uint64)
    orbits = np.zeros((1, 32), np.uint8)
    t = ise_amd.format_aut_table(7, 3, 1, rows[None], np.array([168], np.uint64), orbits)
    lines = t.splitlines()
    assert lines[0] == "BIBD(  7,   7,  3,  3,  1)"
    assert lines[4].split() == ["168", "1", "1", "1", "1"]
    assert lines[-1].split() == ["Total:", "1", "1", "1", "1"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["7_3_1", "6_3_2", "6_3_4", "6_3_6", "6_3_8", "6_3_10", "6_3_12"])
def test_aut_table_of_gpu_enumeration_matches_reference_file(gpu_ctx, name):
    print(f"case summary table {name}", flush=True)
    want = golden("format_golden.json")["tables"][name]["lines"]
    v, k, lam = map(int, name.split("_"))
    _, b, r, _, _ = ise_amd.design_params(v, k, lam)
    designs, auts, _, _ = gpu_ctx.enumerate_bibd(v, k, lam)
    rows = np.zeros((len(designs), gpu_ctx.stride), np.uint64)
    rows[:, :v] = designs
    meta = np.full(len(designs), v | (b << 8) | (ise_amd.SAVE_ROW_PERMUTATIONS << 24), np.uint32)
    res = gpu_ctx.check_batch(rows, meta)
    assert np.array_equal(res["group_order"], auts)
    text = ise_amd.format_aut_table(v, k, lam, designs, auts, res["orbits"])
    assert text.splitlines() == want


@pytest.mark.parametrize("name", ["7_3_1", "6_3_2", "6_3_4", "6_3_6", "6_3_8", "6_3_10", "6_3_12"])
def test_aut_table_of_oracle_enumeration_matches_reference_file(oracle, name):
    want = golden("format_golden.json")["tables"][name]["lines"]
    v, k, lam = map(int, name.split("_"))
    _, b, r, _, _ = ise_amd.design_params(v, k, lam)
    designs, auts, _, _ = oracle.enumerate(v, k, lam)
    rows = np.zeros((len(designs), 32), np.uint64)
    rows[:, :v] = designs
    meta = np.full(len(designs), v | (b << 8) | (ise_amd.SAVE_ROW_PERMUTATIONS << 24), np.uint32)
    res = oracle.check_batch(rows, meta)
    text = ise_amd.format_aut_table(v, k, lam, designs, auts, res["orbits"])
    assert text.splitlines() == want


def test_kfact_blocks_verbatim():
    """1-factorizations printed exactly as the k-Sys result files print them:
    the first 60 of the 396 K_10 ones, the perfect ones of K_10 and K_14 and
    the uniform ones of K_8, K_10 and K_12 -- header with |Aut(M)| and the
    cycle-type counts, and the quoted rows."""
    n_blocks = 0
    for f in golden("format_golden.json")["kfact"]:
        for blk in f["blocks"]:
            head = re.match(r"\s*(\d+): \|Aut\(M\)\| =\s+(\d+)", blk[0])
            number, aut = int(head.group(1)), int(head.group(2))
            mat = [[int(x) for x in line.strip().strip('"').split()] for line in blk[1:]]
            text = ise_amd.kf_format_matrix(np.array(mat, np.uint8), number, aut)
            assert text == "\n\n" + "\n".join(blk) + "\n\n", (f["source"], number)
            n_blocks += 1
    assert n_blocks >= 80
