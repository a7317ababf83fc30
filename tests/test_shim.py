"""The reference-side C++ binding (incidence-system-enumeration_amd/shim/
ise_canonicity_checker.hpp) compiled against stand-ins of the reference types
it touches (tests/shim/shim_test.cpp) and run on matrices: what the
reference's caller observes -- the verdict, rowToChange in *pRowOut under
t_saveRowToChange, and the CRowSolution calls TestCanonicity makes
(CanonicityChecker.h:367-405): removeNoncanonicalSolutions(0) when the witness
fails before the last row, moveNoncanonicalSolutions with reconstructSolution's
vector when it fails on the last row, nothing for the last solution -- against
the oracle."""
import os
import subprocess

import numpy as np
import pytest

import ise_amd
from _helpers import unpack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "shim", "build.sh")


def build(tmp_path):
    exe = str(tmp_path / "shim_test")
    subprocess.run([BUILD, exe], check=True)
    return exe


def test_shim_compiles(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_shim_reports_what_the_reference_caller_sees(tmp_path, oracle):
    print("case C++ shim", flush=True)
    exe = build(tmp_path)
    cases = []
    for v, k, lam in ((13, 3, 1), (15, 3, 1), (16, 6, 2), (15, 7, 3)):
        b = ise_amd.design_params(v, k, lam)[1]
        rows, meta = ise_amd.random_extensions(v, k, lam, 300, 3, v, 0x5E + v)
        for i in range(len(meta)):
            d = int(meta[i]) & 0xff
            mat = unpack(rows[i, :d], b)
            out_info = 4 if i % 5 == 0 else 0
            last = 1 if i % 7 == 0 else 0
            cases.append((mat, out_info, last))
    text = "".join(f"{m.shape[0]} {m.shape[1]} {oi} {last}\n" + "\n".join("".join(map(str, r)) for r in m) + "\n"
                   for m, oi, last in cases)
    run = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stderr
    lines = run.stdout.strip().splitlines()
    assert len(lines) == len(cases)
    seen = {"none": 0, "remove": 0, "move": 0, "row_out": 0}
    for (mat, out_info, last), line in zip(cases, lines):
        f = line.split()
        verdict, fail_row, row_out, action = int(f[0]), int(f[1]), int(f[2]), f[3]
        o = oracle.test_canonicity(mat, out_info=out_info)
        d = mat.shape[0]
        assert verdict == o["verdict"][0]
        if verdict:
            assert action == "none" and row_out == 255
            continue
        assert fail_row == o["fail_row"][0]
        if out_info & 4:
            assert row_out == o["row_to_change"][0] and action == "none"
            seen["row_out"] += 1
        elif last:
            assert action == "none"
        elif fail_row + 1 < d:
            assert action == "remove"
            seen["remove"] += 1
        else:
            assert action == "move"
            assert [int(x) for x in f[4:]] == list(oracle.improved_solution(mat))
            seen["move"] += 1
    assert min(seen["remove"], seen["move"], seen["row_out"]) >= 20, seen
