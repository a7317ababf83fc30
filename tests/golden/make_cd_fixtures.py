"""Expected outputs of runCanonizer for full-day K_10 factorizations, written by
the CPU oracle (orc_cd_run_canonizer) into cd_canon.json.

A full-day run canonizes all 9! = 362880 day orders, about 40 s of oracle work
per k-system, too slow for a GPU test to recompute.  The oracle is pinned
against brute force and invariance checks in tests/test_oracle_cdtools.py;
these vectors only cache its answers.  Inputs: the first matrix of the k10,
p10 and u10 fixtures of kfact.json (published 1-factorizations of K_10).

Run: python tests/golden/make_cd_fixtures.py   (builds nothing; needs the
oracle library, oracle/liboracle_ise.so)
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

CASES = ["k10", "p10", "u10"]


def one(name):
    import numpy as np
    import _oracle
    kf = json.load(open(os.path.join(HERE, "kfact.json")))
    mat = np.array(kf[name]["matrices"][0]["rounds"], np.uint8)
    v = mat.shape[1]
    out = _oracle.load().cd_run_canonizer(mat, v, 2)
    rows = ["".join(str(int(x)) for x in r) for r in out]
    return {"fixture": name, "index": 0, "v": v, "k": 2, "day": 0, "canon": rows}


def main():
    with ProcessPoolExecutor(len(CASES)) as ex:
        res = list(ex.map(one, CASES))
    with open(os.path.join(HERE, "cd_canon.json"), "w") as f:
        json.dump({"source": "oracle orc_cd_run_canonizer (see make_cd_fixtures.py)", "cases": res}, f, indent=1)


if __name__ == "__main__":
    main()
