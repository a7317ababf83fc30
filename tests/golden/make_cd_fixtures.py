"""Expected outputs of runCanonizer for full-day k-systems, written by the CPU
oracle (orc_cd_run_canonizer) into cd_canon.json.

A full-day K_10 run canonizes all 9! = 362880 day orders, about 40 s of oracle
work per k-system, too slow for a GPU test to recompute.  The oracle is pinned
against brute force, invariance and |Aut(M)| checks in
tests/test_oracle_cdtools.py; these vectors only cache its answers.  Inputs:
the first matrix of the k10, p10 and u10 fixtures of kfact.json (published
1-factorizations of K_10) and the seven k15x3 k-systems (15 players, seven
days of five triples: the seven Kirkman schoolgirl solutions).

Run: python tests/golden/make_cd_fixtures.py   (builds nothing; needs the
oracle library, oracle/liboracle_ise.so)
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

CASES = [("k10", 0, 2), ("p10", 0, 2), ("u10", 0, 2)] + [("k15x3", i, 3) for i in range(7)]


def one(case):
    import numpy as np
    import _oracle
    name, index, k = case
    kf = json.load(open(os.path.join(HERE, "kfact.json")))
    mat = np.array(kf[name]["matrices"][index]["rounds"], np.uint8)
    v = mat.shape[1]
    out = _oracle.load().cd_run_canonizer(mat, v, k)
    rows = ["".join(str(int(x)) for x in r) for r in out]
    return {"fixture": name, "index": index, "v": v, "k": k, "day": 0, "canon": rows}


def main():
    with ProcessPoolExecutor(4) as ex:
        res = list(ex.map(one, CASES))
    with open(os.path.join(HERE, "cd_canon.json"), "w") as f:
        json.dump({"source": "oracle orc_cd_run_canonizer (see make_cd_fixtures.py)", "cases": res}, f, indent=1)


if __name__ == "__main__":
    main()
