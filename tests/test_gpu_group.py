"""Orbits of row solutions under a group on the GPU
(CPermutStorage::findSolutionIndex): the largest image of every solution
equals a brute-force maximum over the group, and representatives map
solutions of one orbit to the same listed solution."""
import numpy as np
import pytest

import ise_amd

pytestmark = pytest.mark.gpu


def test_solution_orbits_match_brute_force(gpu_ctx):
    print("case solution orbits S_5 x C_6 on 12 points", flush=True)
    n = 12
    swap = np.arange(n, dtype=np.uint8)
    swap[[0, 1]] = [1, 0]
    c5 = np.arange(n, dtype=np.uint8)
    c5[:5] = [1, 2, 3, 4, 0]
    c6 = np.arange(n, dtype=np.uint8)
    c6[6:] = [7, 8, 9, 10, 11, 6]
    G = ise_amd.group_closure(np.stack([swap, c5, c6]))
    assert len(G) == 120 * 6
    rng = np.random.default_rng(5)
    sols = rng.integers(0, 3, size=(400, n)).astype(np.uint8)
    # add images of listed solutions so that orbits repeat inside the list
    extra = np.stack([sols[i][G[rng.integers(len(G))]] for i in range(100)])
    sols = np.concatenate([sols, extra])
    canon, rep = gpu_ctx.solution_orbits(G, sols)
    for s in range(len(sols)):
        images = sols[s][G]
        best = max(map(tuple, images.tolist()))
        assert tuple(canon[s].tolist()) == best, s
    listed = {tuple(x): i for i, x in reversed(list(enumerate(sols.tolist())))}
    for s in range(len(sols)):
        want = listed.get(tuple(canon[s].tolist()), 0xffffffff)
        assert rep[s] == want
    # solutions of one orbit share the representative
    for i in range(100):
        assert rep[400 + i] == rep[i] and np.array_equal(canon[400 + i], canon[i])


def test_solution_orbits_rejects_bad_input(gpu_ctx):
    G = np.array([[0, 1, 2], [1, 2, 5]], np.uint8)   # 5 is not a position
    with pytest.raises(ise_amd.ISEError):
        gpu_ctx.solution_orbits(G, np.zeros((2, 3), np.uint8))
