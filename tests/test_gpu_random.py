"""The GPU generator of seeded random valid extensions runs the host walk one
thread per candidate: its batches equal the host's, with and without
rejection sampling of the depth."""
import numpy as np
import pytest

import ise_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("v,k,lam,dmin,dmax,attempts,n", [(15, 3, 1, 4, 15, 4096, 1536), (25, 4, 1, 8, 16, 1, 4096),
                                                          (16, 6, 2, 5, 16, 64, 2048), (19, 9, 4, 5, 19, 16, 1024)])
def test_device_batch_equals_host_batch(gpu_ctx, v, k, lam, dmin, dmax, attempts, n):
    print(f"case random {(v, k, lam)} depths {dmin}-{dmax} attempts {attempts} n {n}", flush=True)
    host = ise_amd.random_extensions(v, k, lam, n, dmin, dmax, 0xA5 + v, stride=32, attempts=attempts)
    dev = ise_amd.random_extensions(v, k, lam, n, dmin, dmax, 0xA5 + v, stride=32, attempts=attempts, ctx=gpu_ctx)
    assert np.array_equal(host[1], dev[1])
    assert np.array_equal(host[0], dev[0])
