"""The device white noise (nnx_noise_normal; the leapfrog step's draws)
against the oracle's restatement of the same generator: every normal bit
for bit, over a million neurons at several steps and seeds, including the
far tails the ln and sqrt of u1 near 0 produce."""
import numpy as np
import pytest

import oracle
from neuronephys_clean_amd.devmem import DeviceArray as DA

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,step,first", [(0, 0, 0), (0, 1, 0), (12345, 987654321, 3), (2 ** 63 + 5, 7, 1 << 30)])
def test_device_normals_equal_oracle(ctx, seed, step, first):
    n = 1 << 20
    d = DA.zeros(ctx, n, np.float64)
    ctx.call("nnx_noise_normal", seed, step, first, n, d.ptr)
    got = d.numpy()
    ref = oracle.noise_normal(seed, step, n, first=first)
    bad = np.nonzero(got.view(np.uint64) != ref.view(np.uint64))[0]
    print(f"noise seed={seed} step={step} first={first}: {bad.size} differing, range [{got.min():.3f}, {got.max():.3f}]",
          flush=True)
    assert bad.size == 0, (bad[:5], got[bad[:5]], ref[bad[:5]])
    assert abs(got.mean()) < 5e-3 and abs(got.var() - 1) < 1e-2
