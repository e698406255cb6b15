"""Block V-cycle (n x k right-hand sides) on the GPU: every column equals the single-vector
V-cycle bit for bit, and the FP64 cycle matches the oracle.  Block widths cover one column
per lane (k < 4), four columns per lane with 16-byte loads (k a multiple of 4) and four
columns per lane with single loads and the last group masked (5, 7, 13)."""
import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hier():
    A = H.heterogeneous_diffusion(20, 2.0, 4)[0]
    return A, H.build(A, H.SetupParams(max_coarse=200))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k", [2, 3, 5, 7, 8, 13, 16])
def test_block_columns_equal_single(gpu, hier, prec, k):
    A, h = hier
    M = gpu.AMGPreconditioner(h, prec)
    B = np.asfortranarray(np.stack([H.rng_uniform(c, 2, A.nrows, -1, 1) for c in range(k)], axis=1))
    X = M(B)
    for c in range(k):
        np.testing.assert_array_equal(X[:, c], M(np.ascontiguousarray(B[:, c])))


def test_block_fp64_matches_oracle(gpu, hier):
    A, h = hier
    M = gpu.AMGPreconditioner(h, "fp64")
    ref = oracle.Hier(h)
    B = np.asfortranarray(np.stack([H.rng_uniform(c, 3, A.nrows, -1, 1) for c in range(5)], axis=1))
    X, Xr = M(B), ref.vcycle(B)
    assert np.linalg.norm(X - Xr) <= 1e-12 * np.linalg.norm(Xr)


def test_block_rejects_too_many_columns(gpu, hier):
    A, h = hier
    M = gpu.AMGPreconditioner(h, "fp64")
    with pytest.raises(RuntimeError):
        M(np.zeros((A.nrows, 17), order="F"))
