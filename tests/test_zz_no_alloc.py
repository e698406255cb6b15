"""chr_pcg_solve_noalloc: PCG_CJ whose working memory is the library's fixed device scratch
array.  Checked against chr_pcg_solve on the same device inputs: same iterations, residual
history and solution bit for bit, on two calls in a row (the second smaller, so it runs over
the first call's leftovers in the array)."""
import numpy as np
import pytest

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _case(N, seed):
    A, _ = H.heterogeneous_diffusion(N, 1.5, seed)
    return A, H.build(A), H.rng_uniform(seed, 2, A.nrows, -1.0, 1.0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_equals_pcg_solve_twice(gpu, prec):
    import torch

    for N, seed in ((24, 3), (14, 4)):
        A, hier, bh = _case(N, seed)
        print(f"no-alloc PCG {prec} {N}^3 seed {seed}", flush=True)
        M = gpu.AMGPreconditioner(hier, prec)
        b = torch.from_numpy(bh).cuda()
        x0 = torch.from_numpy(H.rng_uniform(seed, 3, A.nrows, -0.5, 0.5)).cuda()
        for start in (torch.zeros_like(b), x0):
            x1, x2 = start.clone(), start.clone()
            it1, rv1, fl1 = M.pcg_device(b.data_ptr(), x1.data_ptr(), itmax=200, tol=1e-8)
            it2, rv2, fl2 = M.pcg_noalloc_device(b.data_ptr(), x2.data_ptr(), itmax=200, tol=1e-8)
            torch.cuda.synchronize()
            assert fl1 == fl2 == 0 and it1 == it2 and it1 > 0
            np.testing.assert_array_equal(rv1, rv2)
            np.testing.assert_array_equal(x1.cpu().numpy(), x2.cpu().numpy())
            res = np.linalg.norm(bh - A.matvec(x2.cpu().numpy())) / np.linalg.norm(bh)
            assert res <= 1e-8


def test_noalloc_edge_cases_and_limit(gpu):
    import torch

    A, hier, bh = _case(12, 5)
    M = gpu.AMGPreconditioner(hier, "fp64")
    # b = 0: x = 0 after 0 iterations (PCG_CJ.m:11-16), whatever x0 was.
    z = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
    x = torch.ones_like(z)
    it, rv, fl = M.pcg_noalloc_device(z.data_ptr(), x.data_ptr(), itmax=50, tol=1e-8)
    torch.cuda.synchronize()
    assert it == 0 and fl == 0 and not x.cpu().numpy().any()
    # itmax = 1: one iteration, not converged, the same history as chr_pcg_solve.
    b = torch.from_numpy(bh).cuda()
    x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
    r1 = M.pcg_device(b.data_ptr(), x1.data_ptr(), itmax=1, tol=1e-8)
    r2 = M.pcg_noalloc_device(b.data_ptr(), x2.data_ptr(), itmax=1, tol=1e-8)
    torch.cuda.synchronize()
    assert r1[0] == r2[0] == 1 and r1[2] == r2[2] == 1
    np.testing.assert_array_equal(r1[1], r2[1])
    np.testing.assert_array_equal(x1.cpu().numpy(), x2.cpu().numpy())
    # A residual history beyond the scratch array is refused before any work.
    x = torch.zeros_like(b)
    with pytest.raises(RuntimeError, match="scratch"):
        M.pcg_noalloc_device(b.data_ptr(), x.data_ptr(), itmax=10_000_000, tol=1e-8)
    assert not x.cpu().numpy().any()
