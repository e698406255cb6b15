"""Differentiation row: the adjoint gradient through torch.autograd against the FP64 CPU
oracle gradient and against central finite differences."""
import time

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from chronoslab_amd import amg, hierarchy as H

pytestmark = pytest.mark.gpu


def test_fused_grad_kernel_matches_oracle(gpu):
    N = 10
    lnk = H.lnk_field(N, 1.0, 3)
    k = np.exp(lnk)
    A = H.diffusion(N, k)
    x = H.rng_uniform(1, 2, N ** 3, -1, 1)
    lam = H.rng_uniform(2, 2, N ** 3, -1, 1)
    Adev = gpu.DeviceMatrix(A)
    g = np.empty(N ** 3)
    amg._check(amg._lib().chr_fv_grad(None, Adev.ptr, k.ctypes.data, x.ctypes.data, lam.ctypes.data,
                                      g.ctypes.data, amg.PTR_HOST))
    gr = oracle.fv_grad(N, k, x, lam)
    assert np.linalg.norm(g - gr) <= 1e-13 * np.linalg.norm(gr)


def _loss_and_grad(N, lnk, lnk_obs, b, prec, tol):
    import torch

    from chronoslab_amd.autograd import FVDiffusionSolve

    dev = torch.device("cuda:0")
    bt = torch.from_numpy(b).to(dev)
    x_obs = FVDiffusionSolve.apply(torch.from_numpy(lnk_obs).to(dev), bt, N, tol, 1000, "fp64").detach()
    l = torch.from_numpy(lnk).to(dev).requires_grad_(True)
    x = FVDiffusionSolve.apply(l, bt, N, tol, 1000, prec)
    loss = 0.5 * ((x - x_obs) ** 2).sum()
    loss.backward()
    return loss.item(), l.grad.cpu().numpy(), x.detach().cpu().numpy(), x_obs.cpu().numpy()


def _cpu_grad(N, lnk, x_obs, b, tol):
    A = H.diffusion(N, np.exp(lnk))
    h = H.build(A)
    oh = oracle.Hier(h)
    x, it, rv, fl = oh.pcg(A, b, itmax=1000, tol=tol)
    lam, it2, rv2, fl2 = oh.pcg(A, x - x_obs, itmax=1000, tol=tol)
    return oracle.fv_grad(N, np.exp(lnk), x, lam), x


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_small_vs_cpu_and_fd(gpu, prec):
    N = 12
    lnk, lnk_obs = H.lnk_field(N, 1.0, 21), H.lnk_field(N, 1.0, 22)
    b = H.rng_uniform(5, 2, N ** 3, -1, 1)
    loss, g, x, x_obs = _loss_and_grad(N, lnk, lnk_obs, b, prec, 1e-10)
    gc, xc = _cpu_grad(N, lnk, x_obs, b, 1e-12)
    rel = np.linalg.norm(g - gc) / np.linalg.norm(gc)
    print(f"{prec}: |g - g_cpu|/|g_cpu| = {rel:.2e}", flush=True)
    assert rel <= 1e-6
    # Central differences with exact (direct) solves.
    def L(l):
        xx = spla.spsolve(H.diffusion(N, np.exp(l)).to_scipy().tocsc(), b)
        return 0.5 * np.sum((xx - x_obs) ** 2)
    for c in (0, 17, 500, 1000, 1727):
        e = np.zeros(N ** 3)
        e[c] = 1e-6
        fd = (L(lnk + e) - L(lnk - e)) / 2e-6
        assert g[c] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())


def test_gradient_96cube_config(gpu):
    """96^3, ln k ~ N(0,1): forward FP32-cycle PCG to 1e-10, adjoint, dL/d(ln k); checked
    against the FP64 CPU gradient (rel 1e-6) and central differences (step 1e-6, 32 cells)."""
    import torch

    N = 96
    lnk, lnk_obs = H.lnk_field(N, 1.0, 31), H.lnk_field(N, 1.0, 32)
    b = H.rng_uniform(9, 2, N ** 3, -1, 1)
    t = time.perf_counter()
    loss, g, x, x_obs = _loss_and_grad(N, lnk, lnk_obs, b, "fp32", 1e-10)
    t_gpu = time.perf_counter() - t
    t = time.perf_counter()
    gc, xc = _cpu_grad(N, lnk, x_obs, b, 1e-10)
    t_cpu = time.perf_counter() - t
    rel = np.linalg.norm(g - gc) / np.linalg.norm(gc)
    print(f"96^3: loss {loss:.6e}; |g-g_cpu|/|g_cpu| = {rel:.2e}; gpu {t_gpu:.1f}s (incl. 2 setups), "
          f"cpu {t_cpu:.1f}s", flush=True)
    assert rel <= 1e-6
    # Finite differences: the perturbed operators are preconditioned by the unperturbed
    # hierarchy (the preconditioner does not change the converged solution).
    A0 = H.diffusion(N, np.exp(lnk))
    M = amg.AMGPreconditioner(H.build(A0), "fp64")
    rng = np.random.default_rng(4)
    cells = rng.choice(N ** 3, 32, replace=False)
    h = 1e-6
    fds = []
    for c in cells:
        vals = []
        for sgn in (1, -1):
            l = lnk.copy()
            l[c] += sgn * h
            xx, it, rv, fl = M.pcg(b, itmax=500, tol=1e-13, A=amg.DeviceMatrix(H.diffusion(N, np.exp(l))))
            vals.append(0.5 * np.sum((xx - x_obs) ** 2))
        fds.append((vals[0] - vals[1]) / (2 * h))
    fds = np.array(fds)
    err = np.abs(g[cells] - fds) / np.maximum(np.abs(fds), 1e-3 * np.abs(g).max())
    print(f"96^3 FD check: max rel err {err.max():.2e} over 32 cells", flush=True)
    assert err.max() <= 1e-3
