"""Differentiable AMG-PCG solve of the finite-volume diffusion problem (torch consumer of
the C ABI).

    x = FVDiffusionSolve.apply(lnk, b, N)          # A(exp(lnk)) x = b, AMG-PCG on the GPU
    loss = 0.5 * ((x - x_obs) ** 2).sum(); loss.backward()

forward: setup of the AMG hierarchy for A(k) (classical coarsening on the host, aFSAI factors
and smoother damping on the GPU), upload, PCG_CJ with
the chosen V-cycle precision.  backward: the adjoint A^T lam = dL/dx is solved with the same
hierarchy (A is SPD, so A^T = A), and one fused kernel (chr_fv_grad) forms dL/d(ln k) on A's
sparsity pattern; dL/db = lam.  The reference has no counterpart.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import amg, hierarchy as H


def fv_grad(M: amg.AMGPreconditioner, k: torch.Tensor, x: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
    """dL/d(ln k) from the forward field x and the adjoint field lam (device tensors)."""
    g = torch.empty_like(x)
    torch.cuda.current_stream().synchronize()
    amg._check(amg._lib().chr_fv_grad(M.ptr, None, k.data_ptr(), x.data_ptr(), lam.data_ptr(), g.data_ptr(),
                                      amg.PTR_DEVICE), "chr_fv_grad")
    return g


class FVDiffusionSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lnk, b, N: int, tol: float = 1e-10, itmax: int = 1000, precision: str = "fp32",
                params: H.SetupParams | None = None):
        k_host = np.exp(lnk.detach().double().cpu().numpy())
        A = H.diffusion(N, k_host)
        if params is None:  # aFSAI factors and omega on the GPU
            params = H.SetupParams(gpu_afsai=True, gpu_omega=True, gpu_galerkin=True)
        M = amg.AMGPreconditioner(H.build(A, params), precision)
        bb = b.detach().double().contiguous()
        x = torch.zeros_like(bb)
        torch.cuda.current_stream().synchronize()
        it, rv, fl = M.pcg_device(bb.data_ptr(), x.data_ptr(), itmax=itmax, tol=tol)
        if fl != 0:
            raise RuntimeError(f"forward PCG did not reach {tol:g} in {itmax} iterations")
        ctx.M, ctx.tol, ctx.itmax = M, tol, itmax
        ctx.forward_iters = it
        kd = torch.from_numpy(k_host).to(bb.device)
        ctx.save_for_backward(kd, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        kd, x = ctx.saved_tensors
        gxc = gx.detach().double().contiguous()
        lam = torch.zeros_like(x)
        torch.cuda.current_stream().synchronize()
        it, rv, fl = ctx.M.pcg_device(gxc.data_ptr(), lam.data_ptr(), itmax=ctx.itmax, tol=ctx.tol)
        if fl != 0:
            raise RuntimeError("adjoint PCG did not converge")
        ctx.adjoint_iters = it
        g = fv_grad(ctx.M, kd, x, lam)
        return g, lam, None, None, None, None, None
