"""Stored zeros are dropped before upload: products, V-cycles and PCG runs keep their bits.

A row sum starts at +0 and adds rounded products; adding a +-0 product changes no sum, so a
matrix uploaded without its stored zeros gives the same products for finite operands.
CHRONOS_KEEP_ZEROS=1 uploads the matrices as given, for comparison."""
import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _with_env(monkeypatch, env, make):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        return make()
    finally:
        for k in env:
            monkeypatch.delenv(k)


def _with_zeros(A: H.Csr, every: int) -> H.Csr:
    """A copy of A in which every every-th entry above the diagonal and its mirror hold stored
    zeros (the pattern unchanged, so the matrix stays symmetric)."""
    rp, col = A.rowptr, A.col
    val = A.val.copy()
    k = 0
    for i in range(A.nrows):
        for e in range(rp[i], rp[i + 1]):
            j = col[e]
            if j <= i:
                continue
            k += 1
            if k % every:
                continue
            val[e] = 0.0
            row_j = col[rp[j]:rp[j + 1]]
            val[rp[j] + int(np.searchsorted(row_j, i))] = 0.0
    return H.Csr(A.nrows, A.ncols, rp.copy(), col.copy(), val)


def test_spmv_with_stored_zeros_equals_oracle(gpu, monkeypatch):
    A = _with_zeros(H.heterogeneous_diffusion(16, 2.0, 3)[0], 3)
    assert int((A.val == 0.0).sum()) > 0
    x = np.random.default_rng(2).standard_normal(A.ncols)
    ref = oracle.spmv(A, x)  # the stored zeros included
    for keep in ("0", "1"):
        M = _with_env(monkeypatch, {"CHRONOS_KEEP_ZEROS": keep}, lambda: gpu.DeviceMatrix(A))
        np.testing.assert_array_equal(M(x), ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_elasticity_cycle_and_pcg_without_zeros(gpu, monkeypatch, prec):
    """The Q1 elasticity operator stores exact zeros (clamped couplings and vanishing element
    terms); its hierarchy uploaded with and without them gives the same V-cycle bit for bit
    and the same PCG iterations."""
    A, _ = H.elasticity(8)
    assert int((A.val == 0.0).sum()) > 0
    A, _ = H.bc_fix(A)
    hier = H.build(A, H.SetupParams(prolong="bamg", soc="dom", soc_tau=0.01, prol_emin="emin",
                                    tv=H.rigid_body_modes(8), max_coarse=200))
    Md = _with_env(monkeypatch, {"CHRONOS_KEEP_ZEROS": "0"}, lambda: gpu.AMGPreconditioner(hier, prec))
    Mk = _with_env(monkeypatch, {"CHRONOS_KEEP_ZEROS": "1"}, lambda: gpu.AMGPreconditioner(hier, prec))
    b = H.rng_uniform(5, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(Md(b), Mk(b))
    xd, itd, _, fd = Md.pcg(b, itmax=200, tol=1e-8)
    xk, itk, _, fk = Mk.pcg(b, itmax=200, tol=1e-8)
    print(prec, "iterations", itd, itk, flush=True)
    assert fd == fk == 0 and itd == itk
    assert np.linalg.norm(xd - xk) <= 1e-10 * np.linalg.norm(xk)
