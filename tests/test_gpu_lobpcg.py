"""LOBPCG on the GPU against the oracle (same hierarchy, same X0) and against the closed-form
spectrum of the Dirichlet 7-point Laplacian."""
import time

import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def poisson_eigs(N, k):
    j = np.arange(1, N + 1)
    l1 = 2 - 2 * np.cos(j * np.pi / (N + 1))
    return np.sort((l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel())[:k]


def x0(n, k, seed=1):
    return np.asfortranarray(np.stack([H.rng_uniform(seed, 10 + c, n) for c in range(k)], axis=1))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lobpcg_matches_oracle(gpu, prec):
    """Iteration count within +-1 and eigenvalues within 1e-10 of the oracle on a spectrum
    without repeated eigenvalues (heterogeneous diffusion)."""
    A = H.heterogeneous_diffusion(14, 1.0, 5)[0]
    h = H.build(A)
    M = gpu.AMGPreconditioner(h, prec)
    X0 = x0(A.nrows, 8)
    it, lam, X, rh, lh, ierr = gpu.lobpcg(None, M, X0, itmax=200, tol=1e-8)
    ito, lamo, Xo, rho, lho, ierro = oracle.lobpcg(A, oracle.Hier(h), X0, itmax=200, tol=1e-8)
    ev = np.linalg.eigvalsh(A.to_scipy().toarray())[:8]
    print(f"{prec}: gpu it {it} ierr {ierr}; oracle it {ito} ierr {ierro}; "
          f"max rel diff {np.max(np.abs(lam - lamo) / lamo):.2e}", flush=True)
    assert ierr == 0 and ierro == 0
    assert abs(it - ito) <= 1
    np.testing.assert_allclose(lam, lamo, rtol=1e-10)
    np.testing.assert_allclose(lam, ev, rtol=1e-9)
    R = A.to_scipy() @ X - X * lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-6 * lam.max()


def test_lobpcg_gram_table_equals_single_products(gpu, monkeypatch):
    """The Rayleigh-Ritz Gram blocks computed in one pass over the blocks (a table of the
    matrix-core Gram kernel) give the same bits as one pass per product
    (CHRONOS_GRAM_MULTI=0): same iterations, eigenvalues and eigenvectors.  n = 13^3 and five
    pairs, so neither the rows nor the columns fill whole 16 x 4 instruction tiles."""
    A = H.heterogeneous_diffusion(13, 1.0, 7)[0]
    h = H.build(A)
    M = gpu.AMGPreconditioner(h, "fp64")
    X0 = x0(A.nrows, 5)
    r1 = gpu.lobpcg(None, M, X0, itmax=60, tol=1e-9)
    monkeypatch.setenv("CHRONOS_GRAM_MULTI", "0")
    r2 = gpu.lobpcg(None, M, X0, itmax=60, tol=1e-9)
    print(f"iterations {r1[0]} / {r2[0]}, ierr {r1[5]} / {r2[5]}", flush=True)
    assert r1[5] == 0 and r1[0] == r2[0]
    np.testing.assert_array_equal(r1[1], r2[1])
    np.testing.assert_array_equal(r1[2], r2[2])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lobpcg_poisson_closed_form(gpu, prec):
    """16^3 Poisson: eigenvalues within 1e-8 of the closed form.  The spectrum has exactly
    repeated eigenvalues, so eigenvectors inside a cluster are not unique and the iteration
    path may differ from the oracle's by rounding; only the eigenvalues are compared."""
    N, k = 16, 8
    A = H.poisson7(N)
    h = H.build(A)
    M = gpu.AMGPreconditioner(h, prec)
    it, lam, X, rh, lh, ierr = gpu.lobpcg(None, M, x0(A.nrows, k), itmax=200, tol=1e-8)
    print(f"{prec}: poisson 16^3 it {it} ierr {ierr}", flush=True)
    assert ierr == 0
    np.testing.assert_allclose(lam, poisson_eigs(N, k), rtol=1e-8)


def test_lobpcg_unpreconditioned_largest(gpu):
    A = H.heterogeneous_diffusion(6, 1.0, 4)[0]
    ev = np.linalg.eigvalsh(A.to_scipy().toarray())
    X0 = x0(A.nrows, 4, 2)
    it, lam, X, rh, lh, ierr = gpu.lobpcg(A, None, X0, largest=True, itmax=400, tol=1e-9)
    ito, lamo, Xo, rho, lho, ierro = oracle.lobpcg(A, None, X0, largest=True, itmax=400, tol=1e-9)
    np.testing.assert_allclose(lam, np.sort(ev[::-1][:4]), rtol=1e-7)
    np.testing.assert_allclose(lam, lamo, rtol=1e-10)
    assert abs(it - ito) <= 1


def test_lobpcg_rank_deficient_start(gpu):
    A = H.poisson7(6)
    it, lam, X, rh, lh, ierr = gpu.lobpcg(A, None, np.ones((A.nrows, 3)), itmax=20, tol=1e-8)
    assert ierr == 1


def test_lobpcg_96cube_16_pairs(gpu):
    """The LOBPCG configuration: 96^3 Poisson, 16 smallest pairs, AMG preconditioner,
    FP32 V-cycle, FP64 Rayleigh-Ritz; eigenvalue relative error <= 1e-8 vs closed form."""
    N, k = 96, 16
    A = H.poisson7(N)
    h = H.build(A)
    M = gpu.AMGPreconditioner(h, "fp32")
    X0 = x0(A.nrows, k, 7)
    t = time.perf_counter()
    it, lam, X, rh, lh, ierr = gpu.lobpcg(None, M, X0, itmax=300, tol=1e-8)
    dt = time.perf_counter() - t
    err = np.abs(lam - poisson_eigs(N, k)) / poisson_eigs(N, k)
    print(f"96^3 LOBPCG: {it} iterations, {dt:.2f} s, ierr {ierr}, max eig rel err {err.max():.2e}", flush=True)
    assert ierr == 0
    assert err.max() <= 1e-8


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lobpcg_constraint_block_against_oracle(gpu, prec):
    """The constraint block of EigDriver.m:213-223 (blockY, lobpcg.m:64-75,228-230,384-387):
    Y spans the eigenvectors of the four smallest eigenvalues of the 12^3 Poisson matrix, so
    the constrained solve finds the next three (closed form, numpy.linalg.eigh), with X
    orthogonal to Y; eigenvalues within 1e-10 of the oracle's and iterations within one."""
    A = H.poisson7(12)
    w, V = np.linalg.eigh(A.to_scipy().toarray())
    Y = V[:, :4]
    M = gpu.AMGPreconditioner(H.build(A), prec)
    X0 = x0(A.nrows, 3, 21)
    it, lam, X, rh, lh, ierr = gpu.lobpcg(None, M, X0, itmax=200, tol=1e-9, Y=Y)
    ito, lamo, Xo, rho, lho, ierro = oracle.lobpcg(A, oracle.Hier(M.hier), X0, itmax=200, tol=1e-9, Y=Y)
    assert ierr == 0 and ierro == 0
    np.testing.assert_allclose(lam, w[4:7], rtol=1e-9)
    np.testing.assert_allclose(lam, lamo, rtol=1e-10)
    assert abs(it - ito) <= 1
    assert np.abs(Y.T @ X).max() <= 1e-9
    with pytest.raises(RuntimeError, match="rank deficient"):
        gpu.lobpcg(None, M, X0, itmax=20, tol=1e-9, Y=np.stack([Y[:, 0], Y[:, 0]], 1))
