"""GPU aFSAI setup kernel (chr_afsai) against the host restatement (chs_afsai).

The host factor is pinned separately (closed forms and a dense replay of the greedy growth
in test_setup.py).  The GPU kernel performs the same IEEE operations in the same order, so
the bar here is equality: identical patterns and bit-identical values, and the same
is_posdef.  Cases cover ties in the candidate gradients (constant-coefficient Poisson, which
sends the selection to the sequential partition), step_size > 1, early and late stopping,
coarse-level operators with long rows, rows without a lower part, the reference's own
Flux.mat block (tests/golden), an indefinite matrix and the empty / 1x1 edge cases.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _same(A, nstep=15, step_size=1, eps=1e-3):
    from chronoslab_amd import amg

    Gh, pdh = H.afsai(A, nstep, step_size, eps)
    Gd, pdd = amg.afsai(A, nstep, step_size, eps)
    print(f"n={A.nrows} nnzA={A.nnz} nnzG={Gh.nnz} nstep={nstep} step={step_size} eps={eps}", flush=True)
    assert pdd == pdh
    np.testing.assert_array_equal(Gd.rowptr, Gh.rowptr)
    np.testing.assert_array_equal(Gd.col, Gh.col)
    np.testing.assert_array_equal(Gd.val, Gh.val)
    return Gd, pdd


def _diffusion_levels():
    A, _ = H.heterogeneous_diffusion(22, 2.0, 21)  # levels of 10648, 3531 and 762 rows
    hier = H.build(A)
    assert hier.nlev >= 3
    return [L.A for L in hier.levels[:3]]


def _flux_block():
    d = np.load(os.path.join(GOLDEN, "clas_flux_block1200.npz"))
    return H.Csr(1200, 1200, d["rowptr"], d["col"], d["val"])


CASES = {
    "poisson7_10": lambda: H.poisson7(10),
    "diffusion16": lambda: H.heterogeneous_diffusion(16, 2.0, 3)[0],
    "flux_block1200": _flux_block,
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_afsai_equals_host(gpu, name):
    _same(CASES[name]())


@pytest.mark.parametrize("level", [0, 1, 2])
def test_afsai_coarse_levels(gpu, level):
    _same(_diffusion_levels()[level])


@pytest.mark.parametrize("nstep,step_size,eps", [(5, 2, 1e-3), (4, 3, 1e-3), (15, 1, 0.0), (15, 1, 0.5),
                                                 (0, 1, 1e-3), (1, 1, 1e-3), (31, 2, 1e-6)])
def test_afsai_parameters(gpu, nstep, step_size, eps):
    _same(H.heterogeneous_diffusion(9, 1.5, 8)[0], nstep, step_size, eps)
    _same(H.poisson7(7), nstep, step_size, eps)


def test_afsai_edge_cases(gpu):
    # Diagonal matrix: no row has a lower part; G = diag(1/sqrt(a_ii)).
    D = H.Csr.from_scipy(sp.diags(np.arange(1.0, 51.0)))
    G, pd = _same(D)
    assert pd and G.nnz == 50
    # 1 x 1 and a 2 x 2 SPD matrix.
    _same(H.Csr.from_scipy(sp.csr_matrix(np.array([[4.0]]))))
    _same(H.Csr.from_scipy(sp.csr_matrix(np.array([[4.0, -1.0], [-1.0, 3.0]]))))
    # Indefinite: the Schur complement of the second row is negative.
    G, pd = _same(H.Csr.from_scipy(sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))))
    assert not pd


def test_afsai_empty(gpu):
    from chronoslab_amd import amg

    E = H.Csr(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    G, pd = amg.afsai(E)
    assert G.nrows == 0 and G.nnz == 0 and pd


def test_hierarchy_with_gpu_afsai_equals_host(gpu):
    A, _ = H.heterogeneous_diffusion(18, 2.0, 5)
    h_host = H.build(A)
    h_gpu = H.build(A, H.SetupParams(gpu_afsai=True))
    assert h_host.nlev == h_gpu.nlev
    for a, b in zip(h_host.levels, h_gpu.levels):
        for f in ("A", "P", "G"):
            x, y = getattr(a, f), getattr(b, f)
            if x is None:
                assert y is None
                continue
            np.testing.assert_array_equal(x.rowptr, y.rowptr)
            np.testing.assert_array_equal(x.col, y.col)
            np.testing.assert_array_equal(x.val, y.val)
        assert a.omega == b.omega


# ------------------------------------------------------------------ smoother damping omega
def _jacobi_factor(A):
    d = A.to_scipy().diagonal()
    return H.Csr.from_scipy(sp.diags(1.0 / np.sqrt(d)).tocsr())


def test_omega_closed_form(gpu):
    """G = D^-1/2 on the 7-point Poisson matrix: lambda_max(G A G^T) = 1 + cos(pi/(N+1))."""
    from chronoslab_amd import amg

    N = 8
    A = H.poisson7(N)
    G = _jacobi_factor(A)
    exact = 1.0 + np.cos(np.pi / (N + 1))
    om_d, lam_d = amg.fsai_omega(A, G, seed=3)
    om_h, lam_h = H.fsai_omega(A, G, seed=3)
    print(f"lambda gpu {lam_d!r} host {lam_h!r} exact {exact!r}", flush=True)
    assert abs(lam_d - exact) <= 1e-6 * exact
    assert abs(lam_h - exact) <= 1e-6 * exact
    assert om_d == min(1.0, 1.9 / lam_d)


@pytest.mark.parametrize("level", [0, 1])
def test_omega_matches_host(gpu, level):
    from chronoslab_amd import amg

    A = _diffusion_levels()[level]
    G, _ = H.afsai(A)
    om_d, lam_d = amg.fsai_omega(A, G, seed=5)
    om_h, lam_h = H.fsai_omega(A, G, seed=5)
    print(f"level {level}: lambda gpu {lam_d!r} host {lam_h!r}", flush=True)
    # Lanczos stops at a 1e-7 relative change: the two runs differ in rounding only, which
    # can move the stopping step by one.
    assert abs(lam_d - lam_h) <= 1e-6 * lam_h
    assert abs(om_d - om_h) <= 1e-6 * om_h
    # Dense check of the largest eigenvalue of G A G^T (small level).
    if A.nrows <= 4000:
        M = G.to_scipy() @ A.to_scipy() @ G.to_scipy().T
        ev = np.linalg.eigvalsh(M.toarray())[-1]
        assert abs(lam_d - ev) <= 1e-5 * ev


def test_hierarchy_with_gpu_omega(gpu):
    A, _ = H.heterogeneous_diffusion(18, 2.0, 5)
    h_host = H.build(A)
    h_gpu = H.build(A, H.SetupParams(gpu_afsai=True, gpu_omega=True))
    assert h_host.nlev == h_gpu.nlev
    for a, b in zip(h_host.levels, h_gpu.levels):
        assert abs(a.omega - b.omega) <= 1e-6 * a.omega
        if a.G is not None:
            np.testing.assert_array_equal(a.G.val, b.G.val)
