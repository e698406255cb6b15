"""Host setup (hierarchy producer): generators, SoC, PMIS, classical and extended+i
interpolation, aFSAI.

Pins: both interpolations against the reference's own compiled cores (golden
fixtures in tests/golden, and oracle/_ref live where it is built); aFSAI against its closed
forms ((G A)_ij = 0 on the pattern, diag(G A G^T) = 1) and a dense brute-force replay of the
greedy pattern growth; PMIS against the independent/dominating-set properties.
"""
import glob
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

import oracle
from chronoslab_amd import hierarchy as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------- generators
def test_poisson7_shape_and_spectrum():
    N = 6
    A = H.poisson7(N)
    assert A.nrows == N ** 3 and A.nnz == 7 * N ** 3 - 6 * N ** 2
    M = A.to_scipy().toarray()
    assert np.array_equal(M, M.T)
    k = np.arange(1, N + 1)
    lam1 = 2 - 2 * np.cos(k * np.pi / (N + 1))
    closed = np.sort((lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]).ravel())
    np.testing.assert_allclose(np.linalg.eigvalsh(M), closed, rtol=1e-12, atol=1e-12)


def test_diffusion_matrix_properties():
    N = 7
    A, lnk = H.heterogeneous_diffusion(N, 2.0, 5)
    assert A.nnz == 7 * N ** 3 - 6 * N ** 2
    M = A.to_scipy()
    assert abs(M - M.T).max() == 0.0
    k = np.exp(lnk)
    # Face transmissibility is the harmonic mean 2 k_i k_j / (k_i + k_j).
    i, j = 3 + N * (2 + N * 4), 4 + N * (2 + N * 4)
    assert M[i, j] == pytest.approx(-2 * k[i] * k[j] / (k[i] + k[j]), rel=1e-15)
    # Row sums vanish except on the Dirichlet faces x = 0 and x = N-1 (+2 k_i each).
    rs = np.asarray(M.sum(axis=1)).ravel()
    x = np.arange(N ** 3) % N
    expect = np.where(x == 0, 2 * k, 0.0) + np.where(x == N - 1, 2 * k, 0.0)
    np.testing.assert_allclose(rs, expect, rtol=1e-9, atol=1e-9 * k.max())
    assert np.linalg.eigvalsh(M.toarray()).min() > 0


def test_rng_reproducible_and_distributed():
    a = H.rng_uniform(3, 2, 200000, -1, 1)
    b = H.rng_uniform(3, 2, 10, -1, 1, offset=5)
    np.testing.assert_array_equal(a[5:15], b)
    assert abs(a.mean()) < 0.01 and abs(a.var() - 1 / 3) < 0.01
    z = H.rng_normal(3, 1, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert not np.array_equal(H.rng_uniform(3, 2, 10), H.rng_uniform(3, 3, 10))


# ------------------------------------------------------------------------- coarsening
def _symS(A, strong):
    S = sp.csr_matrix(((strong > 0).astype(float), A.col, A.rowptr), shape=(A.nrows, A.nrows))
    S.setdiag(0)
    S.eliminate_zeros()
    return S


def test_soc_cla_m_matrix_measure():
    A = H.poisson7(5)
    s = H.soc_cla(A, 0.25)
    rows = np.repeat(np.arange(A.nrows), np.diff(A.rowptr))
    # Every off-diagonal entry of the Laplacian is -1 = column minimum: measure 1 > tau.
    assert np.all(s[rows != A.col] == 1) and np.all(s[rows == A.col] == 1)
    assert np.all(H.soc_cla(A, 1.0)[rows != A.col] == -1)  # measure 1 is not > 1


def test_soc_cla_weak_positive_entries():
    # Positive off-diagonal entries have a negative measure and are weak.
    M = sp.csr_matrix(np.array([[4.0, -1.0, 0.5], [-1.0, 4.0, -2.0], [0.5, -2.0, 4.0]]))
    A = H.Csr.from_scipy(M)
    s = H.soc_cla(A, 0.25)
    d = dict(zip(zip(np.repeat(np.arange(3), np.diff(A.rowptr)), A.col), s))
    assert d[(0, 2)] == -1 and d[(2, 0)] == -1
    assert d[(0, 1)] == 1  # -1 / max(min_col0=-1, min_col1=-2) = 1
    assert d[(1, 2)] == 1  # -2 / max(-2, -2) = 1


@pytest.mark.parametrize("gen", ["poisson", "diffusion"])
def test_pmis_is_maximal_independent(gen):
    A = H.poisson7(10) if gen == "poisson" else H.heterogeneous_diffusion(10, 2.0, 4)[0]
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(1, 100, A.nrows))
    S = _symS(A, s)
    C = fc >= 0
    assert nc == C.sum() and np.array_equal(np.sort(fc[C]), np.arange(nc))
    assert np.all(np.diff(fc[C]) > 0)  # coarse numbering follows fine order
    Ssym = ((S + S.T) > 0).astype(int)
    assert (Ssym[C][:, C]).nnz == 0  # independent
    dominated = np.asarray(Ssym[:, C].sum(axis=1)).ravel() > 0
    assert np.all(dominated[~C])  # every fine node has a strong coarse neighbour


def test_pmis_dirichlet_rows_are_fine():
    A0 = H.poisson7(5).to_scipy().tolil()
    A0[7, :] = 0
    A0[:, 7] = 0
    A0[7, 7] = 1.0
    A = H.Csr.from_scipy(A0.tocsr())
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(1, 100, A.nrows))
    assert fc[7] == -1


# ---------------------------------------------------------------- classical interpolation
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "clas_*.npz"))))
def test_classical_prolongation_golden(path):
    d = np.load(path)
    n = d["rowptr"].shape[0] - 1
    A = H.Csr(n, n, d["rowptr"], d["col"], d["val"])
    P = H.prol_classical(A, d["strong"], d["fcnode"], int(d["nc"]), nthreads=3)
    np.testing.assert_array_equal(P.rowptr, d["P_rowptr"])
    np.testing.assert_array_equal(P.col, d["P_col"])
    np.testing.assert_array_equal(P.val, d["P_val"])  # bit for bit


def test_golden_fixtures_present():
    assert len(glob.glob(os.path.join(GOLDEN, "clas_*.npz"))) >= 3


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_classical_prolongation_vs_reference_core(seed):
    A = H.heterogeneous_diffusion(12, 1.5 + seed / 2, seed)[0]
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(seed, 100, A.nrows))
    P = H.prol_classical(A, s, fc, nc)
    iat, ja, co = oracle.ref_classical_prol(A, s, fc, nc, nthreads=2)
    np.testing.assert_array_equal(P.rowptr, iat)
    np.testing.assert_array_equal(P.col, ja)
    np.testing.assert_array_equal(P.val, co)


def test_classical_prolongation_invariants():
    A = H.poisson7(8)
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(2, 100, A.nrows))
    P = H.prol_classical(A, s, fc, nc).to_scipy().toarray()
    for i in np.where(fc >= 0)[0]:
        e = np.zeros(nc)
        e[fc[i]] = 1
        np.testing.assert_array_equal(P[i], e)
    # Interior fine rows of a zero-row-sum operator interpolate constants exactly.
    N = 8
    g = np.arange(N ** 3)
    x, y, z = g % N, (g // N) % N, g // (N * N)
    interior = (x > 1) & (x < N - 2) & (y > 1) & (y < N - 2) & (z > 1) & (z < N - 2) & (fc < 0)
    np.testing.assert_allclose(P[interior].sum(axis=1), 1.0, rtol=1e-13)


# ------------------------------------------------------------- extended+i interpolation
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "exti_*.npz"))))
def test_exti_prolongation_golden(path):
    d = np.load(path)
    n = d["rowptr"].shape[0] - 1
    A = H.Csr(n, n, d["rowptr"], d["col"], d["val"])
    P = H.prol_exti(A, d["strong"], d["fcnode"], int(d["nc"]), nthreads=3)
    np.testing.assert_array_equal(P.rowptr, d["P_rowptr"])
    np.testing.assert_array_equal(P.col, d["P_col"])
    np.testing.assert_array_equal(P.val, d["P_val"])  # bit for bit


def test_exti_golden_fixtures_present():
    assert len(glob.glob(os.path.join(GOLDEN, "exti_*.npz"))) >= 3


@pytest.mark.skipif(not oracle.ref_exti_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_exti_prolongation_vs_reference_core(seed):
    A = H.heterogeneous_diffusion(12, 1.5 + seed / 2, seed)[0]
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(seed, 100, A.nrows))
    P = H.prol_exti(A, s, fc, nc)
    iat, ja, co = oracle.ref_exti_prol(A, s, fc, nc, nthreads=2)
    np.testing.assert_array_equal(P.rowptr, iat)
    np.testing.assert_array_equal(P.col, ja)
    np.testing.assert_array_equal(P.val, co)


def test_exti_prolongation_invariants():
    # Coarse rows are unit rows; interior fine rows of a zero-row-sum operator interpolate
    # constants exactly (the extended+i weights and denominator add up to the row sum).
    N = 8
    A = H.poisson7(N)
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(4, 100, A.nrows))
    P = H.prol_exti(A, s, fc, nc).to_scipy().toarray()
    for i in np.where(fc >= 0)[0]:
        e = np.zeros(nc)
        e[fc[i]] = 1
        np.testing.assert_array_equal(P[i], e)
    g = np.arange(N ** 3)
    x, y, z = g % N, (g // N) % N, g // (N * N)
    interior = (x > 1) & (x < N - 2) & (y > 1) & (y < N - 2) & (z > 1) & (z < N - 2) & (fc < 0)
    assert interior.any()
    np.testing.assert_allclose(P[interior].sum(axis=1), 1.0, rtol=1e-13)


# ------------------------------------------------------------------------------ aFSAI
def _afsai_checks(A, G, nstep):
    Ad = A.to_scipy().toarray()
    Gd = G.to_scipy().toarray()
    n = A.nrows
    for i in range(n):
        cols = G.col[G.rowptr[i]:G.rowptr[i + 1]]
        assert cols[-1] == i and np.all(np.diff(cols) > 0) and len(cols) <= nstep + 1
        pat = cols[:-1]
        if len(pat):
            ga = Gd[i] @ Ad[:, pat]
            assert np.abs(ga).max() <= 1e-12 * np.abs(Gd[i]).max() * np.abs(Ad).max()
    np.testing.assert_allclose(np.diag(Gd @ Ad @ Gd.T), 1.0, rtol=1e-12)


@pytest.mark.parametrize("gen", ["poisson", "diffusion"])
def test_afsai_closed_forms(gen):
    A = H.poisson7(6) if gen == "poisson" else H.heterogeneous_diffusion(6, 2.0, 8)[0]
    G, pd = H.afsai(A, 15, 1, 1e-3)
    assert pd
    _afsai_checks(A, G, 15)


def _dense_afsai_pattern(Ad, i, nstep, eps):
    """Greedy growth replayed with dense linear algebra (no ties in the inputs used)."""
    pat = []
    x = np.zeros(0)
    dk_old = 0.0
    for step in range(1, nstep + 1):
        g = Ad[i, :i].copy()
        if pat:
            g += x @ Ad[pat, :i]
        cand = [j for j in range(i) if j not in pat and (Ad[i, j] != 0 or (pat and np.any(Ad[pat, j] != 0)))]
        if not cand:
            break
        j = max(cand, key=lambda c: abs(g[c]))
        pat = sorted(pat + [j])
        a = Ad[np.ix_(pat, pat)]
        x = np.linalg.solve(a, -Ad[i, pat])
        dk = -x @ Ad[i, pat]
        if step == nstep:
            break
        if not (abs(dk - dk_old) >= eps * dk_old and dk != 0):
            break
        dk_old = dk
    return pat


def test_afsai_greedy_pattern_matches_dense_replay():
    A = H.heterogeneous_diffusion(5, 2.0, 21)[0]  # continuous random values: no ties
    G, _ = H.afsai(A, 8, 1, 1e-3)
    Ad = A.to_scipy().toarray()
    for i in range(A.nrows):
        got = list(G.col[G.rowptr[i]:G.rowptr[i + 1] - 1])
        assert got == _dense_afsai_pattern(Ad, i, 8, 1e-3), i


def test_afsai_values_match_dense_solve():
    A = H.heterogeneous_diffusion(5, 1.0, 3)[0]
    G, _ = H.afsai(A, 6, 2, 1e-3)
    Ad = A.to_scipy().toarray()
    for i in range(A.nrows):
        cols = G.col[G.rowptr[i]:G.rowptr[i + 1]]
        vals = G.val[G.rowptr[i]:G.rowptr[i + 1]]
        pat = cols[:-1]
        x = np.linalg.solve(Ad[np.ix_(pat, pat)], -Ad[i, pat]) if len(pat) else np.zeros(0)
        s = 1 / np.sqrt(Ad[i, i] + Ad[i, pat] @ x)
        np.testing.assert_allclose(vals, np.append(s * x, s), rtol=1e-11)


def test_afsai_flags_indefinite_matrix():
    M = sp.csr_matrix(np.array([[1.0, 2.0, 0.0], [2.0, 1.0, 1.0], [0.0, 1.0, 1.0]]))
    G, pd = H.afsai(H.Csr.from_scipy(M), 4, 1, 1e-3)
    assert not pd


def test_fsai_omega_matches_dense_eigenvalue():
    A = H.heterogeneous_diffusion(6, 2.0, 2)[0]
    G, _ = H.afsai(A)
    om, lam = H.fsai_omega(A, G)
    Gd, Ad = G.to_scipy().toarray(), A.to_scipy().toarray()
    lmax = np.linalg.eigvalsh(Gd @ Ad @ Gd.T).max()
    assert lam == pytest.approx(lmax, rel=1e-7)
    assert om == pytest.approx(min(1.0, 1.9 / lmax), rel=1e-7)


def test_galerkin_matches_dense():
    A = H.heterogeneous_diffusion(6, 1.0, 2)[0]
    s = H.soc_cla(A, 0.25)
    fc, nc = H.pmis(A, s, H.rng_uniform(1, 100, A.nrows))
    P = H.prol_classical(A, s, fc, nc)
    Ac = H.galerkin(A, P)
    Pd, Ad = P.to_scipy().toarray(), A.to_scipy().toarray()
    np.testing.assert_allclose(Ac.to_scipy().toarray(), Pd.T @ Ad @ Pd, rtol=1e-13, atol=1e-13 * abs(Ad).max())
    assert np.all(np.diff(Ac.col)[np.diff(np.repeat(np.arange(nc), np.diff(Ac.rowptr))) == 0] > 0)


def test_bc_fix_rule():
    M = H.poisson7(4).to_scipy().tolil()
    M[3, 3] = 1.0
    A = H.Csr.from_scipy(M.tocsr())
    Af, nf = H.bc_fix(A)
    assert nf == 1
    lmax = np.linalg.eigvalsh(A.to_scipy().toarray()).max()
    assert Af.to_scipy()[3, 3] == pytest.approx(lmax / 10, rel=1e-4)
    assert H.bc_fix(H.poisson7(4))[1] == 0


def test_hierarchy_shapes_and_stopping_rule():
    A = H.poisson7(16)
    h = H.build(A, H.SetupParams(max_coarse=100))
    assert h.levels[-1].A.nrows <= 100
    for l in range(h.nlev - 1):
        L = h.levels[l]
        assert L.P.nrows == L.A.nrows and L.P.ncols == h.levels[l + 1].A.nrows
        assert L.G.nrows == L.A.nrows
        assert 0 < L.omega <= 1
    h2 = H.build(A, H.SetupParams(nlev_max=2))
    assert h2.nlev == 2
