"""The driver's caller semantics around the solve: the two Dirichlet-row rules and the
right-hand-side modes (sources/utils/computeAMG.m:5-10;
sources/Drivers/driver_AMG/aspAMG_main.m:89-90,119-134)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.stats

from chronoslab_amd import hierarchy as H


def _with_unit_entries():
    """7-point Poisson on 5^3 (6 on the diagonal, -1 off it) with unknown 7 eliminated as a
    Dirichlet unit row and column, and two symmetric off-diagonal entries set to exactly 1."""
    M = H.poisson7(5).to_scipy().tolil()
    M[7, :] = 0.0
    M[:, 7] = 0.0
    M[7, 7] = 1.0
    M[20, 21] = 1.0
    M[21, 20] = 1.0
    return H.Csr.from_scipy(M.tocsr())


def _changed(A, Af):
    a, b = A.to_scipy().tocoo(), Af.to_scipy().tocoo()
    assert np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col)
    m = a.data != b.data
    return set(zip(a.row[m].tolist(), a.col[m].tolist())), b.data[m]


def test_bc_diag_rule_changes_unit_diagonal_only():
    A = _with_unit_entries()
    Af, nf = H.bc_fix(A, 1, "diag")
    where, vals = _changed(A, Af)
    assert nf == 1 and where == {(7, 7)}
    lmax = np.abs(np.linalg.eigvalsh(A.to_scipy().toarray())).max()
    np.testing.assert_allclose(vals, lmax / 10, rtol=1e-4)


def test_bc_all_rule_changes_every_unit_entry():
    A = _with_unit_entries()
    Af, nf = H.bc_fix(A, 1, "all")
    where, vals = _changed(A, Af)
    assert nf == 3 and where == {(7, 7), (20, 21), (21, 20)}
    lmax = np.abs(np.linalg.eigvalsh(A.to_scipy().toarray())).max()
    np.testing.assert_allclose(vals, lmax / 10, rtol=1e-4)
    # Entries that are not exactly 1 are untouched, whichever rule.
    assert H.bc_fix(H.poisson7(4), 1, "all")[1] == 0


def test_bc_fix_independent_of_thread_count():
    A, _ = H.elasticity(6)  # clamped face: 3 * 7^2 unit rows
    outs = [H.bc_fix(A, 1, "diag", nthreads=nt) for nt in (1, 3, 8)]
    assert outs[0][1] == 3 * 7 * 7
    for Af, nf in outs[1:]:
        assert nf == outs[0][1] and np.array_equal(Af.val, outs[0][0].val)


def test_bc_fix_rejects_unknown_mode():
    with pytest.raises(KeyError):
        H.bc_fix(H.poisson7(3), 1, "rows")


def test_rhs_unit_sol_and_rand_sol_have_their_solutions():
    A = H.heterogeneous_diffusion(8, 1.0, 3)[0]
    S = A.to_scipy()
    for mode in ("unit_sol", "rand_sol"):
        b, x = H.rhs_build(A, mode, seed=5)
        # The solution the mode promises solves the system.
        np.testing.assert_allclose(sp.linalg.spsolve(S.tocsc(), b), x, rtol=1e-10)
    _, x = H.rhs_build(A, "unit_sol")
    assert np.all(x == 1.0)
    _, u = H.rhs_build(A, "rand_sol", seed=5)
    assert u.min() >= 0.0 and u.max() < 1.0 and u.std() > 0.2
    _, u2 = H.rhs_build(A, "rand_sol", seed=6)
    assert not np.array_equal(u, u2)


def test_rhs_unit_rhs_and_rhs_in():
    A = H.poisson7(4)
    b, x = H.rhs_build(A, "unit_rhs")
    assert x is None and np.all(b == 1.0) and b.shape == (64,)
    given = np.arange(64.0)
    b, _ = H.rhs_build(A, "rhs_in", rhs=given)
    assert np.array_equal(b, given)
    with pytest.raises(ValueError, match="Missing the expected right-hand side"):
        H.rhs_build(A, "rhs_in")
    with pytest.raises(ValueError, match="Wrong value for rhs_build"):
        H.rhs_build(A, "ones")


def test_rhs_rand_rhs_spans_six_decades():
    """rand_rhs = u .* 10.^(6 (v - 0.5)): the decade exponent log10(rhs / u) is uniform on
    [-3, 3) and independent of u."""
    A = H.poisson7(40)  # 64,000 entries
    b, _ = H.rhs_build(A, "rand_rhs", seed=9)
    u = H.rng_uniform(9, 30, A.nrows)
    e = np.log10(b / u)
    assert e.min() >= -3.0 and e.max() < 3.0
    assert scipy.stats.kstest((e + 3.0) / 6.0, "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(u, e)[0, 1]) < 0.02


def test_fsai_nan_diagonal_repair():
    """smoother.m:59-71: rows of the aFSAI factor whose diagonal is NaN become the single
    entry 1/sqrt(a_ii); the other rows are untouched.  Invariant kept: diag(G A G^T) = 1 on
    every row (the aFSAI closed form, and (1/sqrt(a_ii))^2 a_ii = 1 on a repaired row)."""
    A = H.heterogeneous_diffusion(5, 1.0, 4)[0]
    G, pd = H.afsai(A)
    assert pd
    bad = [3, 40, 77]
    v = G.val.copy()
    for i in bad:
        v[G.rowptr[i + 1] - 1] = np.nan  # the diagonal is each row's last entry
    Gn = H.Csr(G.nrows, G.ncols, G.rowptr, G.col, v)
    Gr, nf = H.fsai_nan_repair(A, Gn)
    assert nf == len(bad)
    a = A.to_scipy()
    for i in range(G.nrows):
        lo, hi = Gr.rowptr[i], Gr.rowptr[i + 1]
        if i in bad:
            assert list(Gr.col[lo:hi]) == [i] and Gr.val[lo] == 1.0 / np.sqrt(a[i, i])
        else:
            g0, g1 = G.rowptr[i], G.rowptr[i + 1]
            assert np.array_equal(Gr.col[lo:hi], G.col[g0:g1]) and np.array_equal(Gr.val[lo:hi], G.val[g0:g1])
    d = (Gr.to_scipy() @ a @ Gr.to_scipy().T).diagonal()
    np.testing.assert_allclose(d, 1.0, rtol=1e-12)
    # A factor without NaN comes back unchanged.
    G2, n2 = H.fsai_nan_repair(A, G)
    assert n2 == 0 and np.array_equal(G2.val, G.val)
