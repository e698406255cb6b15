"""Energy-minimising prolongation (EMIN): the pattern of MEX_cptProlPatt.cpp and the projected
Jacobi-PCG of EMIN_matfree.cpp / DEFL_PCG_matfree.cpp (solv_emin = 4, prec_emin = 1).

Pinned by
- the reference's own known answer, Tests/CubeBAMG_EMIN.m (SQMR to 1e-6 in 20 iterations on
  examples/matrices/Cubo_4820.mat, fixture tests/golden/cubo4820.npz), under its acceptance
  rule, for several PMIS seeds;
- the pattern against a dense restatement: the candidate weights |(S P)_ij| from a dense
  product with S = |a_ij| / sqrt(a_ii a_jj) above patt_tau, the budget largest kept;
- the minimisation against a dense KKT solve of min trace(P^T A P) over the pattern under
  each fine row's test-space constraint (numpy least squares and SVD for the constraint
  blocks): with many iterations and no energy stop, the CG lands on the KKT solution;
- invariants: full-rank rows reproduce the test space exactly, the energy decreases, coarse
  rows stay e_c, the result does not depend on the thread count.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from chronoslab_amd import amg, hierarchy as H
from test_bamg import cube_params, cubo, rule


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cube_emin_known_answer(seed):
    As, TV0, rhs = cubo()
    Afix, nfix = H.bc_fix(H.Csr.from_scipy(As), 1)  # computeAMG.m:5-10
    hier = H.build(Afix, H.SetupParams(tv=TV0, **cube_params(seed, "emin")))
    x, info, relres, niter, rv = amg.SQMR(lambda v: As @ v, rhs, 1e-6, 100, oracle.Hier(hier).vcycle)
    print(f"emin seed {seed}: levels {[l.A.nrows for l in hier.levels]}, SQMR {niter} (reference 20)")
    assert info == 0 and rule(20, niter), niter


def _level(ne=4, seed=5):
    """Finest level of the elasticity set-up: A, strength, C/F split, BAMG P, test space."""
    A, _ = H.elasticity(ne)
    A, _ = H.bc_fix(A, 1)
    strong = H.soc_dom(A, 0.01)
    rng = np.random.default_rng(seed)
    fc, nc = H.pmis(A, strong, rng.random(A.nrows))
    tv = H.orthonormalize(H.rigid_body_modes(ne))
    P, _ = H.prol_bamg(A, strong, fc, nc, tv)
    return A, fc, nc, tv, P


def _dense_pattern(A, P, fc, ntv, tau, nnzr_max):
    Ad = A.to_scipy().toarray()
    d = np.diag(Ad)
    S = np.abs(Ad) / np.sqrt(np.outer(d, d))
    S[(S <= tau) | (Ad == 0)] = 0.0
    Pd = P.to_scipy().toarray()
    W = S @ Pd
    reach = ((S != 0).astype(float) @ (Pd != 0).astype(float)) > 0
    nnzr = max(nnzr_max, ntv)
    rows = []
    for i in range(A.nrows):
        if fc[i] >= 0:
            rows.append((None, None, None, None))
            continue
        keep = np.flatnonzero(Pd[i] != 0)
        # weights at rounding level (a cancelling sum) may come out zero or not in either order
        tiny = 1e-13 * np.abs(W[i]).max()
        sure = np.setdiff1d(np.flatnonzero(np.abs(W[i]) > tiny), keep)
        amb = np.setdiff1d(np.flatnonzero(reach[i] & (np.abs(W[i]) <= tiny)), keep)
        rows.append((set(keep.tolist()), set(sure.tolist()), set(amb.tolist()), W[i]))
    return rows, nnzr


@pytest.mark.parametrize("nnzr_max", [30, 12])
def test_pattern_against_dense_product(nnzr_max):
    A, fc, nc, tv, P = _level()
    prm = H.SetupParams(nnzr_max=nnzr_max)
    T = H.emin_pattern(A, P, fc, tv.shape[1], prm).to_scipy()
    rows, nnzr = _dense_pattern(A, P, fc, tv.shape[1], prm.patt_tau, nnzr_max)
    grown = 0
    for i, (keep, sure, amb, w) in enumerate(rows):
        got = set(T.indices[T.indptr[i]:T.indptr[i + 1]].tolist())
        if keep is None:
            assert not got
            continue
        assert keep <= got
        extra = got - keep
        budget = max(0, nnzr - len(keep))
        assert extra <= sure | amb
        assert len(extra & sure) == min(budget, len(sure))
        assert len(extra) <= min(budget, len(sure) + len(amb))
        if len(sure) > budget > 0:
            # every chosen weight at least every rejected one (up to rounding of the product)
            chosen = np.abs(w[sorted(extra)])
            rest = np.abs(w[sorted(sure - extra)])
            assert chosen.min() >= rest.max() * (1 - 1e-12)
        grown += len(extra)
    assert grown > 0
    # patt_pow = 1: the pattern is P's fine rows.
    T1 = H.emin_pattern(A, P, fc, tv.shape[1], H.SetupParams(patt_pow=1)).to_scipy()
    Pf = P.to_scipy().tolil()
    for i in np.flatnonzero(fc >= 0):
        Pf.rows[i] = []
        Pf.data[i] = []
    assert (abs(T1 - (Pf.tocsr() != 0)) > 0).nnz == 0


def _kkt_reference(A, P, T, fc, tv, condmax):
    """Dense minimiser of trace(P^T A P) over T's entries under the row constraints."""
    Ad = A.to_scipy().toarray()
    n, nc = P.nrows, P.ncols
    Pd = P.to_scipy().toarray()
    c2g = np.empty(nc, int)
    c2g[fc[fc >= 0]] = np.flatnonzero(fc >= 0)
    ent = [(i, j) for i in range(n) for j in T.indices[T.indptr[i]:T.indptr[i + 1]]]
    nt = len(ent)
    x0 = np.array([Pd[i, j] for i, j in ent])
    cons, rhs = [], []
    for i in np.flatnonzero(fc < 0):
        ks = np.arange(T.indptr[i], T.indptr[i + 1])
        if len(ks) == 0:
            continue
        B = tv[c2g[T.indices[ks]], :]  # nr x ntv
        U, s, Vt = np.linalg.svd(B, full_matrices=False)
        r = 0
        if s[0] > 1e-15:
            r = 1
            while r < len(s) and s[0] < condmax * s[r]:
                r += 1
        # least-squares correction onto the constraint, then Q^T x fixed there
        g = tv[i] - B.T @ x0[ks]
        x0[ks] += U[:, :r] @ ((Vt[:r] @ g) / s[:r])
        for q in range(r):
            row = np.zeros(nt)
            row[ks] = U[:, q]
            cons.append(row)
            rhs.append(U[:, q] @ x0[ks])
    # energy x^T K x + 2 f.x with K = A (x) I on the pattern, f = A(i, c(j))
    idx = {e: k for k, e in enumerate(ent)}
    K = np.zeros((nt, nt))
    f = np.zeros(nt)
    for k, (i, j) in enumerate(ent):
        f[k] = Ad[i, c2g[j]]
        for m in np.flatnonzero(Ad[i]):
            kk = idx.get((m, j))
            if kk is not None:
                K[k, kk] = Ad[i, m]
    C = np.array(cons)
    KKT = np.block([[2 * K, C.T], [C, np.zeros((len(C), len(C)))]])
    sol = np.linalg.lstsq(KKT, np.concatenate([-2 * f, rhs]), rcond=None)[0]
    energy = lambda x: x @ K @ x + 2 * f @ x
    return ent, x0, sol[:nt], energy


def _values(Pm, ent):
    Pd = Pm.to_scipy().toarray()
    return np.array([Pd[i, j] for i, j in ent])


def test_svd_rows_reproduce_least_squares():
    # patt_pow = 1: the pattern is BAMG's P, at most ntv entries per row, so every row's block
    # goes to the rank-revealing SVD branch; a full-row-rank block leaves no freedom and the
    # result is the least-squares correction of P itself.
    A, fc, nc, tv, P = _level(ne=3, seed=2)
    prm = H.SetupParams(patt_pow=1)
    T = H.emin_pattern(A, P, fc, tv.shape[1], prm).to_scipy()
    assert np.diff(T.indptr).max() <= tv.shape[1]
    ent, x0, xk, energy = _kkt_reference(A, P, T, fc, tv, prm.condmax_emin)
    Pe, it = H.emin_improve(A, P, H.Csr.from_scipy(T), fc, tv, prm)
    x = _values(Pe, ent)
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-12 * np.abs(x0).max())
    np.testing.assert_allclose(x, xk, rtol=0, atol=1e-9 * np.abs(xk).max())


def test_minimiser_matches_dense_kkt():
    A, fc, nc, tv, P = _level(ne=3, seed=2)
    T = H.emin_pattern(A, P, fc, tv.shape[1]).to_scipy()
    assert T.nnz > P.nnz - nc  # the pattern grew
    ent, x0, xk, energy = _kkt_reference(A, P, T, fc, tv, 1e10)
    # many iterations, no energy stop: the projected CG converges to the constrained minimum
    Pe, it = H.emin_improve(A, P, H.Csr.from_scipy(T), fc, tv, H.SetupParams(itmax_emin=2000, energ_tol=0.0))
    x = _values(Pe, ent)
    print(f"CG {it}, |x - x_kkt| / |x_kkt| = {np.linalg.norm(x - xk) / np.linalg.norm(xk):.2e}")
    assert np.linalg.norm(x - xk) <= 1e-7 * np.linalg.norm(xk)
    assert abs(energy(x) - energy(xk)) <= 1e-10 * abs(energy(xk))
    # the reference's stopping rule (10 iterations, 1e-3 energy ratio) lowers the energy from
    # the constraint-corrected start
    P10, it10 = H.emin_improve(A, P, H.Csr.from_scipy(T), fc, tv)
    x10 = _values(P10, ent)
    assert 1 <= it10 <= 10
    assert energy(xk) <= energy(x10) <= energy(x0)


def test_constraint_and_coarse_rows():
    A, fc, nc, tv, P = _level(ne=5, seed=9)
    T = H.emin_pattern(A, P, fc, tv.shape[1])
    Pe, _ = H.emin_improve(A, P, T, fc, tv)
    Pd = Pe.to_scipy().toarray()
    c2g = np.empty(nc, int)
    c2g[fc[fc >= 0]] = np.flatnonzero(fc >= 0)
    tvc = tv[c2g]
    Tc = T.to_scipy()
    exact = 0
    for i in range(A.nrows):
        if fc[i] >= 0:
            assert np.array_equal(np.flatnonzero(Pd[i]), [fc[i]]) and Pd[i, fc[i]] == 1.0
            continue
        cols = Tc.indices[Tc.indptr[i]:Tc.indptr[i + 1]]
        assert set(np.flatnonzero(Pd[i]).tolist()) <= set(cols.tolist())
        B = tvc[cols]
        if len(cols) > tv.shape[1] and np.linalg.cond(B) < 1e8:
            np.testing.assert_allclose(Pd[i] @ tvc, tv[i], atol=1e-10)
            exact += 1
    assert exact > 0


def test_emin_independent_of_thread_count():
    A, fc, nc, tv, P = _level()
    T = H.emin_pattern(A, P, fc, tv.shape[1], nthreads=1)
    out = [H.emin_improve(A, P, T, fc, tv, nthreads=nt)[0].to_scipy() for nt in (1, 3, 8)]
    for T2 in (H.emin_pattern(A, P, fc, tv.shape[1], nthreads=nt).to_scipy() for nt in (3, 8)):
        assert (abs(T2 - T.to_scipy()) > 0).nnz == 0
    for o in out[1:]:
        assert (o != out[0]).nnz == 0


def test_emin_needs_the_bamg_test_space():
    A, _ = H.elasticity(4)
    A, _ = H.bc_fix(A, 1)
    with pytest.raises(Exception):
        H.build(A, H.SetupParams(prolong="exti", prol_emin="emin", nlev_max=2, max_coarse=50))


def _ops_struct(open_fn, apply_fn, close_fn):
    import ctypes as C

    OPEN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                       C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_void_p))
    APPLY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))
    CLOSE = C.CFUNCTYPE(None, C.c_void_p)
    cbs = (OPEN(open_fn), APPLY(apply_fn), CLOSE(close_fn))
    addr = lambda f: C.cast(f, C.c_void_p).value
    return H._CEminOps(addr(cbs[0]), addr(cbs[1]), addr(cbs[2]), None), cbs


def _improve_with_ops(A, P, T, fc, tv, ops):
    import ctypes as C

    ha, hp, ht = H._to_handle(A), H._to_handle(P), H._to_handle(T)
    fc = np.ascontiguousarray(fc, np.int32)
    tvf = np.asfortranarray(tv, dtype=np.float64)
    ep = H._cemin(H.SetupParams())
    out, it = C.c_void_p(), C.c_int32()
    rc = H._lib().chs_emin_improve_ops(ha.ptr, hp.ptr, ht.ptr, fc.ctypes.data_as(H._I32P), tvf.shape[1],
                                       tvf.ctypes.data_as(H._F64P), C.cast(C.pointer(ep), C.c_void_p),
                                       C.cast(C.pointer(ops), C.c_void_p), 0, C.byref(out), C.byref(it))
    return rc, (H._take(out) if rc == 0 else None), it.value


def test_operator_hook():
    """The CG-operator hook of chs_build (chs_emin_kp_ops): a declining open falls back to the
    host product (same bits), an external product (here numpy's, same operator, other
    summation order) gives the same P to rounding, and a failing apply is an error."""
    import ctypes as C

    A, fc, nc, tv, P = _level(ne=4, seed=5)
    T = H.emin_pattern(A, P, fc, tv.shape[1])
    Ph, ith = H.emin_improve(A, P, T, fc, tv)
    calls = {"open": 0, "apply": 0, "close": 0}
    state = {}

    def open_decline(ctx, n, ncc, rpA, colA, valA, rpT, colT, handle):
        calls["open"] += 1
        return 1

    def open_np(ctx, n, ncc, rpA, colA, valA, rpT, colT, handle):
        calls["open"] += 1
        nnz, nt = rpA[n], rpT[n]
        Am = sp.csr_matrix((np.ctypeslib.as_array(valA, (nnz,)).copy(), np.ctypeslib.as_array(colA, (nnz,)).copy(),
                            np.ctypeslib.as_array(rpA, (n + 1,)).copy()), shape=(n, n))
        tr, tc = np.ctypeslib.as_array(rpT, (n + 1,)).copy(), np.ctypeslib.as_array(colT, (nt,)).copy()
        state["A"], state["rows"], state["cols"], state["shape"] = Am, np.repeat(np.arange(n), np.diff(tr)), tc, (n, ncc)
        handle[0] = C.c_void_p(1)
        return 0

    def apply_np(h, x, y):
        calls["apply"] += 1
        nt = len(state["cols"])
        X = sp.csr_matrix((np.ctypeslib.as_array(x, (nt,)).copy(), (state["rows"], state["cols"])), shape=state["shape"])
        Y = (state["A"] @ X).tocsr()
        np.ctypeslib.as_array(y, (nt,))[:] = np.asarray(Y[state["rows"], state["cols"]]).ravel()
        return 0

    def apply_fail(h, x, y):
        calls["apply"] += 1
        return -1

    def close(h):
        calls["close"] += 1

    ops, keep = _ops_struct(open_decline, apply_np, close)
    rc, Pd, itd = _improve_with_ops(A, P, T, fc, tv, ops)
    assert rc == 0 and itd == ith and calls["open"] == 1 and calls["apply"] == 0 and calls["close"] == 0
    assert np.array_equal(Pd.rowptr, Ph.rowptr) and np.array_equal(Pd.col, Ph.col) and np.array_equal(Pd.val, Ph.val)
    ops, keep = _ops_struct(open_np, apply_np, close)
    rc, Pn, itn = _improve_with_ops(A, P, T, fc, tv, ops)
    assert rc == 0 and itn == ith and calls["apply"] == ith + 1 and calls["close"] == 1
    assert np.array_equal(Pn.col, Ph.col)
    np.testing.assert_allclose(Pn.val, Ph.val, rtol=0, atol=1e-12 * np.abs(Ph.val).max())
    ops, keep = _ops_struct(open_np, apply_fail, close)
    rc, _, _ = _improve_with_ops(A, P, T, fc, tv, ops)
    assert rc != 0 and calls["close"] == 2
