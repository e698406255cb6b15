"""The host dense routines of the GPU library's LOBPCG Rayleigh-Ritz steps (dense.hpp).

chs_sygv solves A v = lambda B v (symmetric A, SPD B) as MATLAB's eig(A, B) does in
sources/EigenSolver/lobpcg.m:99, :325 and :437; chs_cond_sym is cond of a symmetric matrix
(sources/EigenSolver/B_orthonormalize.m:9).  They are checked against LAPACK through scipy
and numpy (eigenvalues, condition numbers) and against what the mathematics fixes: the
residual A V - B V diag(w), B-orthonormality V^T B V = I, ascending order, repeated
eigenvalues, matrices that are already diagonal or tridiagonal, and a B that is not
positive definite."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

from chronoslab_amd import _lib

_F64P = C.POINTER(C.c_double)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def sygv(A, B):
    lib = _lib.load_setup_lib()
    m = A.shape[0]
    A, B = _f(A), _f(B)
    w = np.zeros(m)
    V = np.zeros((m, m), order="F")
    rc = lib.chs_sygv(m, A.ctypes.data_as(_F64P), B.ctypes.data_as(_F64P), w.ctypes.data_as(_F64P),
                      V.ctypes.data_as(_F64P))
    return rc, w, V


def cond_sym(M):
    M = _f(M)
    return _lib.load_setup_lib().chs_cond_sym(M.shape[0], M.ctypes.data_as(_F64P))


def _spd(rng, m, shift=1.0):
    X = rng.standard_normal((m, m))
    return X @ X.T / m + shift * np.eye(m)


def _sym(rng, m):
    X = rng.standard_normal((m, m))
    return 0.5 * (X + X.T)


def _check(A, B, w, V, tol=1e-11):
    m = A.shape[0]
    assert np.all(np.diff(w) >= 0)
    scale = max(np.linalg.norm(A, 2), 1e-300)
    res = A @ V - (B @ V) * w
    assert np.linalg.norm(res) <= tol * scale * max(np.linalg.norm(V), 1.0) * m
    assert np.allclose(V.T @ B @ V, np.eye(m), atol=tol * m)


@pytest.mark.parametrize("m", [1, 2, 3, 7, 16, 31, 48])
def test_sygv_against_lapack(m):
    rng = np.random.default_rng(100 + m)
    A, B = _sym(rng, m), _spd(rng, m)
    rc, w, V = sygv(A, B)
    assert rc == 0
    ref = sla.eigh(A, B, eigvals_only=True)
    assert np.allclose(w, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    _check(A, B, w, V)


def test_sygv_repeated_eigenvalues():
    """A cluster of equal eigenvalues (as the Rayleigh-Ritz blocks of symmetric grids have):
    the eigenvalues exact to rounding, the vectors still B-orthonormal with small residual."""
    rng = np.random.default_rng(7)
    m = 48
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    lam = np.repeat(np.arange(1.0, 17.0), 3)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    for B in (np.eye(m), _spd(rng, m, 2.0)):
        rc, w, V = sygv(A, B)
        assert rc == 0
        ref = sla.eigh(A, B, eigvals_only=True)
        assert np.allclose(w, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        _check(A, B, w, V)
    rc, w, _ = sygv(A, np.eye(m))
    assert np.allclose(w, lam, rtol=0, atol=1e-12 * lam.max())


def test_sygv_structured_inputs():
    rng = np.random.default_rng(3)
    m = 20
    d = rng.standard_normal(m)
    rc, w, V = sygv(np.diag(d), np.eye(m))  # already diagonal: no reflector, no rotation
    assert rc == 0
    assert np.array_equal(w, np.sort(d))
    assert np.allclose(np.abs(V), np.eye(m)[:, np.argsort(d, kind="stable")])
    T = np.diag(2.0 * np.ones(m)) - np.diag(np.ones(m - 1), 1) - np.diag(np.ones(m - 1), -1)
    rc, w, V = sygv(T, np.eye(m))  # 1-D Laplacian: closed-form eigenvalues
    k = np.arange(1, m + 1)
    assert np.allclose(w, 2.0 - 2.0 * np.cos(k * np.pi / (m + 1)), rtol=0, atol=1e-13)
    _check(T, np.eye(m), w, V)
    rc, w, V = sygv(np.zeros((m, m)), np.eye(m))
    assert rc == 0 and np.array_equal(w, np.zeros(m))
    assert np.allclose(V.T @ V, np.eye(m))
    for s in (1e-9, 1e9):  # scaled problems: eigenvalues scale with A / B
        A, B = _sym(rng, m), _spd(rng, m)
        rc, w, V = sygv(s * A, B)
        ref = sla.eigh(s * A, B, eigvals_only=True)
        assert rc == 0 and np.allclose(w, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_sygv_rejects_indefinite_b():
    rng = np.random.default_rng(5)
    m = 6
    A = _sym(rng, m)
    B = np.eye(m)
    B[4, 4] = -1.0
    assert sygv(A, B)[0] == 1
    assert sygv(A, np.zeros((m, m)))[0] == 1


@pytest.mark.parametrize("m", [1, 2, 5, 16, 48])
def test_cond_sym_against_numpy(m):
    rng = np.random.default_rng(40 + m)
    for M in (_sym(rng, m), _spd(rng, m, 1e-3)):
        assert np.isclose(cond_sym(M), np.linalg.cond(M, 2), rtol=1e-10)
    M = _spd(rng, m)
    if m > 1:
        u = rng.standard_normal(m)
        Ms = np.outer(u, u)  # rank one: singular for m > 1
        c = cond_sym(Ms)
        assert c == np.inf or c > 1e14
    assert cond_sym(np.eye(m)) == 1.0

