"""Host checks of the symmetric diagonal layout used by the GPU library's single-vector
products and its PCG operator product.

chs_dia_apply builds the layout exactly as the GPU library does (dia_layout.hpp) and replays
spmv_dia_kernel's index arithmetic and summation order on the CPU, failing on any access
outside the stored diagonals.  Its product must equal the plain sequential CSR product bit
for bit: in FP64 the oracle's spmv, in FP32 a loop with FP32 products and sums.  Matrices
that are not candidates (unsymmetric values, stored zeros, too many diagonals, unsorted
rows, rectangular or triangular) must be declined."""
import ctypes as C

import numpy as np
import scipy.sparse as sp
import pytest

import oracle
from chronoslab_amd import _lib
from chronoslab_amd import hierarchy as H
from test_sell_layout import csr_seq_fp32

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


def dia_apply(A: H.Csr, x: np.ndarray, fp32: bool):
    lib = _lib.load_setup_lib()
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.col, dtype=np.int32)
    v = np.ascontiguousarray(A.val, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.full(A.nrows, np.nan)
    m = C.c_int32(0)
    offs = np.zeros(16, np.int32)
    rc = lib.chs_dia_apply(A.nrows, A.ncols, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P),
                           v.ctypes.data_as(_F64P), int(fp32), xx.ctypes.data_as(_F64P),
                           y.ctypes.data_as(_F64P), C.byref(m), offs.ctypes.data_as(_I32P))
    return rc, y, list(offs[: m.value])


def stencil27(N: int, rng) -> H.Csr:
    """Symmetric matrix with the 27-point pattern of an N^3 grid (14 stored diagonals) and
    random values, diagonally dominant."""
    idx = np.arange(N ** 3).reshape(N, N, N)
    rows, cols = [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dz, dy, dx) <= (0, 0, 0):
                    continue
                a = idx[max(0, -dz):N - max(0, dz), max(0, -dy):N - max(0, dy), max(0, -dx):N - max(0, dx)]
                b = idx[max(0, dz):N - max(0, -dz), max(0, dy):N - max(0, -dy), max(0, dx):N - max(0, -dx)]
                rows.append(a.ravel())
                cols.append(b.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = -rng.uniform(0.1, 1.0, r.size)
    U = sp.coo_matrix((v, (r, c)), shape=(N ** 3, N ** 3))
    M = (U + U.T).tocsr()
    M = M + sp.diags(-np.asarray(M.sum(axis=1)).ravel() + 1.0)
    return H.Csr.from_scipy(M)


def _check_equal(A: H.Csr, seed: int = 1):
    x = np.random.default_rng(seed).standard_normal(A.ncols)
    rc, y, offs = dia_apply(A, x, fp32=False)
    assert rc == 0
    np.testing.assert_array_equal(y, oracle.spmv(A, x))
    rc, y32, _ = dia_apply(A, x, fp32=True)
    assert rc == 0
    np.testing.assert_array_equal(y32, csr_seq_fp32(A, x))
    return offs


def test_poisson_stencil_offsets_and_bits():
    N = 12
    offs = _check_equal(H.poisson7(N))
    assert offs == [0, 1, N, N * N]


def test_heterogeneous_diffusion_with_boundary_rows():
    A, _ = H.heterogeneous_diffusion(14, 2.0, seed=3)
    assert _check_equal(A) == [0, 1, 14, 196]
    # the boundary rule changes diagonals only: still a candidate, same bits
    A2, _ = H.bc_fix(A)
    _check_equal(A2, seed=2)


def test_27_point_pattern():
    A = stencil27(9, np.random.default_rng(4))
    offs = _check_equal(A)
    assert len(offs) == 14 and offs[0] == 0


def test_diagonal_matrix_and_one_row():
    D = H.Csr.from_scipy(sp.diags(np.arange(1.0, 8.0)).tocsr())
    assert _check_equal(D) == [0]
    one = H.Csr(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5]))
    assert _check_equal(one) == [0]


def test_missing_main_diagonal_entries():
    """Rows without a diagonal entry (the slot holds a zero, skipped): same bits."""
    M = sp.lil_matrix((40, 40))
    rng = np.random.default_rng(5)
    for i in range(40):
        if i % 3:
            M[i, i] = rng.uniform(1, 2)
        if i + 5 < 40:
            v = rng.uniform(-1, 1)
            M[i, i + 5] = v
            M[i + 5, i] = v
    _check_equal(H.Csr.from_scipy(M.tocsr()))


def test_declined_matrices():
    x = np.ones(64)
    A = H.poisson7(4)
    # one value differing from its mirror in the last bit
    B = H.Csr(A.nrows, A.ncols, A.rowptr.copy(), A.col.copy(), A.val.copy())
    e = int(np.nonzero(B.col > np.repeat(np.arange(64), np.diff(B.rowptr)))[0][0])
    B.val[e] = np.nextafter(B.val[e], 0.0)
    assert dia_apply(B, x, False)[0] == 1
    # an explicitly stored zero
    Z = H.Csr(A.nrows, A.ncols, A.rowptr.copy(), A.col.copy(), A.val.copy())
    Z.val[e] = 0.0
    assert dia_apply(Z, x, False)[0] == 1
    # unsorted row
    U = H.Csr(A.nrows, A.ncols, A.rowptr.copy(), A.col.copy(), A.val.copy())
    a, b = U.rowptr[10], U.rowptr[10] + 1
    U.col[[a, b]] = U.col[[b, a]]
    U.val[[a, b]] = U.val[[b, a]]
    assert dia_apply(U, x, False)[0] == 1
    # lower triangular (the FSAI factor's shape)
    L = H.Csr.from_scipy(sp.tril(A.to_scipy()).tocsr())
    assert dia_apply(L, x, False)[0] == 1
    # rectangular
    R = H.Csr.from_scipy(A.to_scipy()[:, :60].tocsr())
    assert dia_apply(R, np.ones(60), False)[0] == 1
    # more than 16 stored diagonals
    rng = np.random.default_rng(6)
    n = 200
    M = sp.lil_matrix((n, n))
    for d in range(1, 18):
        for i in range(n - d):
            v = rng.uniform(-1, 1)
            M[i, i + d] = v
            M[i + d, i] = v
    M.setdiag(20.0)
    assert dia_apply(H.Csr.from_scipy(M.tocsr()), np.ones(n), False)[0] == 1
