"""Host checks of the packed-code stream layout used by the GPU library's single-vector
products (pack_layout.hpp, spmv_pack_kernel).

chs_pack_apply builds the layout exactly as the GPU library does and replays the kernel's
index arithmetic on the CPU, failing on any access outside the layout's buffers, on a
column decoded out of range, and on any row not written exactly once.  For rows that fit a
block its product must equal the plain sequential CSR product bit for bit (FP64: the
oracle's spmv; FP32: FP32 products and sums).  A row longer than a block has more distinct
offsets than one dictionary holds, so such matrices are never packed."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from chronoslab_amd import _lib
from chronoslab_amd import hierarchy as H
from test_sell_layout import csr_seq_fp32, random_csr_fast

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


def pack_apply(A: H.Csr, x: np.ndarray, per: int, fp32: bool):
    lib = _lib.load_setup_lib()
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.col, dtype=np.int32)
    v = np.ascontiguousarray(A.val, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.full(A.nrows, np.nan)
    nblk = C.c_int32()
    rc = lib.chs_pack_apply(A.nrows, A.ncols, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P),
                            v.ctypes.data_as(_F64P), per, int(fp32), xx.ctypes.data_as(_F64P),
                            y.ctypes.data_as(_F64P), C.byref(nblk))
    return rc, y, nblk.value


def _fsai_mats():
    A, _ = H.heterogeneous_diffusion(14, 2.0, seed=6)
    hier = H.build(A)
    out = []
    for lev in hier.levels[:-1]:
        out += [lev.A, lev.G, H.transpose(lev.G)]
    return out


@pytest.mark.parametrize("per", [4, 8, 16])
def test_fsai_factors_and_transposes_bit_exact(per):
    rng = np.random.default_rng(per)
    for M in _fsai_mats():
        x = rng.standard_normal(M.ncols)
        rc, y, nblk = pack_apply(M, x, per, fp32=False)
        assert rc == 0, rc
        np.testing.assert_array_equal(y, oracle.spmv(M, x))
        rc, y32, _ = pack_apply(M, x, per, fp32=True)
        assert rc == 0
        np.testing.assert_array_equal(y32, csr_seq_fp32(M, x))


def test_wide_column_span_and_windows():
    # Offsets spread over +-100000 (beyond 16-bit block offsets) but few distinct per
    # window: 2-byte packed codes apply; several 16384-row windows and ragged rows.
    rng = np.random.default_rng(1)
    n = 60000
    offs = [-98304, -16384, -128, -1, 0, 1, 128, 16384, 98304] + list(range(-40, 41, 8))
    A = random_csr_fast(n, 12, offs, rng)
    x = rng.standard_normal(n)
    rc, y, nblk = pack_apply(A, x, 8, fp32=False)
    assert rc == 0 and nblk > n // 256
    np.testing.assert_array_equal(y, oracle.spmv(A, x))


def test_empty_rows_and_long_rows():
    rng = np.random.default_rng(2)
    n = 3000
    M = sp.lil_matrix((n, n))
    for i in range(0, n, 3):  # two of every three rows empty
        M[i, i] = 4.0
        if i + 1 < n:
            M[i, i + 1] = -1.0
    A = H.Csr.from_scipy(M.tocsr())
    x = rng.standard_normal(n)
    rc, y, _ = pack_apply(A, x, 4, fp32=False)
    assert rc == 0
    np.testing.assert_array_equal(y, oracle.spmv(A, x))
    # A row longer than a block has more distinct offsets than a dictionary holds: such a
    # matrix is never packed (it keeps the stream layout and its long-row path).
    M[5, :2500] = rng.standard_normal(2500)
    assert pack_apply(H.Csr.from_scipy(M.tocsr()), x, 4, fp32=False)[0] == 1


def test_not_candidates():
    rng = np.random.default_rng(3)
    P = H.Csr(5, 3, np.array([0, 1, 2, 3, 4, 5], np.int32), np.array([0, 1, 2, 0, 1], np.int32), np.ones(5))
    assert pack_apply(P, np.ones(3), 8, False)[0] == 1
    # More than 256 distinct row-relative offsets in one window.
    n = 2000
    A = random_csr_fast(n, 8, list(range(-1000, 1000)), rng)
    assert pack_apply(A, np.ones(n), 8, False)[0] == 1
