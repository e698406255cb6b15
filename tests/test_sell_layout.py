"""Host checks of the sliced-ELL layout used by the GPU library's single-vector products.

chs_sell_apply builds the layout exactly as the GPU library does (sell_layout.hpp) and
replays spmv_sell_kernel's index arithmetic on the CPU, failing on any access outside the
layout's buffers and on any row not written exactly once.  Its product must equal the
plain sequential CSR product bit for bit: in FP64 the oracle's spmv, in FP32 a loop with
FP32 products and sums."""
import ctypes as C

import numpy as np
import scipy.sparse as sp
import pytest

import oracle
from chronoslab_amd import _lib
from chronoslab_amd import hierarchy as H

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


def sell_apply(A: H.Csr, x: np.ndarray, fp32: bool):
    lib = _lib.load_setup_lib()
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.col, dtype=np.int32)
    v = np.ascontiguousarray(A.val, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.full(A.nrows, np.nan)
    info = np.zeros(6, np.int32)
    elems = C.c_int64()
    rc = lib.chs_sell_apply(A.nrows, A.ncols, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P),
                            v.ctypes.data_as(_F64P), int(fp32), xx.ctypes.data_as(_F64P),
                            y.ctypes.data_as(_F64P), info.ctypes.data_as(_I32P), C.byref(elems))
    return rc, y, dict(cb=int(info[0]), sorted=bool(info[1]), wshift=int(info[2]),
                       nslices=int(info[3]), sshift=int(info[4]), ulen=int(info[5]),
                       elems=elems.value)


def csr_seq_fp32(A: H.Csr, x: np.ndarray) -> np.ndarray:
    v = A.val.astype(np.float32)
    xf = x.astype(np.float32)
    y = np.zeros(A.nrows)
    for i in range(A.nrows):
        s = np.float32(0.0)
        for e in range(A.rowptr[i], A.rowptr[i + 1]):
            s = np.float32(s + np.float32(v[e] * xf[A.col[e]]))
        y[i] = float(s)
    return y


def random_csr(n, rowlen, offsets, rng, empty_frac=0.0) -> H.Csr:
    """Square CSR: row i holds up to rowlen(i) distinct columns i - d, d drawn from offsets."""
    rp = [0]
    cols, vals = [], []
    for i in range(n):
        if rng.random() < empty_frac:
            rp.append(rp[-1])
            continue
        cand = [i - d for d in offsets if 0 <= i - d < n]
        k = min(len(cand), rowlen(i))
        c = np.sort(rng.choice(cand, size=k, replace=False)) if k else np.zeros(0, int)
        cols.extend(c.tolist())
        vals.extend(rng.standard_normal(k).tolist())
        rp.append(rp[-1] + k)
    return H.Csr(n, n, np.asarray(rp, np.int32), np.asarray(cols, np.int32), np.asarray(vals))


def random_csr_fast(n, k, offsets, rng) -> H.Csr:
    """Square CSR with up to k columns i - d per row, d drawn from offsets (vectorised)."""
    offs = np.asarray(offsets, np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = rows - offs[rng.integers(0, len(offs), size=rows.size)]
    keep = (cols >= 0) & (cols < n)
    key = np.unique(rows[keep] * n + cols[keep])
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    return H.Csr(n, n, rp.astype(np.int32), c.astype(np.int32), rng.standard_normal(c.size))


def check(A, rng, fp32_exact=True, expect_cb=None, allow_skip=False):
    x = rng.standard_normal(A.ncols)
    rc, y, info = sell_apply(A, x, fp32=False)
    if allow_skip and rc == 1:  # padded by more than half: left to the stream layout
        return None
    assert rc == 0, f"replay failed with {rc}"
    np.testing.assert_array_equal(y, oracle.spmv(A, x))
    if expect_cb is not None:
        assert info["cb"] == expect_cb, info
    rc, y32, info32 = sell_apply(A, x, fp32=True)
    assert rc == 0
    if fp32_exact:
        np.testing.assert_array_equal(y32, csr_seq_fp32(A, x))
    return info, info32


def test_stencil_dictionary_codes():
    # 7-point Poisson: 7 offsets, 8-bit dictionary codes; slices hold a multiple of four
    # elements per lane, so rows of 7 take 8 slots.
    A = H.poisson7(12)
    info, _ = check(A, np.random.default_rng(0), expect_cb=8)
    assert info["elems"] <= 1.25 * A.nnz


def test_uniform_ell(monkeypatch):
    # Unsorted layout of a matrix whose slices are all as long as its longest row: uniform
    # ELL (computed slice offsets).  Rows of one length plus an empty last slice's worth of
    # short rows exercise the zero-length and clamped loads.
    monkeypatch.setenv("CHRONOS_SELL_SORT", "0")
    A = H.poisson7(16)
    info, _ = check(A, np.random.default_rng(5), expect_cb=8)
    assert info["ulen"] == 8 and not info["sorted"] and info["elems"] == info["nslices"] * 8 * 64


def test_fsai_factor_and_transpose():
    A, _ = H.heterogeneous_diffusion(10, 2.0, seed=3)
    hier = H.build(A)
    rng = np.random.default_rng(1)
    for lev in hier.levels[:-1]:
        for M in (lev.A, lev.G, H.transpose(lev.G)):
            check(M, rng)


def test_ragged_rows_sorted_windows():
    # Rows of length 0..40 in a random order: sorting inside windows pays for its perm.
    rng = np.random.default_rng(2)
    lens = rng.integers(0, 41, size=3000)
    A = random_csr(3000, lambda i: int(lens[i]), list(range(-60, 60)), rng)
    info, _ = check(A, rng)
    assert info["sorted"]


@pytest.mark.parametrize("sort", ["0", "8", "10", "12", "14"])
def test_sort_windows(monkeypatch, sort):
    # Ragged rows over several dictionary windows, with each sort window forced in turn:
    # rows are permuted only inside their sort window, and every row is written once.
    rng = np.random.default_rng(7)
    n = 40000
    A = random_csr_fast(n, 12, list(range(-300, 300, 7)), rng)
    keep = rng.random(A.nnz) < 0.85  # ragged rows (padded by less than half unsorted)
    M = A.to_scipy().tocoo()
    A = H.Csr.from_scipy(sp.csr_matrix((M.data[keep], (M.row[keep], M.col[keep])), shape=(n, n)))
    monkeypatch.setenv("CHRONOS_SELL_SORT", sort)
    info, _ = check(A, rng, fp32_exact=False)
    assert info["wshift"] == 14
    if sort == "0":
        assert not info["sorted"]
    else:
        assert info["sorted"] and info["sshift"] == int(sort)


def test_default_sort_window_choice():
    rng = np.random.default_rng(8)
    lens = rng.integers(0, 41, size=20000)
    A = random_csr(20000, lambda i: int(lens[i]), list(range(-60, 60)), rng)
    info, _ = check(A, rng)
    assert info["sorted"] and 8 <= info["sshift"] <= info["wshift"]


def test_sizes_not_multiple_of_slice_and_tiny():
    rng = np.random.default_rng(3)
    for n in (1, 5, 63, 64, 65, 255, 257, 1000):
        A = random_csr(n, lambda i: 1 + i % 9, list(range(-20, 21)), rng, empty_frac=0.1)
        if A.nnz == 0:
            continue
        check(A, rng, allow_skip=n < 512)


def test_many_windows_and_16bit_offsets():
    # More than 255 distinct offsets per window, range below 65535: 16-bit codes.  n spans
    # several 16384-row windows.
    rng = np.random.default_rng(4)
    offs = list(range(-600, 601, 3))
    A = random_csr_fast(40000, 6, offs, rng)
    check(A, rng, fp32_exact=False, expect_cb=16)


def test_wide_offsets_full_columns():
    # Offsets spanning more than 65535 with many distinct values: 32-bit columns.
    rng = np.random.default_rng(5)
    n = 140000
    offs = rng.choice(np.arange(-69999, 70000), size=2000, replace=False)
    A = random_csr_fast(n, 12, offs, rng)
    check(A, rng, fp32_exact=False, expect_cb=32)


def test_not_candidates():
    rng = np.random.default_rng(6)
    # Rectangular.
    P = H.Csr(5, 3, np.array([0, 1, 2, 3, 4, 5], np.int32), np.array([0, 1, 2, 0, 1], np.int32), np.ones(5))
    assert sell_apply(P, np.ones(3), False)[0] == 1
    # A row longer than 256.
    n = 600
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = 1
    rp[1] = 300
    rp = np.cumsum(np.r_[0, [300] + [1] * (n - 1)]).astype(np.int32)
    cols = np.r_[np.arange(300), np.arange(1, n)].astype(np.int32)
    A = H.Csr(n, n, rp, cols, rng.standard_normal(len(cols)))
    assert sell_apply(A, np.ones(n), False)[0] == 1
    # Empty.
    E = H.Csr(4, 4, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert sell_apply(E, np.ones(4), False)[0] == 1
