"""Hierarchy on-disk format: the reference's Ext_bin files (DumpChronos.cpp:27-45,
read_ExtBinMat.m:1-36, read_ExtBinVec.m:1-21).

Pins: the byte layout against a record packed by hand with struct (little-endian int64
header, 1-based int64 row/column and float64 value per entry, the extra ncols word only in
the rectangular header); round trips bit for bit; a hierarchy saved and loaded gives the
same V-cycle (oracle) bit for bit; malformed files are rejected.
"""
import struct

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from chronoslab_amd import hierarchy as H


def test_matrix_bytes_match_the_reference_layout(tmp_path):
    M = H.Csr.from_scipy(sp.csr_matrix(np.array([[0.0, 2.5, 0.0], [-1.0, 0.0, 4.0]])))
    p = tmp_path / "P.Ext_bin"
    H.write_ext_bin_mat(str(p), M, rect=True)
    expect = struct.pack("<qqq", 2, 3, 3) + struct.pack("<qqd", 1, 2, 2.5) + struct.pack("<qqd", 2, 1, -1.0) + \
        struct.pack("<qqd", 2, 3, 4.0)
    assert p.read_bytes() == expect
    S = H.Csr.from_scipy(sp.csr_matrix(np.array([[4.0, -1.0], [-1.0, 4.0]])))
    q = tmp_path / "A.Ext_bin"
    H.write_ext_bin_mat(str(q), S)
    assert q.read_bytes()[:16] == struct.pack("<qq", 2, 4)
    assert len(q.read_bytes()) == 16 + 4 * 24


def test_matrix_and_vector_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for shape, rect in (((300, 300), False), ((300, 120), True)):
        M = H.Csr.from_scipy(sp.random(*shape, density=0.03, random_state=1, format="csr"))
        p = str(tmp_path / f"m{shape[1]}.Ext_bin")
        H.write_ext_bin_mat(p, M, rect=rect)
        R = H.read_ext_bin_mat(p, rect=rect)
        assert (R.nrows, R.ncols) == shape
        np.testing.assert_array_equal(R.rowptr, M.rowptr)
        np.testing.assert_array_equal(R.col, M.col)
        np.testing.assert_array_equal(R.val, M.val)
    v = rng.standard_normal(77)
    H.write_ext_bin_vec(str(tmp_path / "v.Ext_bin"), v)
    np.testing.assert_array_equal(H.read_ext_bin_vec(str(tmp_path / "v.Ext_bin")), v)
    X = rng.standard_normal((50, 6))
    H.write_ext_bin_vec(str(tmp_path / "X.Ext_bin"), X)
    raw = (tmp_path / "X.Ext_bin").read_bytes()
    assert raw[:16] == struct.pack("<qq", 50, 6) and raw[16:24] == struct.pack("<d", X[0, 0])
    assert raw[24:32] == struct.pack("<d", X[0, 1])  # row by row, as DumpChronos writes TV0
    np.testing.assert_array_equal(H.read_ext_bin_vec(str(tmp_path / "X.Ext_bin")), X)


def test_hierarchy_round_trip_same_vcycle(tmp_path):
    A, _ = H.heterogeneous_diffusion(14, 2.0, 9)
    hier = H.build(A)
    H.save_hierarchy(str(tmp_path / "h"), hier)
    back = H.load_hierarchy(str(tmp_path / "h"))
    assert back.nlev == hier.nlev
    for a, b in zip(hier.levels, back.levels):
        assert a.omega == b.omega and a.nupre == b.nupre and a.nupost == b.nupost
        for f in ("A", "P", "G"):
            x, y = getattr(a, f), getattr(b, f)
            if x is None:
                assert y is None
                continue
            assert (x.nrows, x.ncols) == (y.nrows, y.ncols)
            np.testing.assert_array_equal(x.rowptr, y.rowptr)
            np.testing.assert_array_equal(x.col, y.col)
            np.testing.assert_array_equal(x.val, y.val)
    r = H.rng_uniform(3, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(oracle.Hier(back).vcycle(r), oracle.Hier(hier).vcycle(r))


def test_malformed_files_rejected(tmp_path):
    p = tmp_path / "bad.Ext_bin"
    p.write_bytes(struct.pack("<qq", 2, 1) + struct.pack("<qqd", 3, 1, 1.0))  # row 3 of 2
    with pytest.raises(ValueError):
        H.read_ext_bin_mat(str(p))
    p.write_bytes(struct.pack("<qq", 2, 2) + struct.pack("<qqd", 1, 1, 1.0) * 2)  # duplicate
    with pytest.raises(ValueError):
        H.read_ext_bin_mat(str(p))
    p.write_bytes(struct.pack("<qq", 2, 3) + struct.pack("<qqd", 1, 1, 1.0))  # truncated
    with pytest.raises(ValueError):
        H.read_ext_bin_mat(str(p))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reloaded_hierarchy_runs_on_the_gpu(gpu, tmp_path, prec):
    """A 32^3 hierarchy built once, stored in Ext_bin files and reloaded (the cross-run reuse
    the format serves, DumpChronos.cpp:27-45): the HIP V-cycle and PCG_CJ on the reloaded
    hierarchy against the oracle on the original."""
    A = H.poisson7(32)
    hier = H.build(A)
    H.save_hierarchy(str(tmp_path), hier)
    again = H.load_hierarchy(str(tmp_path))
    ref = oracle.Hier(hier)
    b = H.rng_uniform(1, 2, A.nrows, -1.0, 1.0)
    M = gpu.AMGPreconditioner(again, prec)
    x, xo = M(b), ref.vcycle(b)
    assert np.linalg.norm(x - xo) <= (1e-12 if prec == "fp64" else 1e-5) * np.linalg.norm(xo)
    xs, it, rv, fl = M.pcg(b, itmax=100, tol=1e-8)
    xr, itr, rvr, flr = ref.pcg(A, b, itmax=100, tol=1e-8)
    assert fl == 0 and flr == 0
    if prec == "fp64":
        assert it == itr
        assert np.linalg.norm(xs - xr) <= 1e-10 * np.linalg.norm(xr)
    else:
        assert abs(it - itr) <= 1
        assert np.linalg.norm(b - A.matvec(xs)) <= 1e-8 * np.linalg.norm(b)
