"""The symmetric diagonal layout on the GPU: the same bits as the other layouts and the oracle.

CHRONOS_DIA=1 forces the diagonal kernel for every single-vector product of a matrix that
admits it (a few diagonals, values symmetric bit for bit), CHRONOS_DIA=0 excludes it.  The
kernel sums each row in column order with rounded products, so products and V-cycles agree
bit for bit with the other layouts.  The PCG operator product runs on the diagonal layout
whenever the operator admits it; its dot products are then summed over 256-row blocks
instead of the stream layout's blocks, so the two PCG runs agree to rounding."""
import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H
from test_dia_layout import stencil27

pytestmark = pytest.mark.gpu


def _with_env(monkeypatch, env, make):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        return make()
    finally:
        for k in env:
            monkeypatch.delenv(k)


def _mats():
    A, _ = H.heterogeneous_diffusion(20, 2.0, seed=4)
    return [
        ("poisson20", H.poisson7(20)),
        ("diffusion20", A),
        ("diffusion20_bc", H.bc_fix(A)[0]),
        ("stencil27_12", stencil27(12, np.random.default_rng(3))),
    ]


@pytest.mark.parametrize("name,A", _mats())
def test_spmv_dia_equals_oracle(gpu, monkeypatch, name, A):
    x = np.random.default_rng(1).standard_normal(A.ncols)
    ref = oracle.spmv(A, x)
    for mode in ("1", "0"):
        M = _with_env(monkeypatch, {"CHRONOS_DIA": mode}, lambda: gpu.DeviceMatrix(A))
        print(f"{name} CHRONOS_DIA={mode}", flush=True)
        np.testing.assert_array_equal(M(x), ref)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_vcycle_dia_bit_identical_and_pcg(gpu, monkeypatch, prec):
    A, _ = H.heterogeneous_diffusion(24, 2.0, seed=5)
    hier = H.build(A)
    Md = _with_env(monkeypatch, {"CHRONOS_DIA": "1"}, lambda: gpu.AMGPreconditioner(hier, prec))
    Mc = _with_env(monkeypatch, {"CHRONOS_DIA": "0"}, lambda: gpu.AMGPreconditioner(hier, prec))
    ld, lc = Md.layout(), Mc.layout()
    print(prec, "level 0 A:", ld[0]["A"], "vs", lc[0]["A"], flush=True)
    assert ld[0]["A"][0] == "dia" and lc[0]["A"][0] != "dia"
    b = H.rng_uniform(6, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(Md(b), Mc(b))
    xd, itd, rvd, fd = Md.pcg(b, itmax=100, tol=1e-8)
    xc, itc, rvc, fc = Mc.pcg(b, itmax=100, tol=1e-8)
    print(prec, "iterations dia", itd, "stream", itc, flush=True)
    assert fd == fc == 0 and itd == itc
    assert np.linalg.norm(xd - xc) <= 1e-11 * np.linalg.norm(xc)
    ref = oracle.Hier(hier)
    xr, itr, _, _ = ref.pcg(hier.levels[0].A, b, itmax=100, tol=1e-8)
    assert abs(itd - itr) <= (0 if prec == "fp64" else 1)
    if prec == "fp64":
        assert np.linalg.norm(xd - xr) <= 1e-10 * np.linalg.norm(xr)
    true_res = np.linalg.norm(b - A.matvec(xd)) / np.linalg.norm(b)
    assert true_res <= 1e-8


def test_pcg_dia_operator_against_oracle_poisson(gpu):
    """Default tuning (the PCG product on the diagonal layout) on the 7-point Poisson matrix:
    equal iterations and the FP64 solution within 1e-10 of the oracle's."""
    A = H.poisson7(32)
    hier = H.build(A)
    M = gpu.AMGPreconditioner(hier, "fp64")
    b = H.rng_uniform(2, 2, A.nrows, -1, 1)
    x, it, rv, fl = M.pcg(b, itmax=100, tol=1e-8)
    xr, itr, rvr, flr = oracle.Hier(hier).pcg(hier.levels[0].A, b, itmax=100, tol=1e-8)
    print("poisson32 iterations", it, "oracle", itr, flush=True)
    assert fl == flr == 0 and it == itr
    np.testing.assert_allclose(rv, rvr, rtol=1e-7)
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)
