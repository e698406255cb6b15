"""The sliced-ELL layout on the GPU: the same bits as the CSR-stream layout and the oracle.

CHRONOS_SELL=1 forces the sliced-ELL kernel for every square level matrix that admits it,
CHRONOS_SELL=0 excludes it; both layouts sum each row in column order with rounded
products, so V-cycles, products and PCG runs agree bit for bit."""
import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H
from test_sell_layout import random_csr, random_csr_fast

pytestmark = pytest.mark.gpu


def _with_layout(monkeypatch, mode, make):
    monkeypatch.setenv("CHRONOS_SELL", mode)
    try:
        return make()
    finally:
        monkeypatch.delenv("CHRONOS_SELL")


def _mats():
    rng = np.random.default_rng(11)
    A, _ = H.heterogeneous_diffusion(20, 2.0, seed=4)
    hier = H.build(A)
    lens = rng.integers(0, 41, size=3000)
    return [
        ("poisson20", H.poisson7(20)),
        ("G0_diffusion20", hier.levels[0].G),
        ("GT0_diffusion20", H.transpose(hier.levels[0].G)),
        ("A1_diffusion20", hier.levels[1].A),
        ("ragged3000", random_csr(3000, lambda i: int(lens[i]), list(range(-60, 60)), rng)),
        ("off16_40000", random_csr_fast(40000, 6, list(range(-600, 601, 3)), rng)),
        ("col32_70000", random_csr_fast(70000, 4, rng.choice(np.arange(-69999, 70000), 2000, replace=False), rng)),
    ]


@pytest.mark.parametrize("name,A", _mats())
def test_spmv_layouts_equal_oracle(gpu, monkeypatch, name, A):
    x = np.random.default_rng(1).standard_normal(A.ncols)
    ref = oracle.spmv(A, x)
    for mode in ("1", "0"):
        M = _with_layout(monkeypatch, mode, lambda: gpu.DeviceMatrix(A))
        print(f"{name} CHRONOS_SELL={mode}", flush=True)
        np.testing.assert_array_equal(M(x), ref)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_vcycle_layouts_bit_identical(gpu, monkeypatch, prec):
    A, _ = H.heterogeneous_diffusion(24, 2.0, seed=5)
    hier = H.build(A)
    Ms = _with_layout(monkeypatch, "1", lambda: gpu.AMGPreconditioner(hier, prec))
    Mc = _with_layout(monkeypatch, "0", lambda: gpu.AMGPreconditioner(hier, prec))
    ls, lc = Ms.layout(), Mc.layout()
    print(prec, "sell layouts", ls[0], "stream layouts", lc[0], flush=True)
    assert ls[0]["G"][0] == "sell" and ls[0]["T"][0] == "sell" and ls[0]["A"][0] == "sell"
    assert all(row[w][0] != "sell" for row in lc for w in "AGTPR")
    b = H.rng_uniform(6, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(Ms(b), Mc(b))
    xs, its, rvs, fs = Ms.pcg(b, itmax=100, tol=1e-8)
    xc, itc, rvc, fc = Mc.pcg(b, itmax=100, tol=1e-8)
    assert its == itc and fs == fc == 0
    np.testing.assert_array_equal(xs, xc)
    np.testing.assert_array_equal(rvs, rvc)
    # And against the oracle on the same hierarchy.
    ref = oracle.Hier(hier)
    if prec == "fp64":
        xr = ref.vcycle(b)
        assert np.linalg.norm(Ms(b) - xr) <= 1e-12 * np.linalg.norm(xr)


def _with_env(monkeypatch, env, make):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        return make()
    finally:
        for k in env:
            monkeypatch.delenv(k)


@pytest.mark.parametrize("name,A", _mats())
def test_spmv_packed_layout_equals_oracle(gpu, monkeypatch, name, A):
    """The packed-code stream layout (CHRONOS_PACK=1, sliced ELL off) where the matrix admits
    it: same bits as the stream layout and the oracle."""
    x = np.random.default_rng(2).standard_normal(A.ncols)
    ref = oracle.spmv(A, x)
    M = _with_env(monkeypatch, {"CHRONOS_SELL": "0", "CHRONOS_PACK": "1"}, lambda: gpu.DeviceMatrix(A))
    print(f"{name} CHRONOS_PACK=1", flush=True)
    np.testing.assert_array_equal(M(x), ref)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_vcycle_packed_layout_bit_identical(gpu, monkeypatch, prec):
    A, _ = H.heterogeneous_diffusion(24, 2.0, seed=5)
    hier = H.build(A)
    Mp = _with_env(monkeypatch, {"CHRONOS_SELL": "0", "CHRONOS_PACK": "1"},
                   lambda: gpu.AMGPreconditioner(hier, prec))
    Mc = _with_env(monkeypatch, {"CHRONOS_SELL": "0", "CHRONOS_PACK": "0"},
                   lambda: gpu.AMGPreconditioner(hier, prec))
    lp = Mp.layout()
    print(prec, "packed layouts", lp[0], flush=True)
    assert lp[0]["G"][0][0] == "pack" and lp[0]["T"][0][0] == "pack"
    b = H.rng_uniform(7, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(Mp(b), Mc(b))
    xp, itp, rvp, fp_ = Mp.pcg(b, itmax=100, tol=1e-8)
    xc, itc, rvc, fc = Mc.pcg(b, itmax=100, tol=1e-8)
    assert itp == itc and fp_ == fc == 0
    np.testing.assert_array_equal(xp, xc)
