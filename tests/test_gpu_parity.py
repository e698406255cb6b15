"""GPU parity of the HIP solve phase against the CPU oracle on the same hierarchy."""
import numpy as np
import pytest

import oracle
from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _problem(kind, N, seed=3):
    if kind == "poisson":
        return H.poisson7(N)
    return H.heterogeneous_diffusion(N, 2.0, seed)[0]


@pytest.fixture(scope="module")
def case(request):
    return request.param


@pytest.mark.parametrize("kind,N", [("poisson", 12), ("diffusion", 20), ("poisson", 32)])
def test_vcycle_fp64_matches_oracle(gpu, kind, N):
    A = _problem(kind, N)
    hier = H.build(A)
    ref = oracle.Hier(hier)
    M = gpu.AMGPreconditioner(hier, "fp64")
    b = H.rng_uniform(5, 2, A.nrows, -1, 1)
    x = M(b)
    xr = ref.vcycle(b)
    rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(f"{kind}{N} fp64 V-cycle rel diff {rel:.3e}", flush=True)
    assert rel <= 1e-12


@pytest.mark.parametrize("kind,N", [("poisson", 12), ("diffusion", 20)])
def test_vcycle_fp32_close_to_oracle(gpu, kind, N):
    A = _problem(kind, N)
    hier = H.build(A)
    ref = oracle.Hier(hier)
    M = gpu.AMGPreconditioner(hier, "fp32")
    b = H.rng_uniform(5, 2, A.nrows, -1, 1)
    rel = np.linalg.norm(M(b) - ref.vcycle(b)) / np.linalg.norm(ref.vcycle(b))
    print(f"{kind}{N} fp32 V-cycle rel diff {rel:.3e}", flush=True)
    assert rel <= 1e-4  # FP32 cycle: single-precision rounding through ~L levels


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,N", [("poisson", 32), ("diffusion", 24)])
def test_pcg_matches_oracle(gpu, prec, kind, N):
    A = _problem(kind, N)
    hier = H.build(A)
    ref = oracle.Hier(hier)
    b = H.rng_uniform(9, 2, A.nrows, -1, 1)
    xr, itr, rvr, flr = ref.pcg(A, b, itmax=200, tol=1e-8)
    M = gpu.AMGPreconditioner(hier, prec)
    x, it, rv, fl = M.pcg(b, itmax=200, tol=1e-8)
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    tres = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    print(f"{kind}{N} {prec}: it {it} oracle {itr} err {err:.3e} true-res {tres:.3e}", flush=True)
    assert fl == 0 and abs(it - itr) <= 1
    assert tres <= 1e-8
    if prec == "fp64":
        assert err <= 1e-10
        np.testing.assert_allclose(rv[: min(it, itr) + 1], rvr[: min(it, itr) + 1], rtol=1e-7)


def test_pcg_fresh_outer_operators_and_growing_itmax(gpu):
    """Repeated solves with a new outer operator object each time (the graphs captured for a
    freed operator must not be reused even if the new one reuses its address, and the
    partial-sum buffer must fit each operator's row blocks), then a larger itmax."""
    from chronoslab_amd import amg

    A = _problem("diffusion", 16, 7)
    hier = H.build(A)
    M = gpu.AMGPreconditioner(hier, "fp64")
    b = H.rng_uniform(3, 2, A.nrows, -1, 1)
    x_ref, it_ref, _, _ = M.pcg(b, itmax=50, tol=1e-10)
    for trial in range(4):
        x, it, rv, fl = M.pcg(b, itmax=50, tol=1e-10, A=amg.DeviceMatrix(A))
        assert fl == 0 and it == it_ref
        np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-13 * np.abs(x_ref).max())
    x, it, rv, fl = M.pcg(b, itmax=400, tol=1e-10)
    assert it == it_ref and len(rv) == it + 1


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tail_collapse(gpu, monkeypatch, prec):
    """The V-cycle from the first level of <= CHRONOS_TAIL_ROWS rows down becomes one dense
    FP64 operator: the cycle equals the uncollapsed one to rounding, PCG takes the same
    iterations, and the collapsed level is the device hierarchy's last."""
    A = _problem("diffusion", 32, seed=8)
    hier = H.build(A)
    nrows = [L.A.nrows for L in hier.levels]
    c = next(l for l in range(1, hier.nlev - 1) if nrows[l] <= 4096)
    monkeypatch.setenv("CHRONOS_TAIL_ROWS", "0")
    Mfull = gpu.AMGPreconditioner(hier, prec)
    monkeypatch.setenv("CHRONOS_TAIL_ROWS", "4096")
    Mtail = gpu.AMGPreconditioner(hier, prec)
    print(f"levels {nrows}, collapse at level {c}: device levels {Mfull.stats()['nlev']} -> "
          f"{Mtail.stats()['nlev']}", flush=True)
    assert Mfull.stats()["nlev"] == hier.nlev and Mtail.stats()["nlev"] == c + 1
    b = H.rng_uniform(9, 2, A.nrows, -1, 1)
    xf, xt = Mfull(b), Mtail(b)
    rel = np.linalg.norm(xt - xf) / np.linalg.norm(xf)
    print(f"{prec}: collapsed vs full V-cycle rel diff {rel:.3e}", flush=True)
    assert rel <= (1e-13 if prec == "fp64" else 1e-5)
    if prec == "fp64":
        assert np.linalg.norm(xt - oracle.Hier(hier).vcycle(b)) <= 1e-12 * np.linalg.norm(xf)
    _, itf, _, flf = Mfull.pcg(b, itmax=100, tol=1e-8)
    _, itt, _, flt = Mtail.pcg(b, itmax=100, tol=1e-8)
    assert flf == flt == 0 and abs(itf - itt) <= (0 if prec == "fp64" else 1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_levels_against_unfused_and_oracle(gpu, monkeypatch, prec):
    """Levels >= 1 with composite operators ([-P^T A | P^T], [A P | A], [P | omega G^T G],
    omega G^T G: the same linear maps as the V(1,1) sequence, other roundings) against the
    unfused cycle and the oracle: FP64 within 1e-13 of the unfused cycle and 1e-12 of the
    oracle with equal PCG iterations; FP32 within 1e-5 and one iteration."""
    import ctypes as C

    import oracle

    A, _ = H.heterogeneous_diffusion(24, 2.0, 5)
    hier = H.build(A)
    b = H.rng_uniform(5, 2, A.nrows, -1.0, 1.0)
    fused = gpu.AMGPreconditioner(hier, prec)
    monkeypatch.setenv("CHRONOS_FUSE", "0")
    plain = gpu.AMGPreconditioner(hier, prec)
    monkeypatch.delenv("CHRONOS_FUSE")
    # Every level >= 1 below the tail is small here: all four composites are formed.
    per = C.c_int32()
    nd = fused.stats()["nlev"] - 1
    assert nd >= 2
    for l in range(1, nd):
        for w in "ruws":
            assert gpu._lib().chr_hier_layout(fused.ptr, l, w.encode(), C.byref(per), None, None, None) == 0
        assert gpu._lib().chr_hier_layout(plain.ptr, l, b"r", C.byref(per), None, None, None) != 0
    xf, xp, xo = fused(b), plain(b), oracle.Hier(hier).vcycle(b)
    tol_p, tol_o = (1e-13, 1e-12) if prec == "fp64" else (1e-5, 1e-5)
    assert np.linalg.norm(xf - xp) <= tol_p * np.linalg.norm(xp)
    assert np.linalg.norm(xf - xo) <= tol_o * np.linalg.norm(xo)
    _, itf, _, flf = fused.pcg(b, itmax=200, tol=1e-8)
    _, itp, _, flp = plain.pcg(b, itmax=200, tol=1e-8)
    assert flf == flp == 0
    assert itf == itp if prec == "fp64" else abs(itf - itp) <= 1
