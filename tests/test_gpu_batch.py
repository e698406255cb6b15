"""Batched solves: systems of uneven size and depth merged into one device hierarchy."""
import time

import numpy as np
import pytest

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _systems(edges, sigma, seed0):
    mats, hiers = [], []
    for i, N in enumerate(edges):
        A = H.heterogeneous_diffusion(int(N), sigma, seed0 + i)[0]
        mats.append(A)
        hiers.append(H.build(A))
    return mats, hiers


def _stack(sizes, k, seed):
    return np.asfortranarray(np.stack([H.rng_uniform(seed, 20 + c, int(sum(sizes)), -1, 1) for c in range(k)], 1))


@pytest.fixture
def no_tail(monkeypatch):
    # Batched hierarchies keep every level and every product of the V(1,1) sequence; the
    # single-system hierarchies they are compared with bit for bit must neither collapse
    # their tails nor fuse their coarse levels into composite operators.
    monkeypatch.setenv("CHRONOS_TAIL_ROWS", "0")
    monkeypatch.setenv("CHRONOS_FUSE", "0")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_vcycle_equals_each_system(gpu, no_tail, prec):
    edges = [14, 10, 20, 12, 17]  # the 10^3 system is a single (direct) level
    mats, hiers = _systems(edges, 1.5, 40)
    assert min(h.nlev for h in hiers) == 1 and max(h.nlev for h in hiers) >= 3
    bs = gpu.BatchSolver(hiers, prec)
    k = 3
    B = _stack(bs.sizes, k, 1)
    X = bs.vcycle(B)
    for s, h in enumerate(hiers):
        o0, o1 = bs.offsets[s], bs.offsets[s + 1]
        Xs = gpu.AMGPreconditioner(h, prec)(np.asfortranarray(B[o0:o1]))
        np.testing.assert_array_equal(X[o0:o1], Xs)


def test_batch_pcg_matches_single_solves(gpu, no_tail):
    edges = [14, 10, 20, 12]
    mats, hiers = _systems(edges, 1.5, 60)
    bs = gpu.BatchSolver(hiers, "fp64")
    k = 4
    B = _stack(bs.sizes, k, 2)
    X, iters, flags, resv = bs.pcg(B, itmax=200, tol=1e-8)
    assert np.all(flags == 0)
    for s, (A, h) in enumerate(zip(mats, hiers)):
        o0, o1 = bs.offsets[s], bs.offsets[s + 1]
        M = gpu.AMGPreconditioner(h, "fp64")
        for c in range(k):
            b = np.ascontiguousarray(B[o0:o1, c])
            x1, it1, rv1, fl1 = M.pcg(b, itmax=200, tol=1e-8)
            assert abs(iters[s, c] - it1) <= 1
            assert np.linalg.norm(X[o0:o1, c] - x1) <= 1e-10 * np.linalg.norm(x1)
            np.testing.assert_allclose(resv[s * k + c, : min(it1, iters[s, c]) + 1], rv1[: min(it1, iters[s, c]) + 1],
                                       rtol=1e-6, atol=1e-12)


def test_batch_zero_rhs_column_and_itmax(gpu):
    mats, hiers = _systems([12, 15], 1.0, 80)
    bs = gpu.BatchSolver(hiers, "fp64")
    B = _stack(bs.sizes, 2, 3)
    B[: bs.sizes[0], 1] = 0.0  # system 0, column 1: b = 0 -> x = 0, 0 iterations, flag 0
    X, iters, flags, _ = bs.pcg(B, itmax=200, tol=1e-8)
    assert iters[0, 1] == 0 and flags[0, 1] == 0 and not X[: bs.sizes[0], 1].any()
    X, iters, flags, _ = bs.pcg(B, itmax=1, tol=1e-8)
    assert iters[0, 0] == 1 and flags[0, 0] == 1


def test_batch_config_48_systems_8_rhs(gpu):
    """48 heterogeneous-diffusion systems, edge U{16..64}, ln k ~ N(0, 1.5^2), 8 random
    right-hand sides each, FP32 V-cycle, PCG_CJ to 1e-8 in one launch of a 16-iteration graph."""
    rng = np.random.default_rng(48)
    edges = rng.integers(16, 65, size=48)
    t = time.perf_counter()
    mats, hiers = _systems(edges, 1.5, 1000)
    t_setup = time.perf_counter() - t
    bs = gpu.BatchSolver(hiers, "fp32")
    k = 8
    B = _stack(bs.sizes, k, 4)
    bs.pcg(B, itmax=300, tol=1e-8)  # warm-up (graph capture)
    t = time.perf_counter()
    X, iters, flags, _ = bs.pcg(B, itmax=300, tol=1e-8)
    t_solve = time.perf_counter() - t
    worst = 0.0
    for s, A in enumerate(mats):
        o0, o1 = bs.offsets[s], bs.offsets[s + 1]
        R = B[o0:o1] - A.to_scipy() @ X[o0:o1]
        worst = max(worst, (np.linalg.norm(R, axis=0) / np.linalg.norm(B[o0:o1], axis=0)).max())
    print(f"batch: 48 systems ({bs.n} rows total) x 8 rhs: setup {t_setup:.1f}s, solve {t_solve * 1e3:.1f} ms "
          f"(host copies included), iterations {iters.min()}..{iters.max()}, worst true residual {worst:.2e}",
          flush=True)
    assert np.all(flags == 0)
    assert worst <= 1e-8


def test_batch_wider_block_after_narrow_solve(gpu):
    """A batch solved with k = 1 and then with k = 8 at the same itmax: the residual
    histories are regrown with the block width, and every solve equals a fresh batch's."""
    mats, hiers = _systems([12, 15, 10], 1.0, 90)
    bs = gpu.BatchSolver(hiers, "fp64")
    B8 = _stack(bs.sizes, 8, 5)
    bs.pcg(np.asfortranarray(B8[:, :1]), itmax=60, tol=1e-8)
    X, iters, flags, resv = bs.pcg(B8, itmax=60, tol=1e-8)
    X2, iters2, flags2, resv2 = gpu.BatchSolver(hiers, "fp64").pcg(B8, itmax=60, tol=1e-8)
    np.testing.assert_array_equal(iters, iters2)
    np.testing.assert_array_equal(flags, flags2)
    np.testing.assert_array_equal(resv, resv2)
    np.testing.assert_array_equal(X, X2)


def test_batch_graphs_rebuilt_after_wider_vcycle(gpu):
    """pcg(k=2), then a V-cycle with k=8 (which reallocates the cycle vectors the PCG graphs
    were captured with), then pcg(k=2) again: the second solve equals the first."""
    mats, hiers = _systems([12, 15], 1.0, 95)
    bs = gpu.BatchSolver(hiers, "fp64")
    B2 = _stack(bs.sizes, 2, 6)
    X1, it1, fl1, rv1 = bs.pcg(B2, itmax=80, tol=1e-8)
    bs.vcycle(_stack(bs.sizes, 8, 7))
    X2, it2, fl2, rv2 = bs.pcg(B2, itmax=80, tol=1e-8)
    np.testing.assert_array_equal(it1, it2)
    np.testing.assert_array_equal(X1, X2)
    np.testing.assert_array_equal(rv1, rv2)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_pcg_against_oracle(gpu, prec):
    """Each (system, column) solve of a batch against the oracle's PCG_CJ (PCG_CJ.m:1-73) on
    that system's own hierarchy: FP64 cycle the same iteration count and solution within
    1e-10, FP32 cycle within one iteration and the true residual at the tolerance."""
    import oracle

    edges = [14, 10, 20, 12, 17]
    mats, hiers = _systems(edges, 1.5, 110)
    bs = gpu.BatchSolver(hiers, prec)
    k = 3
    B = _stack(bs.sizes, k, 8)
    X, iters, flags, resv = bs.pcg(B, itmax=200, tol=1e-8)
    assert np.all(flags == 0)
    for s, (A, h) in enumerate(zip(mats, hiers)):
        o0, o1 = bs.offsets[s], bs.offsets[s + 1]
        ref = oracle.Hier(h)
        for c in range(k):
            b = np.ascontiguousarray(B[o0:o1, c])
            xr, itr, rvr, flr = ref.pcg(A, b, itmax=200, tol=1e-8)
            x = X[o0:o1, c]
            if prec == "fp64":
                assert iters[s, c] == itr, (s, c, iters[s, c], itr)
                assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)
                np.testing.assert_allclose(resv[s * k + c, : itr + 1], rvr, rtol=1e-7, atol=1e-14)
            else:
                assert abs(iters[s, c] - itr) <= 1, (s, c, iters[s, c], itr)
                assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)


def test_batch_graph_length_does_not_change_results(gpu, monkeypatch):
    """The batch solve as one launch of a 16-iteration graph, as launches of 2-iteration
    graphs, and with itmax cutting a launch short: the same iterations, histories and
    solutions bit for bit (kernels of iterations after every solve stopped exit at once)."""
    mats, hiers = _systems([12, 15, 10], 1.0, 120)
    B = _stack([h.levels[0].A.nrows for h in hiers], 4, 9)
    out = {}
    for gi in ("16", "2"):
        monkeypatch.setenv("CHRONOS_BATCH_GRAPH_ITERS", gi)
        out[gi] = gpu.BatchSolver(hiers, "fp64").pcg(B, itmax=60, tol=1e-8)
    for a, b in zip(out["16"], out["2"]):
        np.testing.assert_array_equal(a, b)
    monkeypatch.setenv("CHRONOS_BATCH_GRAPH_ITERS", "16")
    X, iters, flags, _ = gpu.BatchSolver(hiers, "fp64").pcg(B, itmax=5, tol=1e-8)
    # The 10^3 system is one direct level: one iteration solves it.
    assert np.all(iters[:2] == 5) and np.all(flags[:2] == 1)
    assert np.all(iters[2] == 1) and np.all(flags[2] == 0)
