"""Benchmark: time-to-1e-8 of AMG-preconditioned PCG_CJ on one MI355X.

Workload: 3D heterogeneous diffusion (porous flow), 7-point finite-volume stencil on a
128^3 grid (n = 2,097,152, nnz = 14,581,760), ln k ~ N(0, 2^2) i.i.d., harmonic face
transmissibilities, Dirichlet on the x = 0 and x = 1 faces, RHS U[-1,1].  Classical AMG
(CLA strength 0.25, PMIS, classical interpolation) with the adaptive FSAI smoother
(nstep 15, step 1, eps 1e-3), V(1,1), coarsest size <= 1000.

A step is one PCG_CJ solve from x0 = 0 to relative residual 1e-8 with b already resident in
HBM.  `value` is the mean time-to-tolerance of the FP32-cycle solve (FP32 V-cycle inside
FP64 PCG); the FP64-cycle solve, the V-cycle bandwidth and the CPU oracle are reported
beside it.  The hierarchy setup (host, OpenMP) is timed separately and is not in `value`.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# OpenMP threads of the CPU legs pinned close, one per physical core (read at the first
# OpenMP use in this process, so set before any native library loads).
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# The CPUs this process may use, read before any OpenMP runtime starts: with OMP_PROC_BIND
# set, the runtime binds the initial thread to its first place, after which the main
# thread's affinity mask holds that place alone.
try:
    _AFFINITY = sorted(os.sched_getaffinity(0))
except AttributeError:
    _AFFINITY = list(range(os.cpu_count() or 1))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# rocprofv3 --pmc summary of the roofline kernel (measured HBM bytes per launch).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "spmv_G0_probe_heterodiff128_fp32.md")


def _measured_traffic(n: int):
    """HBM bytes per launch of the probe kernel, from the committed counter summary, when it
    was measured on this workload size (None otherwise)."""
    import re

    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)\s+n=(\d+)", text, re.M)
    if not m or int(m.group(2)) != n:
        return None
    return int(m.group(1))


def _cpu_info():
    """(model name, logical CPUs this process may use, physical cores among them)."""
    model = platform.processor() or "unknown"
    out = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    cpus = _AFFINITY
    # Physical cores among them: distinct (package, core) pairs from sysfs; where sysfs does not
    # tell them apart (some virtual machines report one core), lscpu's threads per core.
    cores = set()
    for c in cpus:
        try:
            t = f"/sys/devices/system/cpu/cpu{c}/topology"
            cores.add((open(f"{t}/physical_package_id").read().strip(), open(f"{t}/core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    tpc = 1
    try:
        for line in out.splitlines():
            if line.startswith("Thread(s) per core:"):
                tpc = max(1, int(line.split(":", 1)[1]))
    except (NameError, ValueError):
        pass
    phys = len(cores) if len(cores) * tpc >= len(cpus) else max(1, len(cpus) // tpc)
    return model, len(cpus), phys


def _time_cpu(fn, reps: int = 5):
    """Median wall time of reps calls after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def vcycle_bytes(hier, s: int, k: int = 1) -> float:
    """Algorithmic bytes of one V-cycle on an n x k block (AMG_Vcycle.m:1-54, V(1,1), zero
    initial guess): every matrix streamed once per product as plain CSR (s-byte values, 4-byte
    columns and row pointers), every vector word read or written once per operation.  Per
    level l below the last: 2 B(A) + 2 B(G) + 2 B(G^T) + B(P) + B(P^T) + k (18 n_l + 2 n_{l+1}) s;
    last level: the dense FP64 inverse once, 8 m^2 + 16 m k."""
    def B(M):
        return M.nnz * (s + 4) + 4 * (M.nrows + 1)

    tot = 0.0
    for l, L in enumerate(hier.levels):
        if L.P is None:
            m = L.A.nrows
            tot += 8.0 * m * m + 16.0 * m * k
            continue
        GT = 4 * (L.G.ncols + 1) + L.G.nnz * (s + 4)
        PT = 4 * (L.P.ncols + 1) + L.P.nnz * (s + 4)
        tot += 2 * B(L.A) + 2 * B(L.G) + 2 * GT + B(L.P) + PT + k * (18 * L.A.nrows + 2 * L.P.ncols) * s
    return tot


def config_lobpcg(np, torch, amg, H, log) -> dict:
    """LOBPCG: 16 smallest eigenpairs of the 96^3 Poisson matrix, AMG preconditioner."""
    out = {}
    dev = torch.device("cuda:0")
    N, k = 96, 16
    A = H.poisson7(N)
    hl = H.build(A)
    M = amg.AMGPreconditioner(hl, "fp32")
    X0 = np.asfortranarray(np.stack([H.rng_uniform(7, 10 + c, A.nrows) for c in range(k)], axis=1))
    X0d = torch.from_numpy(np.ascontiguousarray(X0.T).ravel()).to(dev)  # column-major n x k
    Xd = torch.empty_like(X0d)
    ts = []
    for rep in range(4):  # one warm-up, three timed
        Xd.copy_(X0d)
        torch.cuda.synchronize()
        t = time.perf_counter()
        it, lam, rh, lh, ierr = amg.lobpcg_device(None, M, A.nrows, k, Xd.data_ptr(), itmax=300, tol=1e-8)
        torch.cuda.synchronize()
        if rep:
            ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    j = np.arange(1, N + 1)
    l1 = 2 - 2 * np.cos(j * np.pi / (N + 1))
    exact = np.sort((l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel())[:k]
    # Algorithmic bytes of the iterations (lobpcg.m:141-379), with a = the columns still active
    # (resnorm >= tol at the previous step): block V-cycle on a columns (FP32), FP64 product
    # A R (matrix once, a columns in and out), the Rayleigh-Ritz Gram blocks (the six tall
    # blocks X, R, P and A X, A R, A P read once: 2 (k + 2a) columns), the closing updates
    # (P and X, then their operator sides: 2 (3a + 2k) columns) and the residual (3k columns).
    n8 = A.nrows * 8.0
    BA = A.nnz * 12.0 + 4.0 * (A.nrows + 1)
    it_bytes = 0.0
    for i in range(1, it + 1):
        a = int(np.sum(rh[i - 1] >= 1e-8)) if i - 1 < rh.shape[0] else k
        a = max(a, 1)
        it_bytes += vcycle_bytes(hl, 4, a) + BA + 2 * a * n8 + 2 * (k + 2 * a) * n8 + 2 * (3 * a + 2 * k) * n8 + 3 * k * n8
    gbs = it_bytes / dt / 1e9
    # CPU baseline on a bounded sample: the oracle's lobpcg (same matrix, hierarchy, X0; FP64
    # V-cycle preconditioner) for its first 3 iterations, in which all 16 columns are active.
    import oracle as O

    _, _, cores = _cpu_info()
    O.set_threads(cores)
    ref = O.Hier(hl)
    O.lobpcg(A, ref, X0, itmax=2, tol=1e-8)  # warm-up (the oracle takes itmax >= 2)
    tc = []
    for _ in range(3):
        t = time.perf_counter()
        O.lobpcg(A, ref, X0, itmax=3, tol=1e-8)
        tc.append(time.perf_counter() - t)
    cpu_it = float(np.median(tc)) / 3
    del ref
    out["lobpcg_96cube_16"] = {"seconds": dt, "iterations": it, "ierr": ierr,
                               "cpu_oracle_s_per_iteration": cpu_it, "cpu_threads": cores,
                               "gpu_s_per_iteration": dt / max(it, 1),
                               "max_eig_rel_err_vs_closed_form": float(np.max(np.abs(lam - exact) / exact)),
                               "algorithmic_GB": it_bytes / 1e9, "GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS,
                               "note": "FP32 V-cycle preconditioner, FP64 Rayleigh-Ritz; X0 and X resident on the "
                                       "GPU; median of 3 after a warm-up; GB/s = algorithmic bytes of the iterations "
                                       "(active-column count per step) / whole solve time; CPU: the oracle's lobpcg "
                                       "(FP64 V-cycle) over its first 3 iterations (all 16 columns active), median of 3, "
                                       "per iteration; the GPU figure is the whole solve's mean per iteration"}
    log(f"[bench] LOBPCG 96^3 x16: {dt:.3f} s, {it} it, max eig err "
        f"{out['lobpcg_96cube_16']['max_eig_rel_err_vs_closed_form']:.2e}, {gbs:.0f} GB/s algorithmic "
        f"({gbs / HBM_PEAK_GBS:.2f} of HBM)")
    del M
    return out["lobpcg_96cube_16"]


def config_batch(np, torch, amg, H, log) -> dict:
    """Batching: 48 heterogeneous-diffusion systems, edge U{16..64}, ln k ~ N(0, 1.5^2), 8 rhs."""
    out = {}
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(48)
    edges = rng.integers(16, 65, size=48)
    mats, hiers = [], []
    for i, e in enumerate(edges):
        Ai = H.heterogeneous_diffusion(int(e), 1.5, 1000 + i)[0]
        mats.append(Ai)
        hiers.append(H.build(Ai))
    bs = amg.BatchSolver(hiers, "fp32")
    kk = 8
    B = np.asfortranarray(np.stack([H.rng_uniform(4, 20 + c, bs.n, -1, 1) for c in range(kk)], 1))
    bd = torch.from_numpy(np.ascontiguousarray(B.T).ravel()).to(dev)  # column-major n x k
    xd = torch.zeros_like(bd)
    torch.cuda.synchronize()
    bs.pcg_device(kk, bd.data_ptr(), xd.data_ptr(), itmax=300, tol=1e-8)
    ts = []
    for _ in range(3):
        xd.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        iters, flags = bs.pcg_device(kk, bd.data_ptr(), xd.data_ptr(), itmax=300, tol=1e-8)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    X = xd.cpu().numpy().reshape(kk, bs.n).T
    worst = 0.0
    for s_, Ai in enumerate(mats):
        o0, o1 = bs.offsets[s_], bs.offsets[s_ + 1]
        R = B[o0:o1] - Ai.to_scipy() @ X[o0:o1]
        worst = max(worst, float((np.linalg.norm(R, axis=0) / np.linalg.norm(B[o0:o1], axis=0)).max()))
    # Algorithmic bytes of one batch iteration: every system's block V-cycle on 8 columns (FP32)
    # and its outer FP64 product and vector work (PCG_CJ.m:38-58: 18 n words per column); a
    # converged solve is frozen but its rows are still streamed, so every iteration of the
    # batch moves the whole batch.
    bt_bytes = sum(vcycle_bytes(h_, 4, kk) + A_.nnz * 12.0 + 4.0 * (A_.nrows + 1) + 18.0 * A_.nrows * kk * 8
                   for h_, A_ in zip(hiers, mats))
    t_b = float(np.median(ts))
    gbs = bt_bytes * int(iters.max()) / t_b / 1e9
    # CPU baseline on a bounded sample: the oracle's PCG_CJ (FP64 V-cycle) of every system on
    # its first right-hand side (48 of the 384 solves), median of 3 after a warm-up.
    import oracle as O

    _, _, cores = _cpu_info()
    O.set_threads(cores)
    refs = [O.Hier(h_) for h_ in hiers]
    o0s = [bs.offsets[i] for i in range(len(mats) + 1)]

    def cpu_pass():
        for i_, (r_, A_) in enumerate(zip(refs, mats)):
            r_.pcg(A_, np.ascontiguousarray(B[o0s[i_]:o0s[i_ + 1], 0]), itmax=300, tol=1e-8)

    t_cpu = _time_cpu(cpu_pass, reps=3)
    del refs
    out["batch_48x8"] = {"seconds": t_b, "solves": 48 * kk, "rows": bs.n,
                         "iterations_min_max": [int(iters.min()), int(iters.max())],
                         "all_converged": bool(np.all(flags == 0)), "worst_true_rel_residual": worst,
                         "bytes_per_batch_iteration": bt_bytes, "GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS,
                         "cpu_oracle_s_48_solves": t_cpu, "cpu_threads": cores,
                         "note": "FP32 V-cycle, the batch solve one launch of a graph of 16 iterations, inputs resident on the GPU; GB/s = "
                                 "algorithmic bytes per batch iteration x iterations of the longest solve / solve time; CPU: the "
                                 "oracle's PCG_CJ (FP64 V-cycle) of every system on its first right-hand side (48 of the 384 "
                                 "solves, one after another), median of 3 after a warm-up"}
    log(f"[bench] batch 48x8: {t_b * 1e3:.1f} ms, iters {iters.min()}..{iters.max()}, worst res {worst:.2e}, "
        f"{gbs:.0f} GB/s algorithmic ({gbs / HBM_PEAK_GBS:.2f} of HBM)")
    del bs, hiers, mats
    return out["batch_48x8"]


def config_adjoint(np, torch, amg, H, log) -> dict:
    """Adjoint: 96^3, ln k ~ N(0,1); forward FP32-cycle PCG to 1e-10, adjoint and gradient."""
    out = {}
    dev = torch.device("cuda:0")
    from chronoslab_amd.autograd import FVDiffusionSolve
    import oracle as O  # the FP64 CPU gradient the GPU gradient is checked against

    N = 96
    lnk, lnk_obs = H.lnk_field(N, 1.0, 31), H.lnk_field(N, 1.0, 32)
    bh = H.rng_uniform(9, 2, N ** 3, -1, 1)
    bt = torch.from_numpy(bh).to(dev)
    x_obs = FVDiffusionSolve.apply(torch.from_numpy(lnk_obs).to(dev), bt, N, 1e-10, 1000, "fp64").detach()
    l = torch.from_numpy(lnk).to(dev).requires_grad_(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    x = FVDiffusionSolve.apply(l, bt, N, 1e-10, 1000, "fp32")
    t_fwd = time.perf_counter() - t
    loss = 0.5 * ((x - x_obs) ** 2).sum()
    t = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t
    g = l.grad.cpu().numpy()
    Ak = H.diffusion(N, np.exp(lnk))
    oh = O.Hier(H.build(Ak))
    xc, _, _, _ = oh.pcg(Ak, bh, itmax=1000, tol=1e-10)
    lamc, _, _, _ = oh.pcg(Ak, xc - x_obs.cpu().numpy(), itmax=1000, tol=1e-10)
    gc = O.fv_grad(N, np.exp(lnk), xc, lamc)
    rel = float(np.linalg.norm(g - gc) / np.linalg.norm(gc))
    out["adjoint_96cube"] = {"forward_s_incl_setup": t_fwd, "backward_s": t_bwd,
                             "grad_rel_err_vs_fp64_cpu": rel,
                             "note": "forward includes the host AMG setup of A(k); backward = adjoint "
                                     "PCG (same hierarchy) + fused dL/d(ln k) kernel"}
    log(f"[bench] adjoint 96^3: forward {t_fwd:.2f}s (incl. setup), backward {t_bwd * 1e3:.1f} ms, grad err {rel:.2e}")
    del oh, Ak
    return out["adjoint_96cube"]


def other_configs(np, torch, amg, H, log) -> dict:
    """LOBPCG (96^3 Poisson, 16 pairs), batching (48 systems x 8 rhs), the adjoint gradient
    (96^3) and the elasticity system, each timed with its inputs resident on the GPU."""
    import oracle as O

    return {"lobpcg_96cube_16": config_lobpcg(np, torch, amg, H, log),
            "batch_48x8": config_batch(np, torch, amg, H, log),
            "adjoint_96cube": config_adjoint(np, torch, amg, H, log),
            "elasticity_64cube": elasticity(np, torch, amg, H, O, log)}


def elasticity(np, torch, amg, H, O, log, ne: int = 64, precs=("fp32", "fp64")) -> dict:
    """Q1 linear elasticity on 64^3 hexahedra (823,875 dofs), E = 1, nu = 0.3, clamped z = 0
    face, uniform downward body force.  Set-up: DOM strength (tau 0.01), PMIS, BAMG
    prolongation fitted to the 6 rigid-body modes, then energy minimisation (prol_emin EMIN,
    the reference's default parameters), aFSAI smoother.  PCG_CJ to 1e-8 with the FP32 and the
    FP64 V-cycle, inputs resident on the GPU; error against the FP64 oracle solve."""
    dev = torch.device("cuda:0")
    A, rhs = H.elasticity(ne)
    Afix, _ = H.bc_fix(A, 1)  # computeAMG.m:5-10 (the clamped dofs are unit rows)
    t = time.perf_counter()
    hier = H.build(Afix, H.SetupParams(prolong="bamg", soc="dom", soc_tau=0.01, prol_emin="emin",
                                       tv=H.rigid_body_modes(ne), gpu_emin=True, gpu_afsai=True, gpu_omega=True,
                                       gpu_galerkin=True))
    t_setup = time.perf_counter() - t
    log(f"[bench] elasticity {ne}^3: n={A.nrows} nnz={A.nnz} setup {t_setup:.1f}s {hier.setup_seconds}\n{hier.describe()}")
    b = torch.from_numpy(rhs).to(dev)
    x = torch.zeros_like(b)
    res = {}
    for prec in precs:
        t = time.perf_counter()
        M = amg.AMGPreconditioner(hier, prec)
        t_up = time.perf_counter() - t
        log(f"[bench] elasticity {prec}: layouts {M.layout()}")
        ts = []
        for k in range(4):  # one warm-up, three timed
            x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            it, rv, fl = M.pcg_device(b.data_ptr(), x.data_ptr(), itmax=500, tol=1e-8)
            torch.cuda.synchronize()
            if k:
                ts.append(time.perf_counter() - t)
        xs = x.cpu().numpy()
        vms = M.time_vcycle(10)
        st = M.stats()
        res[prec] = {"mean_s": float(np.mean(ts)), "iterations": it, "flag": fl, "upload_and_autotune_s": t_up,
                     "true_rel_residual": float(np.linalg.norm(rhs - A.matvec(xs)) / np.linalg.norm(rhs)),
                     "vcycle_ms": vms, "vcycle_GBs": st["vcycle_csr_bytes"] / vms / 1e6, "x": xs}
        log(f"[bench] elasticity {prec}: {np.mean(ts) * 1e3:.1f} ms, {it} it, V-cycle {vms:.3f} ms "
            f"({st['vcycle_csr_bytes'] / vms / 1e6:.0f} GB/s algorithmic)")
        del M
    model, logical, cores = _cpu_info()
    O.set_threads(cores)
    t = time.perf_counter()
    xo, ito, _, flo = O.Hier(hier).pcg(Afix, rhs, itmax=500, tol=1e-8)
    t_cpu = time.perf_counter() - t
    O.set_threads(0)
    for prec, r in res.items():
        xs = r.pop("x")
        r["err_vs_fp64_oracle"] = float(np.linalg.norm(xs - xo) / np.linalg.norm(xo))
    log(f"[bench] elasticity oracle: {t_cpu:.2f} s ({ito} it) on {cores} threads (one per physical core)")
    return {"n": A.nrows, "nnz": A.nnz, "nnz_nonzero": int(np.count_nonzero(A.val)), "levels": hier.nlev,
            "setup_s": t_setup,
            "level_rows": [l.A.nrows for l in hier.levels], "level_nnz": [l.A.nnz for l in hier.levels],
            "solve": res,
            "speedup_fp32_vs_fp64": (res["fp64"]["mean_s"] / res["fp32"]["mean_s"]) if len(res) == 2 else None,
            "cpu_oracle_s": t_cpu, "cpu_oracle_iterations": ito, "cpu_threads": cores,
            "note": "BAMG + EMIN prolongation, aFSAI; host coarsening, interpolation and EMIN "
                    "constraints, GPU EMIN operator, aFSAI, omega and Galerkin products; the generated "
                    "operator stores exact zeros (nnz counts them, nnz_nonzero does not), which the upload "
                    "drops (same products); vcycle_GBs counts the matrices as uploaded, as MATLAB's nnz would"}


def config_poisson32(np, torch, amg, H, log) -> dict:
    """The reference's CPU-runnable case: 7-point Poisson on 32^3 (32,768 unknowns, Dirichlet
    boundary eliminated), RHS U[-1,1] from the seed, CLA + PMIS + classical prolongation, aFSAI
    smoother, PCG_CJ to 1e-8 in FP64.  The oracle (the C restatement of the reference's MATLAB
    path) is timed on the host's physical cores, median of 5, and on one thread; the GPU solves
    beside it with b resident."""
    import oracle

    dev = torch.device("cuda:0")
    A = H.poisson7(32)
    hier = H.build(A)
    bh = H.rng_uniform(1, 2, A.nrows, -1.0, 1.0)
    model, logical, cores = _cpu_info()
    oh = oracle.Hier(hier)
    box = {}

    def solve():
        box["o"] = oh.pcg(A, bh, itmax=500, tol=1e-8)

    oracle.set_threads(cores)
    t_cpu = _time_cpu(solve, 5)
    oracle.set_threads(1)
    t_one = _time_cpu(solve, 5)
    oracle.set_threads(0)
    xo, ito, _, _ = box["o"]
    b = torch.from_numpy(bh).to(dev)
    x = torch.zeros_like(b)
    gpu = {}
    for prec in ("fp64", "fp32"):
        M = amg.AMGPreconditioner(hier, prec)

        def gsolve():
            x.zero_()
            torch.cuda.synchronize()
            box["g"] = M.pcg_device(b.data_ptr(), x.data_ptr(), itmax=500, tol=1e-8)
            torch.cuda.synchronize()

        t = _time_cpu(gsolve, 5)
        it, _, fl = box["g"]
        xs = x.cpu().numpy()
        gpu[prec] = {"seconds": t, "iterations": it, "flag": fl,
                     "err_vs_fp64_oracle": float(np.linalg.norm(xs - xo) / np.linalg.norm(xo))}
        del M
    log(f"[bench] poisson 32^3: oracle {t_cpu * 1e3:.1f} ms on {cores} cores ({t_one * 1e3:.1f} ms on 1), {ito} it; "
        f"GPU fp64 {gpu['fp64']['seconds'] * 1e3:.2f} ms, fp32 {gpu['fp32']['seconds'] * 1e3:.2f} ms")
    return {"n": A.nrows, "nnz": A.nnz, "levels": hier.nlev, "cpu_oracle_s": t_cpu, "cpu_oracle_one_thread_s": t_one,
            "cpu_cores": cores, "cpu_model": model, "cpu_iterations": ito, "gpu": gpu,
            "note": "median of 5 after a warm-up for every figure; OpenMP one thread per physical core, "
                    "OMP_PROC_BIND=close"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=128, help="grid edge (default 128)")
    ap.add_argument("--sigma", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--only-c2", action="store_true",
                    help="skip the LOBPCG, batching and adjoint configurations")
    ap.add_argument("--only-elasticity", action="store_true",
                    help="only the 64^3 elasticity configuration (BAMG + EMIN set-up)")
    ap.add_argument("--only-config", choices=["lobpcg", "batch", "adjoint", "poisson32"],
                    help="only the named configuration (its JSON object on one line)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run 8 device-to-device copies of 512 MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_calib:
        import torch

        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(8):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_each": src.numel() * 8, "copies": 8}))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "the solve phase runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design only, not built)"}))
        return 2

    import numpy as np
    import torch

    from chronoslab_amd import amg, hierarchy as H

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    amg.init(0)
    if args.only_config:
        fn = {"lobpcg": config_lobpcg, "batch": config_batch, "adjoint": config_adjoint,
              "poisson32": config_poisson32}[args.only_config]
        print(json.dumps({args.only_config: fn(np, torch, amg, H, log)}), flush=True)
        return 0
    if args.only_elasticity:
        import oracle

        precs = ["fp32", "fp64"] if args.only == "both" else [args.only]
        print(json.dumps({"elasticity_64cube": elasticity(np, torch, amg, H, oracle, log, precs=precs)}),
              flush=True)
        return 0
    N = args.N
    A, lnk = H.heterogeneous_diffusion(N, args.sigma, args.seed)
    b_host = H.rng_uniform(args.seed, 2, A.nrows, -1.0, 1.0)
    # Setup: host classical coarsening and interpolation; aFSAI factors, omega and the Galerkin
    # products on the GPU (bit-identical factors and operators to the all-host setup, timed
    # beside it).
    t0 = time.perf_counter()
    hier = H.build(A, H.SetupParams(gpu_afsai=True, gpu_omega=True, gpu_galerkin=True))
    t_setup = time.perf_counter() - t0
    setup_parts = dict(hier.setup_seconds)
    t_setup_host = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        hh = H.build(A)
        t_setup_host = time.perf_counter() - t0
        same = all(np.array_equal(a.G.val, b.G.val) for a, b in zip(hh.levels[:-1], hier.levels[:-1]))
        same_a = all(np.array_equal(a.A.val, b.A.val) for a, b in zip(hh.levels, hier.levels))
        log(f"[bench] all-host setup {t_setup_host:.2f}s (aFSAI factors identical: {same}, "
            f"Galerkin operators identical: {same_a})")
        del hh
    log(f"[bench] {N}^3 n={A.nrows} nnz={A.nnz} setup {t_setup:.2f}s {setup_parts}\n{hier.describe()}")

    dev = torch.device("cuda:0")
    b = torch.from_numpy(b_host).to(dev)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()

    def run_solves(M, steps, warmup):
        for _ in range(warmup):
            x.zero_()
            torch.cuda.synchronize()
            M.pcg_device(b.data_ptr(), x.data_ptr(), itmax=500, tol=args.tol)
        times, its = [], []
        for _ in range(steps):
            x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            it, rv, fl = M.pcg_device(b.data_ptr(), x.data_ptr(), itmax=500, tol=args.tol)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
            its.append(it)
            if fl != 0:
                raise RuntimeError("PCG did not converge")
        xs = x.cpu().numpy()
        true_res = float(np.linalg.norm(b_host - A.matvec(xs)) / np.linalg.norm(b_host))
        return times, its, xs, true_res

    res = {}
    precs = ["fp32", "fp64"] if args.only == "both" else [args.only]
    upload = {}
    upload_cached = {}
    # The first upload times the layout candidates and records its decisions in a layout
    # cache; a second upload of the same hierarchy (timed after the solves) rebuilds them
    # without timing.
    cache_tmp = None
    if not os.environ.get("CHRONOS_LAYOUT_CACHE"):
        fd, cache_tmp = tempfile.mkstemp(prefix="chronos_layouts_", suffix=".txt")
        os.close(fd)
        os.environ["CHRONOS_LAYOUT_CACHE"] = cache_tmp
    for prec in precs:
        t0 = time.perf_counter()
        M = amg.AMGPreconditioner(hier, prec)
        upload[prec] = time.perf_counter() - t0
        times, its, xs, tres = run_solves(M, args.steps, args.warmup)
        st = M.stats()
        log(f"[bench] {prec}: upload + autotune {upload[prec]:.2f} s; layouts (PER or sell, code bits): {M.layout()}")
        vms = M.time_vcycle(20)
        gms, gbytes = M.time_spmv(0, "G", 20)
        tms, tbytes = M.time_spmv(0, "T", 20)
        ams, abytes = M.time_spmv(0, "A", 20)
        res[prec] = dict(times=times, its=its, x=xs, true_res=tres, stats=st, vcycle_ms=vms,
                         g_ms=gms, g_bytes=gbytes, t_ms=tms, t_bytes=tbytes, a_ms=ams, a_bytes=abytes)
        g_layout = M.layout(stored_bytes=True)[0]["G"]
        res[prec]["g_layout"] = g_layout
        log(f"[bench] {prec}: solve {1e3 * np.mean(times):.3f} ms (min {1e3 * np.min(times):.3f}) iters {its[0]} "
            f"true-res {tres:.2e} | V-cycle {vms:.3f} ms = {st['vcycle_csr_bytes'] / vms / 1e6:.0f} GB/s algorithmic "
            f"({st['vcycle_csr_bytes'] / 1e9:.3f} GB; stored {st['vcycle_bytes'] / 1e9:.3f} GB) | "
            f"G0 {gms * 1e3:.1f} us {gbytes / gms / 1e6:.0f} GB/s | "
            f"GT0 {tms * 1e3:.1f} us {tbytes / tms / 1e6:.0f} GB/s | A0 {ams * 1e3:.1f} us {abytes / ams / 1e6:.0f} GB/s")
        log(f"[bench] {prec}: level bytes (algorithmic) {[round(b / 1e6, 1) for b in st['level_bytes'][:hier.nlev]]} MB")
        lay = M.layout()
        del M
        t0 = time.perf_counter()
        M2 = amg.AMGPreconditioner(hier, prec)
        upload_cached[prec] = {"s": time.perf_counter() - t0, "same_layouts": M2.layout() == lay}
        log(f"[bench] {prec}: upload with the layout cache {upload_cached[prec]['s']:.2f} s, "
            f"same layouts: {upload_cached[prec]['same_layouts']}")
        del M2
    if cache_tmp:
        del os.environ["CHRONOS_LAYOUT_CACHE"]
        os.unlink(cache_tmp)

    # CPU leg: the oracle (FP64 restatement of AMG_Vcycle + PCG_CJ) on the same hierarchy.
    cpu = None
    err_vs_oracle = {}
    if not args.no_cpu:
        import oracle

        model, logical, cores = _cpu_info()
        oh = oracle.Hier(hier)
        res_o = {}

        def solve_oracle():
            res_o["out"] = oh.pcg(A, b_host, itmax=500, tol=args.tol)

        oracle.set_threads(cores)
        t_med = _time_cpu(solve_oracle, 5)
        xo, ito, rvo, flo = res_o["out"]
        oracle.set_threads(1)
        t = time.perf_counter()
        oh.pcg(A, b_host, itmax=500, tol=args.tol)
        t1 = time.perf_counter() - t
        oracle.set_threads(0)
        ts = [t_med]
        for prec, r in res.items():
            err_vs_oracle[prec] = float(np.linalg.norm(r["x"] - xo) / np.linalg.norm(xo))
        cpu = {"value": t_med, "unit": "s", "cores": cores, "kind": "port",
               "sample": (f"full oracle PCG_CJ solve of the same {N}^3 system to {args.tol:g} "
                          f"({ito} iterations), median of 5 after a warm-up, {cores} OpenMP threads, one per "
                          f"physical core (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                          f"OMP_PLACES={os.environ.get('OMP_PLACES')}; {logical} logical CPUs available; "
                          f"{model}); 1-thread solve {t1:.2f} s.  The reference's own path is MATLAB, "
                          f"which cannot run here; the oracle is its C restatement"),
               "one_thread_s": t1, "iters": ito}
        log(f"[bench] cpu oracle: {np.median(ts):.3f} s ({ito} it) on {cores} threads; 1 thread {t1:.2f} s")

    extra = {}
    if not args.no_cpu:
        extra["poisson_32cube_cpu_path"] = config_poisson32(np, torch, amg, H, log)
    if not args.only_c2:
        extra.update(other_configs(np, torch, amg, H, log))

    main_prec = "fp32" if "fp32" in res else "fp64"
    r = res[main_prec]
    val = float(np.mean(r["times"]))
    # Dominant kernel: the level-0 FSAI product (G and G^T carry the largest byte share).
    # achieved = algorithmic bytes (plain CSR with 4-byte indices + vector in and out) per
    # launch / mean launch time; traffic = HBM bytes per launch measured by counters.
    dom_ms = r["g_ms"]
    achieved = r["g_bytes"] / dom_ms / 1e6
    g_layout = r["g_layout"]
    tv = "float" if main_prec == "fp32" else "double"
    if g_layout[0] == "sell":
        kname = f"spmv_sell_kernel<{g_layout[1]}, U, {tv}, {tv}, EpiProbe<{tv}, 'G'>> (U: elements per chunk)"
    elif isinstance(g_layout[0], tuple):
        kname = f"spmv_pack_kernel<{g_layout[0][1]}, {tv}, {tv}, EpiProbe<{tv}, 'G'>>"
    else:
        kname = f"spmv_stream_kernel<{g_layout[0]}, {tv}, ..., EpiProbe<{tv}, 'G'>>"
    out = {
        "metric": "time-to-1e-8 residual (s) & V-cycle HBM GB/s, FP32-cycle vs FP64, 1 GPU",
        "value": val,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": val * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 V-cycle in f64 PCG" if main_prec == "fp32" else "f64",
        "data": "synthetic (seeded counter-based RNG: ln k field and RHS)",
        "config": {"workload": f"heterogeneous diffusion {N}^3 FV 7-pt, ln k~N(0,{args.sigma:g}^2), "
                               f"AMG(CLA+PMIS+classical, aFSAI) PCG_CJ to {args.tol:g}",
                   "n": A.nrows, "nnz": A.nnz, "levels": hier.nlev, "iterations": r["its"][0],
                   "setup_s": t_setup, "setup_parts_s": setup_parts, "setup_all_host_s": t_setup_host,
                   "upload_and_autotune_s": upload, "upload_with_layout_cache": upload_cached},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": _measured_traffic(A.nrows) if main_prec == "fp32" else None,
                     "kernel": f"level-0 G SpMV (FSAI factor), probe instance {kname}",
                     "bytes_per_launch": r["g_bytes"], "stored_matrix_bytes": g_layout[2],
                     "avg_launch_ms": dom_ms},
        "cpu_baseline": cpu,
        "vcycle": {p: {"ms": q["vcycle_ms"], "bytes": q["stats"]["vcycle_csr_bytes"],
                       "stored_bytes": q["stats"]["vcycle_bytes"],
                       "GBs": q["stats"]["vcycle_csr_bytes"] / q["vcycle_ms"] / 1e6,
                       "frac_hbm": q["stats"]["vcycle_csr_bytes"] / q["vcycle_ms"] / 1e6 / HBM_PEAK_GBS,
                       "level_bytes": q["stats"]["level_bytes"][:hier.nlev]}
                   for p, q in res.items()},
        "solve": {p: {"mean_s": float(np.mean(q["times"])), "min_s": float(np.min(q["times"])),
                      "iterations": q["its"][0], "true_rel_residual": q["true_res"],
                      "err_vs_fp64_oracle": err_vs_oracle.get(p)} for p, q in res.items()},
    }
    if "fp32" in res and "fp64" in res:
        out["speedup_fp32_vs_fp64"] = float(np.mean(res["fp64"]["times"]) / np.mean(res["fp32"]["times"]))
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
