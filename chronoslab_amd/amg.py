"""Python consumer of the HIP solve phase, mirroring ChronosLab's solve-phase interface.

Reference interface (paths relative to the reference root) and the mirror here:
  M = @(x) AMG_Vcycle(AMG_prec, A, x)            AMGPreconditioner.__call__(b)
      (sources/utils/computeAMG.m:49)
  [M, T_setup] = computeAMG(A, TV0, param, verb)  computeAMG(A, param) -> (M, T_setup)
      (sources/utils/computeAMG.m:1-51)
  [x,iter,resvec,flag] = PCG_CJ(prodM, prodA, b, x0, itmax, tol)
      (sources/LinearSolver/PCG_CJ.m:1-73)     PCG_CJ(M, A, b, x0, itmax, tol)

All compute runs in libchronos_amg.so on the GPU; a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import hierarchy as H
from ._amg_sigs import ChrCsr, ChrHierStats, ChrLevel
from ._lib import load_amg_lib

FP64 = 0
FP32_CYCLE = 1
PTR_HOST = 0
PTR_DEVICE = 1

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


def _lib():
    return load_amg_lib()


def _release(fn: str, ptr) -> None:
    """Destroys a library object; at interpreter exit the module globals may already be gone,
    and the process teardown releases the device memory anyway."""
    try:
        getattr(_lib(), fn)(ptr)
    except (TypeError, AttributeError):
        pass


def _check(rc: int, what: str = ""):
    if rc != 0:
        raise RuntimeError(f"chronos_amg error {rc} {what}: {_lib().chr_last_error().decode()}")


def _csr(m: H.Csr | None, keep: list) -> ChrCsr:
    if m is None:
        return ChrCsr(0, 0, None, None, None)
    rp = np.ascontiguousarray(m.rowptr, np.int32)
    ci = np.ascontiguousarray(m.col, np.int32)
    v = np.ascontiguousarray(m.val, np.float64)
    keep.extend([rp, ci, v])
    return ChrCsr(m.nrows, m.ncols, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P),
                  v.ctypes.data_as(_F64P))


def _precision(p) -> int:
    if isinstance(p, str):
        return {"fp64": FP64, "fp32": FP32_CYCLE, "fp32_cycle": FP32_CYCLE}[p.lower()]
    return int(p)


def init(device: int = 0) -> None:
    _check(_lib().chr_init(device), "chr_init")


_RESERVE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.POINTER(C.c_int32)),
                       C.POINTER(C.POINTER(C.c_double)))


def galerkin(A: H.Csr, P: H.Csr) -> H.Csr | None:
    """Galerkin coarse operator P^T (A P) on the GPU (chr_galerkin; cpt_aspAMG.m:277), equal bit
    for bit to hierarchy.galerkin's host product.  None when the GPU declines (a product row
    with more than 4096 distinct columns)."""
    keep = []
    ca, cp = _csr(A, keep), _csr(P, keep)
    rp = np.empty(P.ncols + 1, np.int32)
    out = {}

    def reserve(_sink, nnz, col, val):
        out["c"] = np.empty(max(nnz, 1), np.int32)
        out["v"] = np.empty(max(nnz, 1), np.float64)
        out["n"] = nnz
        col[0] = out["c"].ctypes.data_as(_I32P)
        val[0] = out["v"].ctypes.data_as(_F64P)
        return 0

    cb = _RESERVE(reserve)
    rc = _lib().chr_galerkin(C.byref(ca), C.byref(cp), rp.ctypes.data_as(_I32P), None, C.cast(cb, C.c_void_p))
    if rc == 1:
        return None
    _check(rc, "chr_galerkin")
    n = out["n"]
    return H.Csr(P.ncols, P.ncols, rp, out["c"][:n].copy(), out["v"][:n].copy())


def afsai(A: H.Csr, nstep: int = 15, step_size: int = 1, eps: float = 1e-3) -> tuple[H.Csr, bool]:
    """Adaptive symmetric FSAI factor G of A on the GPU (chr_afsai): one batched launch over
    all rows.  Mirrors the reference's SYM_aFSAI_compute gateway
    (sources/Preconditioner/AMG/smoother/MEX_SYM_AFSAI/SYM_aFSAI_compute.cpp:49-149) and
    returns (G, is_posdef); G equals hierarchy.afsai's host factor bit for bit."""
    keep = []
    c = _csr(A, keep)
    n = A.nrows
    kmax = max(1, nstep * step_size) + 1
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(max(1, n * kmax), np.int32)
    v = np.empty(max(1, n * kmax), np.float64)
    nnz = C.c_int64()
    pd = C.c_int32()
    _check(_lib().chr_afsai(C.byref(c), nstep, step_size, eps, rp.ctypes.data_as(_I32P),
                            ci.ctypes.data_as(_I32P), v.ctypes.data_as(_F64P), C.byref(nnz), C.byref(pd)),
           "chr_afsai")
    k = nnz.value
    return H.Csr(n, n, rp, ci[:k].copy(), v[:k].copy()), bool(pd.value)


def fsai_omega(A: H.Csr, G: H.Csr, seed: int = 1) -> tuple[float, float]:
    """omega = min(1, 1.9 / lambda_max(G A G^T)) by Lanczos on the GPU (chr_fsai_omega);
    returns (omega, lambda_max).  The smoother damping of
    sources/Preconditioner/AMG/smoother/smoother.m:72-81."""
    keep = []
    ca, cg = _csr(A, keep), _csr(G, keep)
    om, lam = C.c_double(), C.c_double()
    _check(_lib().chr_fsai_omega(C.byref(ca), C.byref(cg), seed, C.byref(om), C.byref(lam)), "chr_fsai_omega")
    return om.value, lam.value


class DeviceMatrix:
    """A CSR operator resident on the GPU in FP64 (the outer prodA)."""

    def __init__(self, A: H.Csr):
        keep = []
        c = _csr(A, keep)
        self.ptr = C.c_void_p()
        _check(_lib().chr_mat_create(C.byref(c), C.byref(self.ptr)), "chr_mat_create")
        self.shape = (A.nrows, A.ncols)

    def __del__(self):
        if getattr(self, "ptr", None):
            if _release is not None:  # None once the module is torn down at interpreter exit
                _release("chr_mat_destroy", self.ptr)
            self.ptr = None

    def __call__(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.shape[0])
        _check(_lib().chr_spmv(self.ptr, x.ctypes.data, y.ctypes.data, PTR_HOST), "chr_spmv")
        return y


class AMGPreconditioner:
    """The uploaded hierarchy; calling it applies one V-cycle (AMG_Vcycle with x0 = 0)."""

    def __init__(self, hier: H.Hierarchy, precision="fp64"):
        self.hier = hier
        self.precision = _precision(precision)
        keep = []
        nlev = hier.nlev
        arr = (ChrLevel * nlev)()
        for l, L in enumerate(hier.levels):
            arr[l] = ChrLevel(_csr(L.A, keep), _csr(L.P, keep), _csr(L.G, keep), L.omega, L.nupre, L.nupost)
        self.ptr = C.c_void_p()
        _check(_lib().chr_hier_create(nlev, arr, self.precision, C.byref(self.ptr)), "chr_hier_create")
        self.n = hier.levels[0].A.nrows

    def __del__(self):
        if getattr(self, "ptr", None):
            if _release is not None:  # None once the module is torn down at interpreter exit
                _release("chr_hier_destroy", self.ptr)
            self.ptr = None

    def stats(self) -> dict:
        st = ChrHierStats()
        _check(_lib().chr_hier_get_stats(self.ptr, C.byref(st)))
        nl = st.nlev
        return {"nlev": nl, "n": list(st.n[:nl]), "vcycle_bytes": st.vcycle_bytes, "vcycle_csr_bytes": st.vcycle_csr_bytes,
                "level_bytes": list(st.level_bytes[:nl]), "pcg_iter_bytes": st.pcg_iter_bytes,
                "device_bytes": st.device_bytes}

    def layout(self, stored_bytes: bool = False) -> list:
        """Per level, the layout chosen for the single-vector products of A, G, G^T, P, P^T:
        ("sell", code bits) for the sliced-ELL kernel, (PER, code bits) for the stream kernel,
        (("pack", PER), 16) for the packed-code stream kernel, ("dia", 0) for the symmetric
        diagonal kernel,
        with the bytes one product reads of the stored matrix appended when stored_bytes."""
        out = []
        for l in range(self.stats()["nlev"] - 1):  # device levels (the tail may be collapsed)
            row = {}
            for w in "AGTPR":
                per, bits, nnz, sb = C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
                _check(_lib().chr_hier_layout(self.ptr, l, w.encode(), C.byref(per), C.byref(bits),
                                              C.byref(nnz), C.byref(sb)))
                kind = (("dia" if bits.value == 0 else "sell") if per.value == 0
                        else ("pack", -per.value) if per.value < 0 else per.value)
                row[w] = (kind, bits.value) + ((sb.value,) if stored_bytes else ())
            out.append(row)
        return out

    def __call__(self, b: np.ndarray) -> np.ndarray:
        b = np.asfortranarray(b, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        x = np.empty_like(b, order="F")
        _check(_lib().chr_vcycle_apply(self.ptr, k, b.ctypes.data, self.n, x.ctypes.data, self.n, PTR_HOST),
               "chr_vcycle_apply")
        return x

    def apply_device(self, k: int, b_ptr: int, x_ptr: int, ld: int | None = None) -> None:
        ld = self.n if ld is None else ld
        _check(_lib().chr_vcycle_apply(self.ptr, k, b_ptr, ld, x_ptr, ld, PTR_DEVICE), "chr_vcycle_apply")

    def pcg(self, b, x0=None, itmax: int = 1000, tol: float = 1e-8, A: DeviceMatrix | None = None):
        """PCG_CJ with this V-cycle as prodM (host arrays)."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        it, fl = C.c_int32(), C.c_int32()
        resvec = np.zeros(itmax + 1)
        _check(_lib().chr_pcg_solve(self.ptr, A.ptr if A is not None else None, b.ctypes.data, x.ctypes.data,
                                    itmax, tol, PTR_HOST, C.byref(it), resvec.ctypes.data_as(_F64P),
                                    C.byref(fl)), "chr_pcg_solve")
        return x, it.value, resvec[: it.value + 1], fl.value

    def pcg_device(self, b_ptr: int, x_ptr: int, itmax: int = 1000, tol: float = 1e-8,
                   A: DeviceMatrix | None = None):
        it, fl = C.c_int32(), C.c_int32()
        resvec = np.zeros(itmax + 1)
        _check(_lib().chr_pcg_solve(self.ptr, A.ptr if A is not None else None, b_ptr, x_ptr, itmax, tol,
                                    PTR_DEVICE, C.byref(it), resvec.ctypes.data_as(_F64P), C.byref(fl)),
               "chr_pcg_solve")
        return it.value, resvec[: it.value + 1], fl.value

    def pcg_noalloc_device(self, b_ptr: int, x_ptr: int, itmax: int = 1000, tol: float = 1e-8,
                           A: DeviceMatrix | None = None):
        """chr_pcg_solve_noalloc: PCG_CJ on device pointers with the library's fixed scratch
        array as working memory (no device allocation); returns (iter, resvec, flag)."""
        it, fl = C.c_int32(), C.c_int32()
        resvec = np.zeros(itmax + 1)
        _check(_lib().chr_pcg_solve_noalloc(self.ptr, A.ptr if A is not None else None, b_ptr, x_ptr, itmax, tol,
                                            C.byref(it), resvec.ctypes.data_as(_F64P), C.byref(fl)),
               "chr_pcg_solve_noalloc")
        return it.value, resvec[: it.value + 1], fl.value

    def time_vcycle(self, reps: int = 20) -> float:
        ms = C.c_double()
        _check(_lib().chr_time_vcycle(self.ptr, reps, C.byref(ms)), "chr_time_vcycle")
        return ms.value

    def time_spmv(self, level: int, which: str, reps: int = 20) -> tuple[float, float]:
        ms, by = C.c_double(), C.c_double()
        _check(_lib().chr_time_spmv(self.ptr, level, which.encode(), reps, C.byref(ms), C.byref(by)),
               "chr_time_spmv")
        return ms.value, by.value


class BatchSolver:
    """Many independent systems (hierarchies of uneven depth) merged on the device; one
    V-cycle / PCG_CJ call advances every system and every right-hand-side column."""

    def __init__(self, hiers: list, precision="fp32"):
        self._keep = []
        self.nsys = len(hiers)
        arrays = []
        for hh in hiers:
            arr = (ChrLevel * hh.nlev)()
            for l, L in enumerate(hh.levels):
                arr[l] = ChrLevel(_csr(L.A, self._keep), _csr(L.P, self._keep), _csr(L.G, self._keep), L.omega,
                                  L.nupre, L.nupost)
            arrays.append(arr)
        self._arrays = arrays
        ptrs = (C.POINTER(ChrLevel) * self.nsys)(*[C.cast(a, C.POINTER(ChrLevel)) for a in arrays])
        nlev = np.array([hh.nlev for hh in hiers], np.int32)
        self.sizes = [hh.levels[0].A.nrows for hh in hiers]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.ptr = C.c_void_p()
        _check(_lib().chr_batch_create(self.nsys, nlev.ctypes.data_as(_I32P), ptrs, _precision(precision),
                                       C.byref(self.ptr)), "chr_batch_create")

    def __del__(self):
        if getattr(self, "ptr", None):
            if _release is not None:  # None once the module is torn down at interpreter exit
                _release("chr_batch_destroy", self.ptr)
            self.ptr = None

    def vcycle(self, B: np.ndarray) -> np.ndarray:
        B = np.asfortranarray(B, np.float64)
        k = 1 if B.ndim == 1 else B.shape[1]
        X = np.empty_like(B, order="F")
        _check(_lib().chr_batch_vcycle(self.ptr, k, B.ctypes.data, X.ctypes.data, PTR_HOST), "chr_batch_vcycle")
        return X

    def pcg(self, B: np.ndarray, X0=None, itmax: int = 500, tol: float = 1e-8):
        B = np.asfortranarray(B, np.float64)
        k = 1 if B.ndim == 1 else B.shape[1]
        X = np.zeros_like(B, order="F") if X0 is None else np.asfortranarray(X0, np.float64).copy(order="F")
        Q = self.nsys * k
        iters = np.zeros(Q, np.int32)
        flags = np.zeros(Q, np.int32)
        resvec = np.zeros((itmax + 1) * Q)
        _check(_lib().chr_batch_pcg(self.ptr, k, B.ctypes.data, X.ctypes.data, itmax, tol, PTR_HOST,
                                    iters.ctypes.data_as(_I32P), flags.ctypes.data_as(_I32P),
                                    resvec.ctypes.data_as(_F64P)), "chr_batch_pcg")
        return X, iters.reshape(self.nsys, k), flags.reshape(self.nsys, k), resvec.reshape(Q, itmax + 1)

    def pcg_device(self, k: int, b_ptr: int, x_ptr: int, itmax: int = 500, tol: float = 1e-8):
        Q = self.nsys * k
        iters = np.zeros(Q, np.int32)
        flags = np.zeros(Q, np.int32)
        _check(_lib().chr_batch_pcg(self.ptr, k, b_ptr, x_ptr, itmax, tol, PTR_DEVICE,
                                    iters.ctypes.data_as(_I32P), flags.ctypes.data_as(_I32P), None),
               "chr_batch_pcg")
        return iters.reshape(self.nsys, k), flags.reshape(self.nsys, k)


def lobpcg(prodA, prodM: AMGPreconditioner | None, X0: np.ndarray, largest: bool = False,
           reslambda_check: bool = False, itmax: int = 100, tol: float = 1e-8, restart_control: int = 7,
           Y: np.ndarray | None = None):
    """[niter, lambda, X, resnorm_vec, lambda_vec, ierr] = lobpcg(prodA, [], prodM, Y, X0, ...)
    (sources/EigenSolver/lobpcg.m:21-23) on the GPU.  prodA: DeviceMatrix, host Csr or None
    (the preconditioner's level-0 operator); prodM: the uploaded hierarchy or None; Y: the
    constraint block (n x m, m <= 48) or None."""
    if isinstance(prodA, H.Csr):
        prodA = DeviceMatrix(prodA)
    if prodA is None and prodM is None:
        raise ValueError("lobpcg needs prodA or prodM")
    X = np.asfortranarray(X0, dtype=np.float64).copy(order="F")
    n, neig = X.shape
    Yf = None if Y is None else np.asfortranarray(Y, dtype=np.float64)
    m = 0 if Yf is None else Yf.shape[1]
    niter, ierr = C.c_int32(), C.c_int32()
    lam = np.zeros(neig)
    rh = np.zeros((itmax + 1, neig), order="F")
    lh = np.zeros((itmax + 1, neig), order="F")
    _check(_lib().chr_lobpcg(prodM.ptr if prodM is not None else None, prodA.ptr if prodA is not None else None,
                             n, neig, X.ctypes.data, n, Yf.ctypes.data if Yf is not None else None, m, n,
                             int(largest), int(reslambda_check), itmax, tol, restart_control, PTR_HOST,
                             C.byref(niter), lam.ctypes.data_as(_F64P), rh.ctypes.data_as(_F64P),
                             lh.ctypes.data_as(_F64P), C.byref(ierr)), "chr_lobpcg")
    k = niter.value + 1
    return niter.value, lam, X, rh[:k], lh[:k], ierr.value


def lobpcg_device(prodA, prodM: AMGPreconditioner | None, n: int, neig: int, x_ptr: int, largest: bool = False,
                  reslambda_check: bool = False, itmax: int = 100, tol: float = 1e-8, restart_control: int = 7):
    """lobpcg as above on a device block: x_ptr holds X0 (n x neig, column-major FP64) on entry
    and the eigenvectors on exit.  Returns (niter, lambda, resnorm_hist, lambda_hist, ierr)."""
    if isinstance(prodA, H.Csr):
        prodA = DeviceMatrix(prodA)
    if prodA is None and prodM is None:
        raise ValueError("lobpcg needs prodA or prodM")
    niter, ierr = C.c_int32(), C.c_int32()
    lam = np.zeros(neig)
    rh = np.zeros((itmax + 1, neig), order="F")
    lh = np.zeros((itmax + 1, neig), order="F")
    _check(_lib().chr_lobpcg(prodM.ptr if prodM is not None else None, prodA.ptr if prodA is not None else None,
                             n, neig, x_ptr, n, None, 0, n, int(largest), int(reslambda_check), itmax, tol,
                             restart_control, PTR_DEVICE, C.byref(niter), lam.ctypes.data_as(_F64P),
                             rh.ctypes.data_as(_F64P), lh.ctypes.data_as(_F64P), C.byref(ierr)), "chr_lobpcg")
    k = niter.value + 1
    return niter.value, lam, rh[:k], lh[:k], ierr.value


def computeAMG(A: H.Csr, param: H.SetupParams | None = None, precision="fp64", seed: int = 1):
    """computeAMG.m:1-51: BC rule (unit diagonal -> lambda_max/10), setup, upload.

    Returns (M, T_setup) with M the uploaded preconditioner (callable like the handle
    `@(x) AMG_Vcycle(AMG_prec, A, x)`, A being the BC-modified operator)."""
    Afix, _ = H.bc_fix(A, seed)
    t0 = time.perf_counter()
    hier = H.build(Afix, param)
    T_setup = time.perf_counter() - t0
    return AMGPreconditioner(hier, precision), T_setup


def SQMR(Afun, rhs, tol: float, itmax: int, lprec, rprec=None, x0=None):
    """[pnew, info, relres, niter, resvec] = SQMR(Afun, rhs, tol, itmax, lprec, rprec, x0)
    (sources/LinearSolver/SQMR.m:1-108): symmetric QMR with the preconditioner split as
    lprec then rprec (rprec None: identity), the residual recurrence's relative norm as the
    stopping test, info 0 converged / 1 itmax reached / 10 small sigma / 20 small rho0.  The
    host loop runs on vectors of the caller's type; the preconditioner is typically the GPU
    V-cycle (AMGPreconditioner)."""
    rprec = rprec or (lambda v: v)
    rhs = np.asarray(rhs, np.float64).ravel()
    nn = rhs.shape[0]
    d = np.zeros(nn)
    p = np.zeros(nn) if x0 is None else np.array(x0, np.float64).ravel()
    resvec = np.zeros(itmax)
    info, niter, done = 0, 0, False
    bnorm = np.linalg.norm(rhs)
    r = rhs - Afun(p)
    vscr = lprec(r)
    q = rprec(vscr)
    tau = np.linalg.norm(vscr)
    rho0 = r @ q
    theta0 = 0.0
    while not done and niter < itmax:
        niter += 1
        vscr = Afun(q)
        sigma = q @ vscr
        if abs(sigma) < 1e-20 * rho0:
            info = 10
            break
        alpha = rho0 / sigma
        r = r - alpha * vscr
        vscr = lprec(r)
        theta1 = np.linalg.norm(vscr) / tau
        gamma = 1.0 / np.sqrt(1.0 + theta1 * theta1)
        tau = tau * theta1 * gamma
        d = d * gamma * gamma * theta0 * theta0 + gamma * gamma * alpha * q
        theta0 = theta1
        p = p + d
        resnorm = np.linalg.norm(r) / bnorm
        if resnorm < tol:
            done = True
        else:
            u = rprec(vscr)
            rho1 = r @ u
            if abs(rho0) < 1e-20 * rho1:
                info = 20
                break
            beta = rho1 / rho0
            rho0 = rho1
            q = u + beta * q
        resvec[niter - 1] = resnorm
    if niter >= itmax:
        info = 1
    relres = np.linalg.norm(rhs - Afun(p)) / bnorm
    return p, info, relres, niter, resvec[:niter]


def PCG_CJ(prodM: AMGPreconditioner, prodA, b, x0, itmax: int, tol: float):
    """[x, iter, resvec, flag] = PCG_CJ(prodM, prodA, b, x0, itmax, tol) on the GPU.

    prodA is a DeviceMatrix, a host Csr (uploaded for the call), or None for the
    preconditioner's own level-0 operator."""
    if isinstance(prodA, H.Csr):
        prodA = DeviceMatrix(prodA)
    x, it, resvec, flag = prodM.pcg(b, x0, itmax, tol, prodA)
    return x, it, resvec, flag
