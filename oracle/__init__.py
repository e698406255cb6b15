"""ctypes bindings of the CPU FP64 oracle — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
chronoslab_amd package.  `ref_classical_prol` binds oracle/_ref/libref_clas.so, the
reference's own classical-prolongation core compiled from the reference tree by
oracle/Makefile (absent where the reference tree is absent).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_clas.so")
REF_EXTI_LIB = os.path.join(HERE, "_ref", "libref_exti.so")

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


class OrcCsr(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("rowptr", _I32P), ("col", _I32P),
                ("val", _F64P)]


class OrcLevel(C.Structure):
    _fields_ = [("A", OrcCsr), ("P", OrcCsr), ("G", OrcCsr), ("omega", C.c_double),
                ("nupre", C.c_int32), ("nupost", C.c_int32)]


_lib = None
_ref = None
_ref_exti = None


def sources() -> list:
    return [os.path.join(HERE, f) for f in ("oracle.c", "oracle.h", "Makefile")]


def is_current() -> bool:
    """liboracle.so exists and was built from oracle.c / oracle.h as they are now."""
    import hashlib

    h = hashlib.sha256()
    for p in sources():
        with open(p, "rb") as f:
            h.update(f.read())
    try:
        with open(LIB + ".srchash") as f:
            return os.path.exists(LIB) and f.read().strip() == h.hexdigest()
    except OSError:
        return False


def build(reference: bool = True) -> None:
    import hashlib

    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    h = hashlib.sha256()
    for p in sources():
        with open(p, "rb") as f:
            h.update(f.read())
    with open(LIB + ".srchash", "w") as f:
        f.write(h.hexdigest() + "\n")
    if reference and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(reference=False)
        L = C.CDLL(LIB)
        L.orc_hier_create.argtypes = [C.c_int32, C.POINTER(OrcLevel), C.POINTER(C.c_void_p)]
        L.orc_hier_free.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_spmv.argtypes = [C.POINTER(OrcCsr), _F64P, _F64P]
        L.orc_vcycle.argtypes = [C.c_void_p, C.c_int32, _F64P, C.c_int64, _F64P, C.c_int64]
        L.orc_pcg.argtypes = [C.c_void_p, C.POINTER(OrcCsr), _F64P, _F64P, C.c_int32, C.c_double,
                              _I32P, _F64P, _I32P]
        L.orc_lobpcg.argtypes = [C.POINTER(OrcCsr), C.c_void_p, _F64P, _F64P, C.c_int32, C.c_int32, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, _I32P, _F64P, _F64P,
                                 _F64P, _I32P]
        L.orc_fv_grad.argtypes = [C.c_int32, _F64P, _F64P, _F64P, _F64P]
        L.orc_sygv.argtypes = [C.c_int32, _F64P, _F64P, _F64P, _F64P]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class _Keep:
    pass


def _csr(m, keep):
    rp = np.ascontiguousarray(m.rowptr, np.int32)
    ci = np.ascontiguousarray(m.col, np.int32)
    v = np.ascontiguousarray(m.val, np.float64)
    keep.extend([rp, ci, v])
    return OrcCsr(m.nrows, m.ncols, _p(rp, _I32P), _p(ci, _I32P), _p(v, _F64P))


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


class Hier:
    """Oracle hierarchy over a chronoslab_amd.hierarchy.Hierarchy (same matrices)."""

    def __init__(self, h):
        self._keep = []
        nlev = h.nlev
        arr = (OrcLevel * nlev)()
        empty = OrcCsr(0, 0, None, None, None)
        for l, L in enumerate(h.levels):
            A = _csr(L.A, self._keep)
            if L.P is not None:
                arr[l] = OrcLevel(A, _csr(L.P, self._keep), _csr(L.G, self._keep), L.omega, L.nupre, L.nupost)
            else:
                arr[l] = OrcLevel(A, empty, empty, 1.0, 1, 1)
        self._levels = arr
        self.ptr = C.c_void_p()
        rc = lib().orc_hier_create(nlev, arr, C.byref(self.ptr))
        if rc != 0:
            raise RuntimeError(f"orc_hier_create failed: {rc}")
        self.n = h.levels[0].A.nrows
        self.A0 = h.levels[0].A

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_hier_free(self.ptr)
            self.ptr = None

    def vcycle(self, b: np.ndarray) -> np.ndarray:
        b = np.asfortranarray(b, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        x = np.zeros_like(b, order="F")
        rc = lib().orc_vcycle(self.ptr, k, _p(b, _F64P), self.n, _p(x, _F64P), self.n)
        assert rc == 0
        return x

    def pcg(self, A, b: np.ndarray, x0: np.ndarray | None = None, itmax: int = 200, tol: float = 1e-8):
        keep = []
        Ac = _csr(A, keep)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        it, fl = C.c_int32(), C.c_int32()
        resvec = np.zeros(itmax + 1)
        rc = lib().orc_pcg(self.ptr, C.byref(Ac), _p(b, _F64P), _p(x, _F64P), itmax, tol, C.byref(it),
                           _p(resvec, _F64P), C.byref(fl))
        assert rc == 0
        return x, it.value, resvec[: it.value + 1], fl.value


def spmv(A, x: np.ndarray) -> np.ndarray:
    keep = []
    Ac = _csr(A, keep)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.nrows)
    lib().orc_spmv(C.byref(Ac), _p(x, _F64P), _p(y, _F64P))
    return y


def lobpcg(A, hier: "Hier | None", X0: np.ndarray, largest: bool = False, reslambda_check: bool = False,
           itmax: int = 100, tol: float = 1e-8, restart_control: int = 7, Y: np.ndarray | None = None):
    """lobpcg(prodA, [], prodM, Y, X0, ...); Y: the constraint block (n x m) or None."""
    keep = []
    Ac = _csr(A, keep)
    X = np.asfortranarray(X0, dtype=np.float64).copy(order="F")
    n, neig = X.shape
    Yf = None if Y is None else np.asfortranarray(Y, dtype=np.float64)
    m = 0 if Yf is None else Yf.shape[1]
    niter, ierr = C.c_int32(), C.c_int32()
    lam = np.zeros(neig)
    rh = np.zeros((itmax + 1, neig), order="F")
    lh = np.zeros((itmax + 1, neig), order="F")
    rc = lib().orc_lobpcg(C.byref(Ac), hier.ptr if hier is not None else None, _p(X, _F64P),
                          _p(Yf, _F64P) if Yf is not None else None, m, n, neig,
                          int(largest), int(reslambda_check), itmax, tol, restart_control, C.byref(niter),
                          _p(lam, _F64P), _p(rh, _F64P), _p(lh, _F64P), C.byref(ierr))
    if rc != 0:
        raise RuntimeError(f"orc_lobpcg failed: {rc}")
    return niter.value, lam, X, rh[: niter.value + 1], lh[: niter.value + 1], ierr.value


def fv_grad(N: int, k: np.ndarray, x: np.ndarray, lam: np.ndarray) -> np.ndarray:
    g = np.empty(N ** 3)
    k, x, lam = (np.ascontiguousarray(a, np.float64) for a in (k, x, lam))
    lib().orc_fv_grad(N, _p(k, _F64P), _p(x, _F64P), _p(lam, _F64P), _p(g, _F64P))
    return g


def sygv(A: np.ndarray, B: np.ndarray):
    m = A.shape[0]
    A = np.asfortranarray(A, np.float64)
    B = np.asfortranarray(B, np.float64)
    w = np.zeros(m)
    V = np.zeros((m, m), order="F")
    rc = lib().orc_sygv(m, _p(A, _F64P), _p(B, _F64P), _p(w, _F64P), _p(V, _F64P))
    if rc != 0:
        raise RuntimeError("orc_sygv: B not positive definite")
    return w, V


# ---------------------------------------------------------------------------------------
# The reference's own classical-prolongation core (oracle/_ref), when built.
# ---------------------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_classical_prol(A, strong: np.ndarray, fcnode: np.ndarray, nc: int, nthreads: int = 1):
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        R.ref_classical_prol.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _I32P, _I32P, C.c_int,
                                         _I32P, C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(_F64P)]
        R.ref_free.argtypes = [_I32P, _I32P, _F64P]
        _ref = R
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    v = np.ascontiguousarray(A.val, np.float64)
    s = np.ascontiguousarray(strong, np.int32)
    fc = np.ascontiguousarray(fcnode, np.int32)
    nt = C.c_int32()
    ip, jp, cp = _I32P(), _I32P(), _F64P()
    rc = _ref.ref_classical_prol(nthreads, A.nrows, _p(rp, _I32P), _p(ci, _I32P), _p(v, _F64P), _p(s, _I32P),
                                 _p(fc, _I32P), nc, C.byref(nt), C.byref(ip), C.byref(jp), C.byref(cp))
    if rc != 0:
        raise RuntimeError(f"reference classical prolongation failed: {rc}")
    n = int(np.sum(fc >= -1))  # rows: the nodes not marked -2
    iat = np.ctypeslib.as_array(ip, shape=(n + 1,)).copy()
    ja = np.ctypeslib.as_array(jp, shape=(max(nt.value, 1),))[: nt.value].copy()
    co = np.ctypeslib.as_array(cp, shape=(max(nt.value, 1),))[: nt.value].copy()
    _ref.ref_free(ip, jp, cp)
    return iat, ja, co


# The reference's own extended+i prolongation core (oracle/_ref/libref_exti.so), when built.
def ref_exti_available() -> bool:
    return os.path.exists(REF_EXTI_LIB)


def ref_exti_prol(A, strong: np.ndarray, fcnode: np.ndarray, nc: int, nthreads: int = 1):
    global _ref_exti
    if _ref_exti is None:
        R = C.CDLL(REF_EXTI_LIB)
        R.ref_exti_prol.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _I32P, _I32P, C.c_int,
                                    _I32P, C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(_F64P)]
        R.ref_exti_free.argtypes = [_I32P, _I32P, _F64P]
        _ref_exti = R
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    v = np.ascontiguousarray(A.val, np.float64)
    s = np.ascontiguousarray(strong, np.int32)
    fc = np.ascontiguousarray(fcnode, np.int32)
    nt = C.c_int32()
    ip, jp, cp = _I32P(), _I32P(), _F64P()
    rc = _ref_exti.ref_exti_prol(nthreads, A.nrows, _p(rp, _I32P), _p(ci, _I32P), _p(v, _F64P), _p(s, _I32P),
                                 _p(fc, _I32P), nc, C.byref(nt), C.byref(ip), C.byref(jp), C.byref(cp))
    if rc != 0:
        raise RuntimeError(f"reference EXTI prolongation failed: {rc}")
    n = int(np.sum(fc >= -1))  # rows: the nodes not marked -2
    iat = np.ctypeslib.as_array(ip, shape=(n + 1,)).copy()
    ja = np.ctypeslib.as_array(jp, shape=(max(nt.value, 1),))[: nt.value].copy()
    co = np.ctypeslib.as_array(cp, shape=(max(nt.value, 1),))[: nt.value].copy()
    _ref_exti.ref_exti_free(ip, jp, cp)
    return iat, ja, co
