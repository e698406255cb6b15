"""Host-side hierarchy producer and problem generators (ctypes over libchronos_setup.so).

The setup mirrors ChronosLab's `cpt_aspAMG` recursion
(sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:68-351): classical (CLA) or diagonal-
dominance (DOM) strength of connection, PMIS, classical, extended+i or BAMG interpolation
(optionally smoothed), adaptive FSAI or Jacobi smoother with omega = min(1, 1.9/lambda_max),
Galerkin coarse operators, and a last level of size <= maxCoarseSZ or at level nLevMax.  The
aFSAI factors, omega and the Galerkin products can be formed on the GPU (same results).

Matrices cross this module as `Csr` (0-based int32 row pointers and columns, float64 values,
columns sorted).  `scipy.sparse` is used only for convenience conversions.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from ._lib import load_setup_lib

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


@dataclasses.dataclass
class Csr:
    nrows: int
    ncols: int
    rowptr: np.ndarray  # int32 [nrows+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.nrows, self.ncols))

    @staticmethod
    def from_scipy(m) -> "Csr":
        m = m.tocsr()
        m.sort_indices()
        m.eliminate_zeros()
        return Csr(m.shape[0], m.shape[1], m.indptr.astype(np.int32), m.indices.astype(np.int32),
                   m.data.astype(np.float64))

    def matvec(self, x: np.ndarray) -> np.ndarray:
        return self.to_scipy() @ x


@dataclasses.dataclass
class Level:
    A: Csr
    P: Csr | None = None
    G: Csr | None = None
    omega: float = 1.0
    nupre: int = 1
    nupost: int = 1


@dataclasses.dataclass
class Hierarchy:
    levels: list  # list[Level]; the last one carries only A (the coarsest operator)
    setup_seconds: dict

    @property
    def nlev(self) -> int:
        return len(self.levels)

    @property
    def coarse(self) -> Csr:
        return self.levels[-1].A

    def describe(self) -> str:
        rows = []
        for l, L in enumerate(self.levels):
            g = L.G.nnz if L.G is not None else 0
            p = L.P.nnz if L.P is not None else 0
            rows.append(f"L{l}: n={L.A.nrows} nnzA={L.A.nnz} nnzG={g} nnzP={p} omega={L.omega:.6f}")
        return "\n".join(rows)


def _lib():
    return load_setup_lib()


def _check(rc: int):
    if rc != 0:
        msg = _lib().chs_last_error().decode()
        raise RuntimeError(f"chronos setup error {rc}: {msg}")


class _Mat:
    """Owns a chs_mat* handle."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib().chs_mat_free(self.ptr)
            self.ptr = None


def _to_handle(m: Csr) -> _Mat:
    lib = _lib()
    out = C.c_void_p()
    rp = np.ascontiguousarray(m.rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(m.col, dtype=np.int32)
    v = np.ascontiguousarray(m.val, dtype=np.float64)
    _check(lib.chs_mat_from_csr(m.nrows, m.ncols, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P),
                                v.ctypes.data_as(_F64P), C.byref(out)))
    return _Mat(out)


def _from_ptr(ptr) -> Csr:
    lib = _lib()
    nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    _check(lib.chs_mat_dims(ptr, C.byref(nr), C.byref(nc), C.byref(nnz)))
    rp = np.empty(nr.value + 1, np.int32)
    ci = np.empty(nnz.value, np.int32)
    v = np.empty(nnz.value, np.float64)
    _check(lib.chs_mat_copy(ptr, rp.ctypes.data_as(_I32P), ci.ctypes.data_as(_I32P), v.ctypes.data_as(_F64P)))
    return Csr(nr.value, nc.value, rp, ci, v)


def _take(ptr) -> Csr:
    m = _from_ptr(ptr)
    _lib().chs_mat_free(ptr)
    return m


# ---------------------------------------------------------------------------------------
# RNG and generators
# ---------------------------------------------------------------------------------------
def rng_uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0, offset: int = 0) -> np.ndarray:
    out = np.empty(n, np.float64)
    _lib().chs_rng_uniform(C.c_uint64(seed), C.c_uint64(stream), C.c_int64(offset), C.c_int64(n),
                           C.c_double(lo), C.c_double(hi), out.ctypes.data_as(_F64P))
    return out


def rng_normal(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    out = np.empty(n, np.float64)
    _lib().chs_rng_normal(C.c_uint64(seed), C.c_uint64(stream), C.c_int64(offset), C.c_int64(n),
                          out.ctypes.data_as(_F64P))
    return out


def poisson7(N: int) -> Csr:
    out = C.c_void_p()
    _check(_lib().chs_gen_poisson7(N, C.byref(out)))
    return _take(out)


def lnk_field(N: int, sigma: float, seed: int) -> np.ndarray:
    out = np.empty(N ** 3, np.float64)
    _lib().chs_gen_lnk(N, C.c_double(sigma), C.c_uint64(seed), out.ctypes.data_as(_F64P))
    return out


def diffusion(N: int, k: np.ndarray) -> Csr:
    k = np.ascontiguousarray(k, dtype=np.float64)
    if k.shape != (N ** 3,):
        raise ValueError("k must have N^3 entries")
    out = C.c_void_p()
    _check(_lib().chs_gen_diffusion(N, k.ctypes.data_as(_F64P), C.byref(out)))
    return _take(out)


def elasticity(ne: int, E: float = 1.0, nu: float = 0.3, clamp: bool = True) -> tuple[Csr, np.ndarray]:
    """Q1 linear elasticity on ne^3 cube elements ((ne+1)^3 nodes x 3 dofs), uniform body force
    -1 in z, the z = 0 face clamped as unit rows (the elasticity configuration).  Returns
    (A, rhs)."""
    n = 3 * (ne + 1) ** 3
    rhs = np.empty(n, np.float64)
    out = C.c_void_p()
    _check(_lib().chs_gen_elasticity(ne, C.c_double(E), C.c_double(nu), int(clamp), C.byref(out),
                                     rhs.ctypes.data_as(_F64P)))
    return _take(out), rhs


def rigid_body_modes(ne: int) -> np.ndarray:
    """The 6 rigid-body modes of the elasticity grid's nodes (mk_rbm_3d.m), n x 6."""
    n = 3 * (ne + 1) ** 3
    R = np.empty((n, 6), np.float64, order="F")
    _lib().chs_gen_rbm(ne, R.ctypes.data_as(_F64P))
    return R


def heterogeneous_diffusion(N: int, sigma: float, seed: int) -> tuple[Csr, np.ndarray]:
    lnk = lnk_field(N, sigma, seed)
    return diffusion(N, np.exp(lnk)), lnk


# ---------------------------------------------------------------------------------------
# Setup components
# ---------------------------------------------------------------------------------------
def soc_cla(A: Csr, tau: float) -> np.ndarray:
    h = _to_handle(A)
    out = np.empty(A.nnz, np.int8)
    _check(_lib().chs_soc_cla(h.ptr, C.c_double(tau), out.ctypes.data_as(C.POINTER(C.c_int8))))
    return out


def soc_dom(A: Csr, tau: float) -> np.ndarray:
    h = _to_handle(A)
    out = np.empty(A.nnz, np.int8)
    _check(_lib().chs_soc_dom(h.ptr, C.c_double(tau), out.ctypes.data_as(C.POINTER(C.c_int8))))
    return out


def agg_coarsen(A: Csr, strong: np.ndarray, seed: int, level: int = 1) -> np.ndarray:
    """Aggressive coarsening (agg_coarsen.m): per node 2 = coarse, 1 = candidate left fine, 0 = fine."""
    h = _to_handle(A)
    strong = np.ascontiguousarray(strong, np.int8)
    kind = np.empty(A.nrows, np.int8)
    _check(_lib().chs_agg_coarsen(h.ptr, strong.ctypes.data_as(C.POINTER(C.c_int8)), C.c_uint64(seed), level,
                                  kind.ctypes.data_as(C.POINTER(C.c_int8))))
    return kind


def prol_bamg(A: Csr, strong: np.ndarray, fcnode: np.ndarray, nc: int, tv: np.ndarray,
              params: SetupParams | None = None, nthreads: int = 0) -> tuple[Csr, np.ndarray]:
    """BAMG least-squares prolongation from the test space tv (n x k); returns (P, c_mark)."""
    p = params or SetupParams()
    h = _to_handle(A)
    strong = np.ascontiguousarray(strong, np.int8)
    fcnode = np.ascontiguousarray(fcnode, np.int32)
    tv = np.asfortranarray(tv, dtype=np.float64)
    bp = _CBamg(p.itmax_vol, p.tol_vol, p.dist_min, p.dist_max, p.maxcond, p.maxrownrm, p.eps_prol)
    mark = np.zeros(A.nrows, np.int32)
    nm = C.c_int32()
    out = C.c_void_p()
    _check(_lib().chs_prol_bamg(h.ptr, strong.ctypes.data_as(C.POINTER(C.c_int8)), fcnode.ctypes.data_as(_I32P),
                                nc, tv.shape[1], tv.ctypes.data_as(_F64P), C.cast(C.pointer(bp), C.c_void_p),
                                nthreads, C.byref(out), mark.ctypes.data_as(_I32P), C.byref(nm)))
    return _take(out), mark


def emin_pattern(A: Csr, P: Csr, fcnode: np.ndarray, ntv: int, params: SetupParams | None = None,
                 nthreads: int = 0) -> Csr:
    """EMIN interpolation pattern (MEX_cptProlPatt.cpp): fine rows of P grown by the strongest
    entries of S^(patt_pow-1) P, at most max(nnzr_max, ntv) per row; values 1."""
    p = params or SetupParams()
    ha, hp = _to_handle(A), _to_handle(P)
    fcnode = np.ascontiguousarray(fcnode, np.int32)
    ep = _cemin(p)
    out = C.c_void_p()
    _check(_lib().chs_emin_pattern(ha.ptr, hp.ptr, fcnode.ctypes.data_as(_I32P), ntv,
                                   C.cast(C.pointer(ep), C.c_void_p), nthreads, C.byref(out)))
    return _take(out)


def emin_improve(A: Csr, P: Csr, patt: Csr, fcnode: np.ndarray, tv: np.ndarray,
                 params: SetupParams | None = None, nthreads: int = 0) -> tuple[Csr, int]:
    """Energy-minimised P over the pattern patt under the test-space constraint of every fine
    row (EMIN_matfree.cpp with solv_emin = 4, prec_emin = 1); returns (P, CG iterations)."""
    p = params or SetupParams()
    ha, hp, ht = _to_handle(A), _to_handle(P), _to_handle(patt)
    fcnode = np.ascontiguousarray(fcnode, np.int32)
    tv = np.asfortranarray(tv, dtype=np.float64)
    ep = _cemin(p)
    ops = _cemin_ops(p.gpu_emin)
    out = C.c_void_p()
    it = C.c_int32()
    _check(_lib().chs_emin_improve_ops(ha.ptr, hp.ptr, ht.ptr, fcnode.ctypes.data_as(_I32P), tv.shape[1],
                                       tv.ctypes.data_as(_F64P), C.cast(C.pointer(ep), C.c_void_p),
                                       C.cast(C.pointer(ops), C.c_void_p), nthreads, C.byref(out), C.byref(it)))
    return _take(out), it.value


def orthonormalize(V: np.ndarray) -> np.ndarray:
    """Q factor (positive R diagonal) of the thin QR of V, as the test-space set-up forms it."""
    Q = np.array(V, dtype=np.float64, order="F", copy=True)
    _lib().chs_orthonormalize(Q.shape[0], Q.shape[1], Q.ctypes.data_as(_F64P))
    return Q


def pmis(A: Csr, strong: np.ndarray, rand01: np.ndarray) -> tuple[np.ndarray, int]:
    h = _to_handle(A)
    strong = np.ascontiguousarray(strong, np.int8)
    rand01 = np.ascontiguousarray(rand01, np.float64)
    fc = np.empty(A.nrows, np.int32)
    nc = C.c_int32()
    _check(_lib().chs_pmis(h.ptr, strong.ctypes.data_as(C.POINTER(C.c_int8)), rand01.ctypes.data_as(_F64P),
                           fc.ctypes.data_as(_I32P), C.byref(nc)))
    return fc, nc.value


def prol_classical(A: Csr, strong: np.ndarray, fcnode: np.ndarray, nc: int, nthreads: int = 0) -> Csr:
    h = _to_handle(A)
    strong = np.ascontiguousarray(strong, np.int8)
    fcnode = np.ascontiguousarray(fcnode, np.int32)
    out = C.c_void_p()
    _check(_lib().chs_prol_classical(h.ptr, strong.ctypes.data_as(C.POINTER(C.c_int8)),
                                     fcnode.ctypes.data_as(_I32P), nc, nthreads, C.byref(out)))
    return _take(out)


def prol_exti(A: Csr, strong: np.ndarray, fcnode: np.ndarray, nc: int, nthreads: int = 0) -> Csr:
    h = _to_handle(A)
    strong = np.ascontiguousarray(strong, np.int8)
    fcnode = np.ascontiguousarray(fcnode, np.int32)
    out = C.c_void_p()
    _check(_lib().chs_prol_exti(h.ptr, strong.ctypes.data_as(C.POINTER(C.c_int8)),
                                fcnode.ctypes.data_as(_I32P), nc, nthreads, C.byref(out)))
    return _take(out)


def afsai(A: Csr, nstep: int = 15, step_size: int = 1, eps: float = 1e-3, nthreads: int = 0) -> tuple[Csr, bool]:
    h = _to_handle(A)
    out = C.c_void_p()
    pd = C.c_int32()
    _check(_lib().chs_afsai(h.ptr, nstep, step_size, C.c_double(eps), nthreads, C.byref(out), C.byref(pd)))
    return _take(out), bool(pd.value)


def fsai_nan_repair(A: Csr, G: Csr) -> tuple[Csr, int]:
    """smoother.m:59-71: rows of G with a NaN diagonal become diag(1/sqrt(a_ii)); returns
    (repaired G, rows repaired)."""
    ha, hg = _to_handle(A), _to_handle(G)
    nf = C.c_int32()
    _check(_lib().chs_fsai_nan_repair(ha.ptr, hg.ptr, C.byref(nf)))
    return _from_ptr(hg.ptr), nf.value


def fsai_omega(A: Csr, G: Csr, seed: int = 1, nthreads: int = 0) -> tuple[float, float]:
    ha, hg = _to_handle(A), _to_handle(G)
    om, lam = C.c_double(), C.c_double()
    _check(_lib().chs_fsai_omega(ha.ptr, hg.ptr, C.c_uint64(seed), nthreads, C.byref(om), C.byref(lam)))
    return om.value, lam.value


def galerkin(A: Csr, P: Csr, nthreads: int = 0) -> Csr:
    ha, hp = _to_handle(A), _to_handle(P)
    out = C.c_void_p()
    _check(_lib().chs_galerkin(ha.ptr, hp.ptr, nthreads, C.byref(out)))
    return _take(out)


def transpose(M: Csr) -> Csr:
    h = _to_handle(M)
    out = C.c_void_p()
    _check(_lib().chs_transpose(h.ptr, C.byref(out)))
    return _take(out)


BC_DIAG = 0
BC_ALL = 1


def bc_fix(A: Csr, seed: int = 1, mode: str = "diag", nthreads: int = 0) -> tuple[Csr, int]:
    """Dirichlet-row rescaling by lambda_max(A) (Lanczos); returns (A_fixed, entries changed).

    mode "diag": diagonal entries equal to 1 become lambda_max/10 (sources/utils/computeAMG.m:5-10);
    mode "all": every stored entry equal to 1 does, as the driver's A(A==1) = lmax/10
    (sources/Drivers/driver_AMG/aspAMG_main.m:89-90)."""
    m = {"diag": BC_DIAG, "all": BC_ALL}[mode]
    h = _to_handle(A)
    out = C.c_void_p()
    nf = C.c_int32()
    _check(_lib().chs_bc_fix(h.ptr, C.c_uint64(seed), m, nthreads, C.byref(out), C.byref(nf)))
    return _take(out), nf.value


RHS_MODES = ("unit_sol", "unit_rhs", "rand_sol", "rand_rhs", "rhs_in")


def rhs_build(A: Csr, mode: str, seed: int = 1, rhs: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray | None]:
    """The driver's right-hand-side modes (sources/Drivers/driver_AMG/aspAMG_main.m:119-134);
    returns (rhs, exact solution or None).

      unit_sol  rhs = A 1                          (solution 1)
      unit_rhs  rhs = 1
      rand_sol  rhs = A u, u ~ U[0,1)               (solution u)
      rand_rhs  rhs_i = u_i 10^(6 (v_i - 0.5)), u, v ~ U[0,1)
      rhs_in    rhs as given (an error when it is missing)

    MATLAB's rand stream is replaced by the shared counter RNG: u is stream 30 and v stream 31
    of `seed`.  The matrix is the one the solver sees (after the Dirichlet rescaling)."""
    mode = mode.lower()
    n = A.nrows
    if mode == "unit_sol":
        x = np.ones(n)
        return A.matvec(x), x
    if mode == "unit_rhs":
        return np.ones(n), None
    if mode == "rand_sol":
        x = rng_uniform(seed, 30, n)
        return A.matvec(x), x
    if mode == "rand_rhs":
        return rng_uniform(seed, 30, n) * 10.0 ** (6.0 * (rng_uniform(seed, 31, n) - 0.5)), None
    if mode == "rhs_in":
        if rhs is None:
            raise ValueError("Missing the expected right-hand side")
        r = np.asarray(rhs, np.float64)
        if r.shape[0] != n:
            raise ValueError("right-hand side length differs from the matrix size")
        return r, None
    raise ValueError(f"Wrong value for rhs_build: {mode}")


@dataclasses.dataclass
class SetupParams:
    """Fields of parm_AMG / parm_COARSEN / parm_SMOOTH read by the classical + FSAI path."""

    nlev_max: int = 10
    max_coarse: int = 1000
    soc_tau: float = 0.25
    nstep: int = 15
    step_size: int = 1
    fsai_eps: float = 1e-3
    nupre: int = 1
    nupost: int = 1
    smoother: str = "afsai_sym"
    prolong: str = "clas"  # parm_PROLONG proltype: "clas", "exti" or "bamg"
    soc: str = "cla"  # parm_COARSEN SoC_type: "cla" or "dom" (tau in soc_tau)
    nthreads: int = 0
    seed: int = 20240607
    verbose: bool = False
    # Compute the aFSAI factors with the GPU kernel (chr_setup_afsai of libchronos_amg)
    # instead of the host restatement; both give the same factor bit for bit.
    gpu_afsai: bool = False
    # Compute the smoother damping's lambda_max by Lanczos on the GPU (chr_setup_omega).
    gpu_omega: bool = False
    # BAMG: finest-level test space (n x k; padded with uniform draws up to ntv columns) and
    # parm_PROLONG's BAMG fields.
    tv: np.ndarray | None = None
    ntv: int = 6
    itmax_vol: int = 100
    tol_vol: float = 0.01
    dist_min: int = 1
    dist_max: int = 10
    maxcond: float = 1e13
    maxrownrm: float = 5.0
    eps_prol: float = 1e-10
    # parm_PROLONG prol_emin: "none", "smooth" or "emin" (EMIN needs prolong="bamg")
    prol_emin: str = "none"
    # parm_PROLONG's EMIN fields (prec_emin = 1 Jacobi, solv_emin = 4 matrix product).
    patt_pow: int = 2
    patt_tau: float = 0.01
    nnzr_max: int = 30
    itmax_emin: int = 10
    energ_tol: float = 1e-3
    condmax_emin: float = 1e10
    # Run EMIN's CG operator (the masked product A P on the pattern) on the GPU
    # (chr_setup_emin_open/apply/close); same bits as the host product.
    gpu_emin: bool = False
    # Form the Galerkin products P^T (A P) on the GPU (chr_setup_galerkin); same bits.
    gpu_galerkin: bool = False
    # parm_COARSEN nl_agg: levels 1..nl_agg use aggressive coarsening (two-stage EXTI).
    nl_agg: int = 0


def _parm_lines(path: str, kinds: str) -> list:
    """The first token of each of the first len(kinds) lines of a parameter file, as a number
    ('f') or a string ('s'); the rest of each line is a comment (textscan of fgetl)."""
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < len(kinds):
        raise ValueError(f"{path}: expected {len(kinds)} parameter lines")
    out = []
    for ln, k in zip(lines, kinds):
        tok = ln.split()
        if not tok:
            raise ValueError(f"{path}: empty parameter line")
        out.append(float(tok[0]) if k == "f" else tok[0])
    return out


def read_default_params(directory: str, fnames: str = "aspAMG.fnames") -> SetupParams:
    """readDefaultParams (sources/utils/readDefaultParams.m:1-30): the file list (one name per
    line: AMG, SMOOTH, TSPACE, COARSEN, PROLONG, FILTER) and the readers read_amg.m,
    read_smoother.m, read_tspace.m, read_coarsen.m, read_prolong.m and read_filter.m, into
    SetupParams.  The reference overwrites nthread/np with maxNumCompThreads; here the OpenMP
    default applies (nthreads = 0).  Options outside the restated set-up are rejected:
    smoother afsai_nsy, strength AFF/CLASNY/ALG, HYBC prolongation, test-space methods other
    than NONE, EMIN variants other than prec_emin 1 / solv_emin 4, updateCF other than 0, and
    active filters (prolongation filter weight below 100, coarse-operator filter tau above 0)."""
    with open(os.path.join(directory, fnames)) as f:
        names = [ln.split()[0] for ln in f.read().splitlines()[:6]]
    if len(names) < 6:
        raise ValueError(f"{fnames}: expected six parameter file names")
    path = lambda i: os.path.join(directory, names[i])
    nlev_max, max_coarse = _parm_lines(path(0), "ff")
    _nthread, nupre, nupost, nstep, step_size, eps, method = _parm_lines(path(1), "ffffffs")
    ntv, _it_tv, _tol_tv, tv_method, _init = _parm_lines(path(2), "fffsf")
    soc_type, tau, nl_agg = _parm_lines(path(3), "sff")
    pr = _parm_lines(path(4), "ssffffffffffffffffffff")
    _np, filt_wgt, _filt_tol, filt_tau, _min_patt = _parm_lines(path(5), "fffff")
    proltype, prol_emin = pr[0].upper(), pr[1].upper()
    if proltype not in ("BAMG", "EXTI", "CLAS", "HYBC") or prol_emin not in ("EMIN", "SMOOTH", "NONE"):
        raise ValueError(f"Wrong proltype flag: {pr[0]} / {pr[1]}")  # read_prolong.m:31-38
    (itmax_vol, tol_vol, dist_min, dist_max, maxcond, maxrownrm, eps_prol, update_cf, patt_pow, patt_tau,
     nnzr_max, itmax_emin, energ_tol, condmax_emin, prec_emin, solv_emin) = pr[3:19]
    if prol_emin == "EMIN" and not (1 <= prec_emin <= 3 and 1 <= solv_emin <= 4):
        raise ValueError(f"Wrong prec_emin / solv_emin: {prec_emin:g} / {solv_emin:g}")  # read_prolong.m:41-56
    tv_ok = ("none", "smoothing", "srqcg", "lobpcg", "lanczos", "ng-srqcg", "ng-lanczos", "arnoldi",
             "block_arnoldi")
    if tv_method.lower() not in tv_ok:
        raise ValueError("Not existing method")  # read_tspace.m:12-25
    unsupported = []
    if method.lower() not in ("afsai_sym", "jacobi"):
        unsupported.append(f"smoother {method}")
    if soc_type.upper() not in ("CLA", "DOM"):
        unsupported.append(f"strength {soc_type}")
    if proltype == "HYBC":
        unsupported.append("HYBC prolongation")
    if tv_method.lower() != "none":
        unsupported.append(f"test-space method {tv_method}")
    if prol_emin == "EMIN" and (prec_emin != 1 or solv_emin != 4):
        unsupported.append("EMIN other than prec_emin 1 / solv_emin 4")
    if update_cf != 0:
        unsupported.append("updateCF")
    if filt_wgt < 100 or filt_tau > 0:
        unsupported.append("prolongation / coarse-operator filtering")
    if unsupported:
        raise ValueError("parameter set outside the restated set-up: " + ", ".join(unsupported))
    return SetupParams(nlev_max=int(nlev_max), max_coarse=int(max_coarse), soc_tau=float(tau), nstep=int(nstep),
                       step_size=int(step_size), fsai_eps=float(eps), nupre=int(nupre), nupost=int(nupost),
                       smoother=method.lower(), prolong=proltype.lower(), soc=soc_type.lower(), ntv=int(ntv),
                       itmax_vol=int(itmax_vol), tol_vol=float(tol_vol), dist_min=int(dist_min),
                       dist_max=int(dist_max), maxcond=float(maxcond), maxrownrm=float(maxrownrm),
                       eps_prol=float(eps_prol), prol_emin=prol_emin.lower(), patt_pow=int(patt_pow),
                       patt_tau=float(patt_tau), nnzr_max=int(nnzr_max), itmax_emin=int(itmax_emin),
                       energ_tol=float(energ_tol), condmax_emin=float(condmax_emin), nl_agg=int(nl_agg))


def read_general(path: str) -> dict:
    """The driver's parm_GENERAL (sources/Drivers/driver_AMG/aspAMG_main.m): ascii_input,
    rhs_build, symmetric, solv_method, itmax, tol, restart."""
    a, rhs, sym, solver, itmax, tol, restart = _parm_lines(path, "ssssfff")
    return {"ascii_input": a.lower() == "true", "rhs_build": rhs.lower(), "symmetric": sym.lower() == "true",
            "solv_method": solver.upper(), "itmax": int(itmax), "tol": float(tol), "restart": int(restart)}


class _CBamg(C.Structure):
    _fields_ = [("itmax_vol", C.c_int32), ("tol_vol", C.c_double), ("dist_min", C.c_int32),
                ("dist_max", C.c_int32), ("maxcond", C.c_double), ("maxrownrm", C.c_double),
                ("eps_prol", C.c_double)]


class _CEmin(C.Structure):
    _fields_ = [("patt_pow", C.c_int32), ("patt_tau", C.c_double), ("nnzr_max", C.c_int32),
                ("itmax", C.c_int32), ("energ_tol", C.c_double), ("condmax", C.c_double)]


def _cemin(p: SetupParams) -> _CEmin:
    return _CEmin(p.patt_pow, p.patt_tau, p.nnzr_max, p.itmax_emin, p.energ_tol, p.condmax_emin)


class _CEminOps(C.Structure):
    _fields_ = [("open", C.c_void_p), ("apply", C.c_void_p), ("close", C.c_void_p), ("ctx", C.c_void_p)]


def _cemin_ops(gpu: bool) -> _CEminOps:
    if not gpu:
        return _CEminOps(None, None, None, None)
    from ._lib import load_amg_lib

    g = load_amg_lib()
    addr = lambda f: C.cast(f, C.c_void_p).value
    return _CEminOps(addr(g.chr_setup_emin_open), addr(g.chr_setup_emin_apply), addr(g.chr_setup_emin_close), None)


class _CParams(C.Structure):
    _fields_ = [("nlev_max", C.c_int32), ("max_coarse", C.c_int32), ("soc_tau", C.c_double),
                ("nstep", C.c_int32), ("step_size", C.c_int32), ("fsai_eps", C.c_double),
                ("nupre", C.c_int32), ("nupost", C.c_int32), ("smoother", C.c_int32), ("prolong", C.c_int32),
                ("nthreads", C.c_int32), ("seed", C.c_uint64), ("verbose", C.c_int32),
                ("afsai_fn", C.c_void_p), ("afsai_ctx", C.c_void_p), ("omega_fn", C.c_void_p),
                ("omega_ctx", C.c_void_p), ("soc_type", C.c_int32), ("ntv", C.c_int32), ("ntv0", C.c_int32),
                ("tv", _F64P), ("bamg", _CBamg), ("prol_smooth", C.c_int32), ("galerkin_fn", C.c_void_p),
                ("galerkin_ctx", C.c_void_p), ("nl_agg", C.c_int32), ("emin", _CEmin), ("emin_kp", _CEminOps)]


def build(A: Csr, params: SetupParams | None = None) -> Hierarchy:
    p = params or SetupParams()
    if p.nupre != 1 or p.nupost != 1:
        raise ValueError("the V-cycle supports nupre = nupost = 1 (the reference's parameter files)")
    sm = {"afsai_sym": 0, "jacobi": 1}[p.smoother.lower()]
    fn = ofn = gfn = None
    if p.gpu_afsai or p.gpu_omega or p.gpu_galerkin:
        from ._lib import load_amg_lib

        g = load_amg_lib()
        fn = C.cast(g.chr_setup_afsai, C.c_void_p).value if p.gpu_afsai else None
        ofn = C.cast(g.chr_setup_omega, C.c_void_p).value if p.gpu_omega else None
        gfn = C.cast(g.chr_setup_galerkin, C.c_void_p).value if p.gpu_galerkin else None
    pr = {"clas": 0, "exti": 1, "bamg": 2}[p.prolong.lower()]
    soc = {"cla": 0, "dom": 1}[p.soc.lower()]
    tv = None
    if p.tv is not None:
        tv = np.asfortranarray(p.tv, dtype=np.float64).reshape(A.nrows, -1, order="F")
    cp = _CParams(p.nlev_max, p.max_coarse, p.soc_tau, p.nstep, p.step_size, p.fsai_eps, p.nupre,
                  p.nupost, sm, pr, p.nthreads, p.seed, int(p.verbose), fn, None, ofn, None, soc, p.ntv,
                  0 if tv is None else tv.shape[1], None if tv is None else tv.ctypes.data_as(_F64P),
                  _CBamg(p.itmax_vol, p.tol_vol, p.dist_min, p.dist_max, p.maxcond, p.maxrownrm, p.eps_prol),
                  {"none": 0, "smooth": 1, "emin": 2}[p.prol_emin.lower()], gfn, None, p.nl_agg,
                  _cemin(p), _cemin_ops(p.gpu_emin))
    lib = _lib()
    ha = _to_handle(A)
    hh = C.c_void_p()
    _check(lib.chs_build(ha.ptr, C.byref(cp), C.byref(hh)))
    try:
        nlev = lib.chs_hier_nlev(hh)
        levels = []
        for l in range(nlev):
            last = l == nlev - 1
            Al = _from_ptr(lib.chs_hier_mat(hh, l, b"A"))
            if last:
                levels.append(Level(Al))
            else:
                levels.append(Level(Al, _from_ptr(lib.chs_hier_mat(hh, l, b"P")),
                                    _from_ptr(lib.chs_hier_mat(hh, l, b"G")),
                                    lib.chs_hier_omega(hh, l), p.nupre, p.nupost))
        secs = {k: lib.chs_hier_setup_seconds(hh, i)
                for i, k in enumerate(["smoother", "coarsen", "prolong", "galerkin", "total"])}
    finally:
        lib.chs_hier_free(hh)
    # The finest level's A is the operator the caller passed (possibly BC-modified).
    return Hierarchy(levels, secs)


# ---------------------------------------------------------------------------------------
# On-disk format: the reference's Ext_bin binary files (little-endian).
#   matrix (sources/utils/debug_stuff/dump_chronos/DumpChronos.cpp:27-35, read by
#   read_ExtBinMat.m:1-36): int64 nrows, [int64 ncols when rectangular], int64 nterm, then
#   nterm records {int64 i, int64 j, float64 v} with 1-based i, j;
#   vector (read_ExtBinVec.m:1-21): int64 N, int64 M, then N*M float64, row by row.
# A hierarchy is a directory: A_<l>.Ext_bin for every level, P_<l>.Ext_bin (rectangular
# header) and G_<l>.Ext_bin for every level but the last, and hierarchy.txt with one line
# per level "l n omega nupre nupost" (omega written with 17 significant digits, exact).
# ---------------------------------------------------------------------------------------
_EXT_REC = np.dtype([("i", "<i8"), ("j", "<i8"), ("v", "<f8")])


def write_ext_bin_mat(path: str, M: Csr, rect: bool = False) -> None:
    rows = np.repeat(np.arange(M.nrows, dtype=np.int64), np.diff(M.rowptr.astype(np.int64)))
    rec = np.empty(M.nnz, _EXT_REC)
    rec["i"] = rows + 1
    rec["j"] = M.col.astype(np.int64) + 1
    rec["v"] = M.val
    head = [M.nrows, M.ncols, M.nnz] if rect else [M.nrows, M.nnz]
    if not rect and M.nrows != M.ncols:
        raise ValueError("a rectangular matrix needs rect=True")
    with open(path, "wb") as f:
        np.asarray(head, "<i8").tofile(f)
        rec.tofile(f)


def read_ext_bin_mat(path: str, rect: bool = False) -> Csr:
    with open(path, "rb") as f:
        nh = 3 if rect else 2
        head = np.fromfile(f, "<i8", count=nh)
        if head.size != nh:
            raise ValueError(f"{path}: truncated header")
        nrows = int(head[0])
        ncols = int(head[1]) if rect else nrows
        nterm = int(head[-1])
        rec = np.fromfile(f, _EXT_REC, count=nterm)
    if rec.size != nterm:
        raise ValueError(f"{path}: {rec.size} of {nterm} entries")
    i, j = rec["i"] - 1, rec["j"] - 1
    if nterm and (i.min() < 0 or i.max() >= nrows or j.min() < 0 or j.max() >= ncols):
        raise ValueError(f"{path}: index out of range")
    order = np.lexsort((j, i))
    i, j, v = i[order], j[order], rec["v"][order]
    if nterm and np.any((np.diff(i) == 0) & (np.diff(j) == 0)):
        raise ValueError(f"{path}: duplicate entries")
    rp = np.zeros(nrows + 1, np.int64)
    np.cumsum(np.bincount(i, minlength=nrows), out=rp[1:])
    return Csr(nrows, ncols, rp.astype(np.int32), j.astype(np.int32), v.astype(np.float64))


def write_ext_bin_vec(path: str, X: np.ndarray) -> None:
    X = np.asarray(X, np.float64)
    X2 = X.reshape(X.shape[0], -1)
    with open(path, "wb") as f:
        np.asarray([X2.shape[0], X2.shape[1]], "<i8").tofile(f)
        np.ascontiguousarray(X2, "<f8").tofile(f)


def read_ext_bin_vec(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        n, m = (int(t) for t in np.fromfile(f, "<i8", count=2))
        data = np.fromfile(f, "<f8", count=n * m)
    if data.size != n * m:
        raise ValueError(f"{path}: truncated")
    return data.reshape(n, m) if m != 1 else data


def save_hierarchy(directory: str, hier: Hierarchy) -> None:
    """Writes every level of hier in the Ext_bin format (see above)."""
    os.makedirs(directory, exist_ok=True)
    lines = []
    for l, L in enumerate(hier.levels):
        write_ext_bin_mat(os.path.join(directory, f"A_{l}.Ext_bin"), L.A)
        if L.P is not None:
            write_ext_bin_mat(os.path.join(directory, f"P_{l}.Ext_bin"), L.P, rect=True)
            write_ext_bin_mat(os.path.join(directory, f"G_{l}.Ext_bin"), L.G)
        lines.append(f"{l} {L.A.nrows} {L.omega!r} {L.nupre} {L.nupost}")
    with open(os.path.join(directory, "hierarchy.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


def load_hierarchy(directory: str) -> Hierarchy:
    """Reads a hierarchy written by save_hierarchy (values and patterns bit for bit)."""
    with open(os.path.join(directory, "hierarchy.txt")) as f:
        rows = [ln.split() for ln in f if ln.strip()]
    levels = []
    for k, (l, n, omega, nupre, nupost) in enumerate(rows):
        if int(l) != k:
            raise ValueError("hierarchy.txt: levels out of order")
        A = read_ext_bin_mat(os.path.join(directory, f"A_{k}.Ext_bin"))
        if A.nrows != int(n):
            raise ValueError(f"level {k}: A has {A.nrows} rows, manifest says {n}")
        if k < len(rows) - 1:
            P = read_ext_bin_mat(os.path.join(directory, f"P_{k}.Ext_bin"), rect=True)
            G = read_ext_bin_mat(os.path.join(directory, f"G_{k}.Ext_bin"))
            levels.append(Level(A, P, G, float(omega), int(nupre), int(nupost)))
        else:
            levels.append(Level(A, omega=float(omega), nupre=int(nupre), nupost=int(nupost)))
    return Hierarchy(levels, {})
