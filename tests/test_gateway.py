"""The reference-side MEX gateway's walk over the hierarchy struct list, restated in Python.

`amg_hrc` builds the struct list with the fields cpt_aspAMG.m stores
(sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:82-131,288-302,347): per non-last level
level, lmax_A, isPosDef, A (levels > 1 only), P, Pf, S, S_patt, TV, fcnode, omega, Snnz, nupre,
nupost, the closures Minv1/Minv2 (which alone hold the FSAI factor, as smootherOp.left) and
next; on the last level L, S, symm, A and next = 0.  `gateway_levels` reads it the way the
gateway sketched for the reference side does (field G if present, else smootherOp.left from
the Minv1 closure's workspace; computeAMG's handle M supplies AMG_prec and the modified A from
its own workspace).  The CPU tests check that the walk reads only fields that exist and hands
chr_hier_create the hierarchy's own levels; the GPU test applies the V-cycle through them."""
import ctypes as C

import numpy as np
import pytest

from chronoslab_amd import hierarchy as H


class Closure:
    """An anonymous function handle: functions(fh).workspace{1} holds the captured variables."""

    def __init__(self, **captured):
        self.workspace = [dict(captured)]


def _transpose(M):
    return H.Csr.from_scipy(M.to_scipy().T)


def amg_hrc(hier, patched=False):
    """The struct list cpt_aspAMG.m returns for this hierarchy (level 1 first)."""
    nxt = 0
    for l in range(hier.nlev - 1, -1, -1):
        L = hier.levels[l]
        s = {"level": l + 1, "lmax_A": 0.0, "isPosDef": True}
        if l == hier.nlev - 1:
            s.update(L=None, S=None, symm=True, A=L.A, next=0)  # cpt_aspAMG.m:103-131
        else:
            sm = {"left": L.G, "right": _transpose(L.G), "omega": L.omega}  # smoother.m:83-85
            if l > 0:
                s["A"] = L.A  # cpt_aspAMG.m:288-290: level 1 keeps no A
            s.update(P=L.P, Pf=L.P, S=None, S_patt=None, TV=None, fcnode=None, omega=L.omega,
                     Snnz=2 * L.G.nnz, nupre=L.nupre, nupost=L.nupost,
                     Minv1=Closure(smootherOp=sm), Minv2=Closure(smootherOp=sm), next=nxt)
            if patched:
                s["G"] = L.G  # the one-line patch beside cpt_aspAMG.m:301
        nxt = s
    return nxt


def _field(s, name):
    if name not in s:
        raise KeyError(f"hierarchy level has no field {name}")
    return s[name]


def _captured(fh, name):
    ws = fh.workspace[0]
    if name not in ws:
        raise KeyError(f"the handle captures no {name}")
    return ws[name]


def _fsai_factor(L):
    if "G" in L:
        return L["G"]
    return _field(_captured(_field(L, "Minv1"), "smootherOp"), "left")


def gateway_levels(prec, A1):
    """The 'create' command of the gateway: a list of H.Level for chr_hier_create."""
    if isinstance(prec, Closure):  # M = @(x) AMG_Vcycle(AMG_prec, A, x) from computeAMG.m:49
        prec, A1 = _captured(prec, "AMG_prec"), _captured(prec, "A")
    out, L, l = [], prec, 0
    while isinstance(L, dict):
        nxt = _field(L, "next")
        A = A1 if l == 0 else _field(L, "A")
        if isinstance(nxt, dict):
            out.append(H.Level(A, _field(L, "P"), _fsai_factor(L), _field(L, "omega"), _field(L, "nupre"),
                               _field(L, "nupost")))
        else:
            out.append(H.Level(A))
        L, l = nxt, l + 1
    return out


def _same(a, b):
    return a.nrows == b.nrows and a.ncols == b.ncols and np.array_equal(a.rowptr, b.rowptr) and \
        np.array_equal(a.col, b.col) and np.array_equal(a.val, b.val)


@pytest.fixture(scope="module")
def case():
    A, _ = H.heterogeneous_diffusion(12, 1.5, 21)
    Afix, _ = H.bc_fix(A, 1)
    return A, Afix, H.build(Afix)


@pytest.mark.parametrize("how", ["struct", "handle", "patched"])
def test_gateway_walk_yields_the_hierarchy(case, how):
    A, Afix, hier = case
    S = amg_hrc(hier, patched=how == "patched")
    lev = gateway_levels(Closure(AMG_prec=S, A=Afix), None) if how == "handle" else gateway_levels(S, Afix)
    assert len(lev) == hier.nlev
    for a, b in zip(lev, hier.levels):
        assert _same(a.A, b.A)
        if b.P is not None:
            assert _same(a.P, b.P) and _same(a.G, b.G)
            assert (a.omega, a.nupre, a.nupost) == (b.omega, b.nupre, b.nupost)


def test_reference_struct_has_no_smoother_field(case):
    """cpt_aspAMG.m keeps the factor only in the closures: a walk that reads smootherOp as a
    field of the level fails, and one that finds neither G nor the closure's capture fails."""
    _, Afix, hier = case
    S = amg_hrc(hier)
    assert "smootherOp" not in S and "G" not in S
    with pytest.raises(KeyError):
        _field(S, "smootherOp")
    S["Minv1"] = Closure()
    with pytest.raises(KeyError, match="captures no smootherOp"):
        gateway_levels(S, Afix)


@pytest.mark.gpu
def test_gateway_levels_vcycle_equals_oracle(gpu, case):
    import oracle
    from chronoslab_amd._amg_sigs import ChrLevel

    A, Afix, hier = case
    lev = gateway_levels(Closure(AMG_prec=amg_hrc(hier), A=Afix), None)
    keep = []
    arr = (ChrLevel * len(lev))()
    for l, L in enumerate(lev):
        arr[l] = ChrLevel(gpu._csr(L.A, keep), gpu._csr(L.P, keep), gpu._csr(L.G, keep), L.omega, L.nupre, L.nupost)
    h = C.c_void_p()
    gpu._check(gpu._lib().chr_hier_create(len(lev), arr, gpu.FP64, C.byref(h)), "chr_hier_create")
    try:
        n = Afix.nrows
        B = np.asfortranarray(np.stack([H.rng_uniform(4, 40 + c, n, -1, 1) for c in range(3)], 1))
        X = np.empty_like(B, order="F")
        gpu._check(gpu._lib().chr_vcycle_apply(h, 3, B.ctypes.data, n, X.ctypes.data, n, gpu.PTR_HOST))
        Xo = oracle.Hier(hier).vcycle(B)
        assert np.linalg.norm(X - Xo) <= 1e-12 * np.linalg.norm(Xo)
    finally:
        gpu._lib().chr_hier_destroy(h)
