"""Loading of the native libraries built in-tree (see chronoslab_amd/build.py).

libchronos_setup.so  host (CPU) hierarchy producer, include/chronos_setup.h
libchronos_amg.so    HIP solve phase for gfx950, include/chronos_amg.h

Both are loaded from this package directory only; a missing library is an error, never a
silent fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
_lock = threading.Lock()
_setup = None
_amg = None

SETUP_LIB = os.path.join(_HERE, "libchronos_setup.so")
AMG_LIB = os.path.join(_HERE, "libchronos_amg.so")

_VP = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_F64 = C.c_double
_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)
_I8P = C.POINTER(C.c_int8)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def load_setup_lib():
    global _setup
    with _lock:
        if _setup is None:
            if not os.path.exists(SETUP_LIB):
                raise RuntimeError(f"{SETUP_LIB} is not built: run chronoslab_amd.build.build_all()")
            lib = C.CDLL(SETUP_LIB)
            _sig(lib, "chs_last_error", C.c_char_p, [])
            _sig(lib, "chs_mat_from_csr", C.c_int, [_I32, _I32, _I32P, _I32P, _F64P, C.POINTER(_VP)])
            _sig(lib, "chs_mat_dims", C.c_int, [_VP, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64)])
            _sig(lib, "chs_mat_copy", C.c_int, [_VP, _I32P, _I32P, _F64P])
            _sig(lib, "chs_mat_free", None, [_VP])
            _sig(lib, "chs_rng_u64", _U64, [_U64, _U64, _U64])
            _sig(lib, "chs_rng_uniform", None, [_U64, _U64, _I64, _I64, _F64, _F64, _F64P])
            _sig(lib, "chs_rng_normal", None, [_U64, _U64, _I64, _I64, _F64P])
            _sig(lib, "chs_gen_poisson7", C.c_int, [_I32, C.POINTER(_VP)])
            _sig(lib, "chs_gen_diffusion", C.c_int, [_I32, _F64P, C.POINTER(_VP)])
            _sig(lib, "chs_gen_lnk", None, [_I32, _F64, _U64, _F64P])
            _sig(lib, "chs_gen_elasticity", C.c_int, [_I32, _F64, _F64, _I32, C.POINTER(_VP), _F64P])
            _sig(lib, "chs_gen_rbm", None, [_I32, _F64P])
            _sig(lib, "chs_soc_cla", C.c_int, [_VP, _F64, _I8P])
            _sig(lib, "chs_soc_dom", C.c_int, [_VP, _F64, _I8P])
            _sig(lib, "chs_agg_coarsen", C.c_int, [_VP, _I8P, _U64, _I32, _I8P])
            _sig(lib, "chs_prol_bamg", C.c_int, [_VP, _I8P, _I32P, _I32, _I32, _F64P, _VP, C.c_int,
                                                 C.POINTER(_VP), _I32P, C.POINTER(_I32)])
            _sig(lib, "chs_orthonormalize", C.c_int, [C.c_int64, _I32, _F64P])
            _sig(lib, "chs_emin_pattern", C.c_int, [_VP, _VP, _I32P, _I32, _VP, C.c_int, C.POINTER(_VP)])
            _sig(lib, "chs_emin_improve", C.c_int, [_VP, _VP, _VP, _I32P, _I32, _F64P, _VP, C.c_int,
                                                    C.POINTER(_VP), C.POINTER(_I32)])
            _sig(lib, "chs_emin_improve_ops", C.c_int, [_VP, _VP, _VP, _I32P, _I32, _F64P, _VP, _VP, C.c_int,
                                                        C.POINTER(_VP), C.POINTER(_I32)])
            _sig(lib, "chs_pmis", C.c_int, [_VP, _I8P, _F64P, _I32P, C.POINTER(_I32)])
            _sig(lib, "chs_prol_classical", C.c_int, [_VP, _I8P, _I32P, _I32, C.c_int, C.POINTER(_VP)])
            _sig(lib, "chs_prol_exti", C.c_int, [_VP, _I8P, _I32P, _I32, C.c_int, C.POINTER(_VP)])
            _sig(lib, "chs_afsai", C.c_int, [_VP, _I32, _I32, _F64, C.c_int, C.POINTER(_VP), C.POINTER(_I32)])
            _sig(lib, "chs_fsai_nan_repair", C.c_int, [_VP, _VP, C.POINTER(_I32)])
            _sig(lib, "chs_fsai_omega", C.c_int, [_VP, _VP, _U64, C.c_int, C.POINTER(_F64), C.POINTER(_F64)])
            _sig(lib, "chs_transpose", C.c_int, [_VP, C.POINTER(_VP)])
            _sig(lib, "chs_galerkin", C.c_int, [_VP, _VP, C.c_int, C.POINTER(_VP)])
            _sig(lib, "chs_bc_fix", C.c_int, [_VP, _U64, _I32, C.c_int, C.POINTER(_VP), C.POINTER(_I32)])
            _sig(lib, "chs_build", C.c_int, [_VP, _VP, C.POINTER(_VP)])
            _sig(lib, "chs_sell_apply", C.c_int,
                 [_I32, _I32, _I32P, _I32P, _F64P, _I32, _F64P, _F64P, _I32P, C.POINTER(_I64)])
            _sig(lib, "chs_pack_apply", C.c_int,
                 [_I32, _I32, _I32P, _I32P, _F64P, _I32, _I32, _F64P, _F64P, _I32P])
            _sig(lib, "chs_dia_apply", C.c_int,
                 [_I32, _I32, _I32P, _I32P, _F64P, _I32, _F64P, _F64P, _I32P, _I32P])
            _sig(lib, "chs_sygv", C.c_int, [_I32, _F64P, _F64P, _F64P, _F64P])
            _sig(lib, "chs_cond_sym", _F64, [_I32, _F64P])
            _sig(lib, "chs_hier_nlev", C.c_int, [_VP])
            _sig(lib, "chs_hier_mat", _VP, [_VP, _I32, C.c_char])
            _sig(lib, "chs_hier_omega", _F64, [_VP, _I32])
            _sig(lib, "chs_hier_setup_seconds", _F64, [_VP, _I32])
            _sig(lib, "chs_hier_free", None, [_VP])
            _setup = lib
    return _setup


def load_amg_lib():
    """The HIP solve-phase library.  Raises when it is not built.

    PyTorch wheels for ROCm carry their own HIP/HSA runtime (torch/lib/libamdhip64.so, same
    soname as the system one).  Two runtimes in one process cannot both enumerate the GPU,
    so torch, when installed, is imported first: the library then binds to the runtime torch
    already loaded, and torch tensors and this library share one device context."""
    global _amg
    with _lock:
        if _amg is None:
            if not os.path.exists(AMG_LIB):
                raise RuntimeError(f"{AMG_LIB} is not built: run chronoslab_amd.build.build_all()")
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            from . import _amg_sigs

            lib = C.CDLL(AMG_LIB)
            _amg_sigs.declare(lib)
            _amg = lib
    return _amg
