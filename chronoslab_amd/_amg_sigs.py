"""ctypes declarations of include/chronos_amg.h."""
import ctypes as C

_VP = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_F64 = C.c_double
_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


class ChrCsr(C.Structure):
    _fields_ = [("nrows", _I32), ("ncols", _I32), ("rowptr", _I32P), ("col", _I32P), ("val", _F64P)]


class ChrLevel(C.Structure):
    _fields_ = [("A", ChrCsr), ("P", ChrCsr), ("G", ChrCsr), ("omega", _F64), ("nupre", _I32),
                ("nupost", _I32)]


class ChrHierStats(C.Structure):
    _fields_ = [("nlev", _I32), ("precision", _I32), ("n", _I64 * 16), ("vcycle_bytes", _F64), ("vcycle_csr_bytes", _F64),
                ("level_bytes", _F64 * 16), ("pcg_iter_bytes", _F64), ("device_bytes", _F64)]


# Every symbol declared in include/chronos_amg.h (checked by tests/test_abi.py).
EXPORTS = {
    "chr_last_error": (C.c_char_p, []),
    "chr_init": (C.c_int, [_I32]),
    "chr_hier_create": (C.c_int, [_I32, C.POINTER(ChrLevel), _I32, C.POINTER(_VP)]),
    "chr_hier_destroy": (C.c_int, [_VP]),
    "chr_hier_get_stats": (C.c_int, [_VP, C.POINTER(ChrHierStats)]),
    "chr_hier_layout": (C.c_int, [_VP, _I32, C.c_char, _I32P, _I32P, C.POINTER(_I64), C.POINTER(C.c_double)]),
    "chr_vcycle_apply": (C.c_int, [_VP, _I32, _VP, _I64, _VP, _I64, _I32]),
    "chr_mat_create": (C.c_int, [C.POINTER(ChrCsr), C.POINTER(_VP)]),
    "chr_mat_destroy": (C.c_int, [_VP]),
    "chr_spmv": (C.c_int, [_VP, _VP, _VP, _I32]),
    "chr_pcg_solve": (C.c_int, [_VP, _VP, _VP, _VP, _I32, _F64, _I32, _I32P, _F64P, _I32P]),
    "chr_pcg_solve_noalloc": (C.c_int, [_VP, _VP, _VP, _VP, _I32, _F64, _I32P, _F64P, _I32P]),
    "chr_lobpcg": (C.c_int, [_VP, _VP, _I32, _I32, _VP, _I64, _VP, _I32, _I64, _I32, _I32, _I32, _F64, _I32, _I32,
                             _I32P, _F64P, _F64P, _F64P, _I32P]),
    "chr_fv_grad": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _I32]),
    "chr_batch_create": (C.c_int, [_I32, _I32P, C.POINTER(C.POINTER(ChrLevel)), _I32, C.POINTER(_VP)]),
    "chr_batch_destroy": (C.c_int, [_VP]),
    "chr_batch_hier": (_VP, [_VP]),
    "chr_batch_vcycle": (C.c_int, [_VP, _I32, _VP, _VP, _I32]),
    "chr_batch_pcg": (C.c_int, [_VP, _I32, _VP, _VP, _I32, _F64, _I32, _I32P, _I32P, _F64P]),
    "chr_afsai": (C.c_int, [C.POINTER(ChrCsr), _I32, _I32, _F64, _I32P, _I32P, _F64P, C.POINTER(_I64), _I32P]),
    "chr_setup_afsai": (C.c_int, [_VP, _I32, _I32P, _I32P, _F64P, _I32, _I32, _F64, _I32P, _I32P, _F64P,
                                  C.POINTER(_I64), _I32P]),
    "chr_fsai_omega": (C.c_int, [C.POINTER(ChrCsr), C.POINTER(ChrCsr), C.c_uint64, _F64P, _F64P]),
    "chr_setup_omega": (C.c_int, [_VP, _I32, _I32P, _I32P, _F64P, _I32P, _I32P, _F64P, C.c_uint64, _F64P,
                                  _F64P]),
    "chr_galerkin": (C.c_int, [C.POINTER(ChrCsr), C.POINTER(ChrCsr), _I32P, _VP, _VP]),
    "chr_setup_emin_open": (C.c_int, [_VP, _I32, _I32, _I32P, _I32P, _F64P, _I32P, _I32P, C.POINTER(_VP)]),
    "chr_setup_emin_apply": (C.c_int, [_VP, _F64P, _F64P]),
    "chr_setup_emin_close": (None, [_VP]),
    "chr_setup_galerkin": (C.c_int, [_VP, _I32, _I32P, _I32P, _F64P, _I32, _I32P, _I32P, _F64P, _I32P, _VP,
                                     _VP]),
    "chr_time_vcycle": (C.c_int, [_VP, _I32, _F64P]),
    "chr_time_spmv": (C.c_int, [_VP, _I32, C.c_char, _I32, _F64P, _F64P]),
}


def declare(lib):
    for name, (res, args) in EXPORTS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
