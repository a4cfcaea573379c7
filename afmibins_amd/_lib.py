"""ctypes binding of libafm.so (the C ABI declared in include/afm.h).

The library is built in place by `__graft_entry__.build()` (or `make -C
afmibins_amd/csrc`).  Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libafm.so")
CSRC = os.path.join(HERE, "csrc")

_D, _U, _I, _L, _P = C.c_double, C.c_uint, C.c_int, C.c_long, C.c_void_p
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_UP = C.POINTER(C.c_uint)


class Geom(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_uint), ("width", C.c_uint), ("depth", C.c_uint)]


# name -> (restype, argtypes); every entry point declared in include/afm.h
SIGNATURES = {
    "afm_ctx_create": (_I, [C.POINTER(Geom), _I, C.POINTER(_P)]),
    "afm_ctx_destroy": (_I, [_P]),
    "afm_ctx_info": (_I, [_P, _UP, _UP, _UP, _UP]),
    "afm_sync": (_I, [_P]),
    "afm_set_stream": (_I, [_P, _P]),
    "afm_own_stream": (_P, [_P]),
    "afm_update_timing": (_I, [_P, _I]),
    "afm_update_timing_read": (_I, [_P, _DP, C.POINTER(_L)]),
    "afm_kernel_timing_read": (_I, [_P, _I, _DP, C.POINTER(_L)]),
    "afm_alloc": (_I, [_P, C.c_ulong, C.POINTER(_P)]),
    "afm_free": (_I, [_P, _P]),
    "afm_copy_to_device": (_I, [_P, _P, _P, C.c_ulong]),
    "afm_copy_to_host": (_I, [_P, _P, _P, C.c_ulong]),
    "afm_last_error": (C.c_char_p, []),
    "afm_delta": (_I, [_P, _P, _L, _P]),
    "afm_delta_host": (_I, [_P, _L, _U, _P]),
    "afm_interp": (_I, [_P, _I, _P, _P, _L, _P]),
    "afm_spread": (_I, [_P, _I, _P, _P, _P, _L, _P]),
    "afm_laplacian_apply": (_I, [_P, _U, _P, _P]),
    "afm_mg_smooth": (_I, [_P, _U, _P, _P, _P, _I, _I]),
    "afm_mg_restrict": (_I, [_P, _U, _P, _P]),
    "afm_mg_prolong": (_I, [_P, _U, _P, _P]),
    "afm_mg_vcycle": (_I, [_P, _P, _I, _I]),
    "afm_mg_solve": (_I, [_P, _P, _D, _I, _I, _I, _I, _I, _IP, _DP]),
    "afm_laplacian3_apply": (_I, [_P, _U, _U, _P, _P]),
    "afm_mg3_vcycle": (_I, [_P, _U, _U, _P, _I, _I, _D]),
    "afm_mg3_restrict": (_I, [_P, _U, _U, _P, _P]),
    "afm_mg3_prolong": (_I, [_P, _U, _U, _P, _P]),
    "afm_mg3_smooth": (_I, [_P, _U, _U, _P, _P, _I]),
    "afm_laplacian3_cg": (_I, [_P, _U, _U, _P, _P, _D, _I, _IP, _DP]),
    "afm_helmholtz3_boundary": (_I, [_P, _U, _I, _D, _P, _P, _P]),
    "afm_projection3_boundary": (_I, [_P, _U, _P, _P, _P, _P, _P, _P]),
    "afm_csr_spmv": (_I, [_P, C.c_ulong, _P, _P, _P, _P, _P]),
    "afm_csr_cg": (_I, [_P, C.c_ulong, _P, _P, _P, _P, _P, _D, _I, _IP, _DP]),
    "afm_csr_smooth": (_I, [_P, C.c_ulong, _P, _P, _P, _P, _P, _I]),
    "afm_axpy": (_I, [_P, C.c_ulong, _D, _P, _P]),
    "afm_mg3_solve": (_I, [_P, _U, _U, _P, _D, _I, _I, _I, _I, _IP, _DP]),
    "afm_ns3_create": (_I, [_P, _U, _D, _D, _D, _D, _P, _P, _P, C.POINTER(_P)]),
    "afm_ns3_step": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "afm_ns3_get_p": (_I, [_P, _P]),
    "afm_ns3_stats": (_I, [_P, _IP, _IP, _IP]),
    "afm_ns3_destroy": (_I, [_P]),
    "afm_div3": (_I, [_P, _U, _P, _P, _P, _P, _P, _P]),
    "afm_grad3": (_I, [_P, _U, _P, _P, _P, _P]),
    "afm_advect3": (_I, [_P, _U, _P, _P, _P, _P, _P, _P]),
    "afm_helmholtz3_apply": (_I, [_P, _U, _I, _D, _P, _P]),
    "afm_helmholtz3_cg": (_I, [_P, _U, _I, _D, _P, _P, _D, _I, _IP, _DP]),
    "afm_interp3": (_I, [_P, _U, _I, _P, _P, _L, _P]),
    "afm_spread3": (_I, [_P, _U, _I, _P, _P, _P, _L, _P]),
    "afm_helmholtz_apply": (_I, [_P, _D, _P, _P]),
    "afm_helmholtz_cg": (_I, [_P, _D, _P, _P, _D, _I, _I, _IP, _DP]),
    "afm_div": (_I, [_P, _P, _P, _P]),
    "afm_grad_sub": (_I, [_P, _P, _P, _P]),
    "afm_advect": (_I, [_P, _P, _P, _P, _P]),
    "afm_ns_create": (_I, [_P, _D, _D, _D, _D, _I, _I, _P, _P, C.POINTER(_P)]),
    "afm_ns_step": (_I, [_P, _P, _P]),
    "afm_ns_project": (_I, [_P, _P, _P]),
    "afm_ns_get_p": (_I, [_P, _P]),
    "afm_ns_stats": (_I, [_P, _IP, _IP, _IP, _IP]),
    "afm_ns_old": (_I, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "afm_ns_last_residual": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "afm_helmholtz_cg_noalloc": (_I, [_P, _D, _P, _P, _D, _I, _I, _IP, _DP]),
    "afm_ns_get_start": (_I, [_P, _I, _P, _IP]),
    "afm_ns_set_start": (_I, [_P, _I, _P]),
    "afm_ns_destroy": (_I, [_P]),
    "afm_ib_create": (_I, [_P, _L, _P, _P, _P, _P, C.POINTER(_P)]),
    "afm_ib_step": (_I, [_P, _I]),
    "afm_ib_markers": (_I, [_P, _P]),
    "afm_ib_grad": (_I, [_P, _I, _I, _L, _DP, _P, _P]),
    "afm_ib_destroy": (_I, [_P]),
}

_lib = None


class AfmError(RuntimeError):
    pass


def load():
    """Load libafm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfmError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           f"or `make -C {CSRC}` (the HIP path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = load().afm_last_error().decode()
        raise AfmError(f"{what} failed ({rc}): {msg}")
    return rc
