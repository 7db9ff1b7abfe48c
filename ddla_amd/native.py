"""ctypes binding of libddla_hip.so (the C ABI declared in include/ddla_native.h).

The library is required: importing this module raises if the shared object is missing
or lacks a declared entry point, so no caller can silently run without the HIP code.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddla_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ddla_native.h")

# flags (DDLA/Misc/Flags/*.cs)
NO_TRANSPOSE, TRANSPOSE = 0x00, 0x08
LOWER, UPPER = 0xC0, 0x60
NONUNIT_DIAG, UNIT_DIAG = 0x000, 0x100
LEFT, RIGHT = 0, 1
PREC_FP64, PREC_FP32, PREC_FP16, PREC_BF16 = 0, 1, 2, 3

SUCCESS = 0
ERR_INVALID_ARG, ERR_HIP, ERR_NO_DEVICE, ERR_ALLOC, ERR_UNSUPPORTED, ERR_NOT_CONVERGED = (
    -1, -2, -3, -4, -5, -6)

i32, i64, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
cint = ctypes.c_int


class MpirOpts(ctypes.Structure):
    _fields_ = [("factor_prec", i32), ("max_iters", i32), ("tol", ctypes.c_double),
                ("gmres", i32), ("terms", i32), ("pivoting", i32), ("reserved", i32),
                ("x_ref", vp), ("ld_xref", i64)]


class MpirInfo(ctypes.Structure):
    _fields_ = [("status", i32), ("iters", i32), ("inner_iters", i32), ("fallback", i32),
                ("backward_error", ctypes.c_double), ("factor_ms", ctypes.c_double),
                ("refine_ms", ctypes.c_double), ("forward_error", ctypes.c_double)]


# name -> argtypes (restype is always int)
SIGNATURES = {
    "ddla_init": [cint],
    "ddla_finalize": [],
    "ddla_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "ddla_set_stream": [vp],
    "ddla_get_stream": [ctypes.POINTER(vp)],
    "ddla_sync": [],
    "ddla_version": [],
    "ddla_set_wait_limit": [u64],
    "ddla_profile_enable": [cint],
    "ddla_profile_reset": [],
    "ddla_profile_only": [ctypes.c_char_p],
    "ddla_profile_read": [cint, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(i64),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)],
    "ddla_malloc": [ctypes.POINTER(vp), ctypes.c_size_t],
    "ddla_free": [vp],
    "ddla_memcpy_h2d": [vp, vp, ctypes.c_size_t],
    "ddla_memcpy_d2h": [vp, vp, ctypes.c_size_t],
    "ddla_memcpy_d2d": [vp, vp, ctypes.c_size_t],
    "ddla_dmat_set": [i64, i64, vp, i64, i64, vp, i64],
    "ddla_dmat_get": [i64, i64, vp, i64, vp, i64, i64],
    "ddla_drandm_dev": [i64, i64, vp, i64, u64],
    "ddla_dgemm": [cint, cint, i64, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64, i64],
    "ddla_dgemm_dev": [cint, cint, i64, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64, i64],
    "ddla_dgetrf_dev": [i64, vp, i64, vp],
    "ddla_dgetrf_noalloc_dev": [i64, vp, i64, vp],
    "ddla_dgetrs_dev": [cint, i64, i64, vp, i64, vp, vp, i64],
    "ddla_dgetrf": [i64, vp, i64, i64, vp],
    "ddla_dgetrs": [cint, i64, i64, vp, i64, i64, vp, vp, i64, i64],
    "ddla_dlaswp_dev": [cint, cint, i64, i64, vp, i64, vp, i64],
    "ddla_dgetrf_nopiv_dev": [i64, vp, i64],
    "ddla_dgetrs_nopiv_dev": [cint, i64, i64, vp, i64, vp, i64],
    "ddla_dpotrf_dev": [cint, i64, vp, i64],
    "ddla_dpotrs_dev": [cint, i64, i64, vp, i64, vp, i64],
    "ddla_dpotrf": [cint, i64, vp, i64, i64],
    "ddla_dpotrs": [cint, i64, i64, vp, i64, i64, vp, i64, i64],
    "ddla_dldlt_dev": [cint, i64, vp, i64],
    "ddla_dldlt_solve_dev": [cint, i64, i64, vp, i64, vp, i64],
    "ddla_dldlt": [cint, i64, vp, i64, i64],
    "ddla_dldlt_solve": [cint, i64, i64, vp, i64, i64, vp, i64, i64],
    "ddla_dcontract_dev": [ctypes.c_char_p, ctypes.c_double, vp, cint, vp, vp, vp, cint, vp, vp,
                           ctypes.c_double, vp, cint, vp, vp],
    "ddla_qr_t_rows": [i64, i64],
    "ddla_dgeqrf_dev": [i64, i64, vp, i64, vp, i64],
    "ddla_dqr_apply_dev": [cint, i64, i64, vp, i64, vp, i64, i64, vp, i64],
    "ddla_dqr_solve_dev": [i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64],
    "ddla_dgeqrf": [i64, i64, vp, i64, i64, vp, i64, i64],
    "ddla_dqr_solve": [i64, i64, i64, vp, i64, i64, vp, i64, i64, vp, i64, i64, vp, i64, i64],
    "ddla_dtrsm_dev": [cint, cint, cint, cint, i64, i64, dp, vp, i64, vp, i64],
    "ddla_dsyrk_dev": [cint, cint, i64, i64, dp, vp, i64, dp, vp, i64],
    "ddla_damaxv_dev": [i64, vp, i64, ctypes.POINTER(i64)],
    # level-3 BLAS with BLIS argument lists (host and strided-device forms)
    "ddla_dgemmt": [cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64, i64],
    "ddla_dgemmt_dev": [cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64,
                        i64],
    "ddla_dsyrk": [cint, cint, i64, i64, dp, vp, i64, i64, dp, vp, i64, i64],
    "ddla_dsyr2k": [cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64, i64],
    "ddla_dsyr2k_dev": [cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64,
                        i64],
    "ddla_dsymm": [cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp, i64,
                   i64],
    "ddla_dsymm_dev": [cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp,
                       i64, i64],
    "ddla_dtrmm": [cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64],
    "ddla_dtrmm_dev": [cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64],
    "ddla_dtrmm3": [cint, cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp, vp,
                    i64, i64],
    "ddla_dtrmm3_dev": [cint, cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64, dp,
                        vp, i64, i64],
    "ddla_dtrsm": [cint, cint, cint, cint, i64, i64, dp, vp, i64, i64, vp, i64, i64],
    "ddla_dgesv_mpir_dev": [i64, i64, vp, i64, vp, i64, vp, i64, ctypes.POINTER(MpirOpts),
                            ctypes.POINTER(MpirInfo)],
    "ddla_dposv_mpir_dev": [i64, i64, vp, i64, vp, i64, vp, i64, ctypes.POINTER(MpirOpts),
                            ctypes.POINTER(MpirInfo)],
    "ddla_sgetrf_mp_dev": [i64, vp, i64, cint, cint, vp],
    "ddla_sgetf2_dev": [i64, i64, vp, i64, vp],
    "ddla_dgetf2_dev": [i64, i64, vp, i64, vp],
    "ddla_batch_create": [ctypes.POINTER(vp), cint, i64, vp, vp, vp, vp, vp, cint],
    "ddla_batch_execute": [vp],
    "ddla_batch_describe": [vp, ctypes.c_char_p, ctypes.c_size_t],
    "ddla_batch_destroy": [vp],
}

_lib = None
_finalized = False


def finalized() -> bool:
    """True once the interpreter is exiting and the library context has been released: objects
    that still hold device memory then leave it to the process's end instead of calling into
    HIP, whose runtime (or a profiler wrapped around it) may already be torn down."""
    return _finalized


def finalize(at_exit: bool = False):
    """Releases the library context (streams, events, workspaces; ddla_init sets it up again).
    Registered with atexit (at_exit=True) when the library is loaded, so the release runs while
    the HIP runtime is still up, before the interpreter's teardown collects the objects left."""
    global _finalized
    if _lib is None or _finalized:
        return
    if at_exit:
        _finalized = True
    _lib.ddla_finalize()


def header_symbols() -> list[str]:
    """Entry points declared in include/ddla_native.h (for the export check)."""
    import re
    txt = open(HEADER).read()
    return re.findall(r"^int (ddla_\w+)\(", txt, flags=re.M)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m ddla_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)  # AttributeError if the export is missing
        fn.argtypes = args
        fn.restype = cint
    _lib = L
    import atexit
    atexit.register(finalize, True)
    return L


class DDLAError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().ddla_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> int:
    """Raises on a negative status; returns info (>= 0) otherwise."""
    if status < 0:
        raise DDLAError(status, last_error())
    return status
