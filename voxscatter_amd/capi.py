"""ctypes binding of the C ABI declared in include/voxscatter.h.

Loading fails loudly when the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvoxscatter.so"

VS_OK = 0
VS_ERR_ARG = -1
VS_ERR_HIP = -2
VS_ERR_OOM = -3
VS_ERR_NODEV = -4
VS_ERR_UNSUPPORTED = -5
VS_ERR_NAN = -6
VS_ERR_INTERNAL = -7
VS_FP64 = 1
VS_FP32 = 2

_ERR_NAMES = {VS_ERR_ARG: "VS_ERR_ARG", VS_ERR_HIP: "VS_ERR_HIP", VS_ERR_OOM: "VS_ERR_OOM",
              VS_ERR_NODEV: "VS_ERR_NODEV", VS_ERR_UNSUPPORTED: "VS_ERR_UNSUPPORTED",
              VS_ERR_NAN: "VS_ERR_NAN", VS_ERR_INTERNAL: "VS_ERR_INTERNAL"}


class VoxScatterError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


P = C.c_void_p
D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_int64)
INT = C.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "vs_last_error": (C.c_char_p, []),
    "vs_version": (INT, []),
    "vs_device_count": (INT, [C.POINTER(C.c_int)]),
    "vs_ctx_create": (INT, [INT, C.POINTER(P)]),
    "vs_ctx_destroy": (None, [P]),
    "vs_ctx_synchronize": (INT, [P]),
    "vs_dev_alloc": (INT, [P, C.c_int64, C.POINTER(P)]),
    "vs_dev_free": (INT, [P, P]),
    "vs_dev_copy": (INT, [P, P, P, C.c_int64, INT]),
    "vs_alpha_ldr": (INT, [D, C.c_double, C.c_double, D, D, D]),
    "vs_op_create_dda": (INT, [P, INT, INT, INT, D, C.c_double, INT, INT, C.POINTER(P)]),
    "vs_op_create_fN": (INT, [P, INT, INT, INT, D, D, INT, C.POINTER(P)]),
    "vs_op_destroy": (None, [P]),
    "vs_op_create_dense": (INT, [P, C.c_int64, D, INT, C.POINTER(P)]),
    "vs_pc_create_op": (INT, [P, C.POINTER(P)]),
    "vs_op_get_toeplitz": (INT, [P, D]),
    "vs_op_get_fN": (INT, [P, D]),
    "vs_op_set_domain": (INT, [P, I64, C.c_int64, D, C.c_int64, D, C.c_int64, D, C.c_int64]),
    "vs_op_nunknowns": (C.c_int64, [P]),
    "vs_op_create_galerkin": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, C.c_int,
                                        C.POINTER(P)]),
    "vs_galerkin_singular": (C.c_int, [P, C.c_double, C.c_double, D]),
    "vs_galerkin_near_table": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vs_galerkin_singular_directfn": (C.c_int, [C.c_double, C.c_double, C.c_int, D]),
    "vs_op_matvec_variant": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
    "vs_set_fused_krylov": (C.c_int, [C.c_int]),
    "vs_fused_krylov_available": (C.c_int, [P, P, C.c_int, C.POINTER(C.c_int)]),
    "vs_mvp": (INT, [P, D, D, INT]),
    "vs_mvp_dev": (INT, [P, P, P, INT, INT, INT]),
    "vs_mvp_dev_noalloc": (INT, [P, P, P, INT, INT, INT]),
    "vs_mvp_dev_diag": (INT, [P, P, P, P, INT, INT, INT]),
    "vs_circ_1_level": (INT, [P, INT, INT, INT, D, D, D]),
    "vs_circ_2_level": (INT, [P, INT, INT, INT, D, D, D]),
    "vs_invert_blocks": (INT, [P, C.c_int64, INT, D, D, D, D, D]),
    "vs_pc_create_circ": (INT, [P, INT, D, D, D, INT, C.POINTER(P)]),
    "vs_pc_create_blocks": (INT, [P, INT, INT, INT, INT, D, I64, C.c_int64, INT, C.POINTER(P)]),
    "vs_pc_destroy": (None, [P]),
    "vs_pc_apply": (INT, [P, D, D]),
    "vs_pc_apply_dev": (INT, [P, P, P, INT, INT]),
    "vs_bicgstab": (INT, [P, P, D, D, C.c_double, INT, D, C.POINTER(C.c_int), D, D, D, C.POINTER(C.c_int)]),
    "vs_bicgstab_dev": (INT, [P, P, P, P, C.c_double, INT, C.POINTER(C.c_int), D, D, C.POINTER(C.c_int)]),
    "vs_gmres": (INT, [P, P, D, D, INT, C.c_double, INT, D, C.POINTER(C.c_int), D, C.POINTER(C.c_int), D,
                       C.POINTER(C.c_int)]),
    "vs_gmres_dev": (INT, [P, P, P, P, INT, C.c_double, INT, C.POINTER(C.c_int), D, C.POINTER(C.c_int),
                           C.POINTER(C.c_int)]),
    "vs_refine": (INT, [P, P, D, D, C.c_double, INT, INT, C.c_double, INT, INT, D, C.POINTER(C.c_int), D,
                        C.POINTER(C.c_int), C.POINTER(C.c_int), D, C.POINTER(C.c_int)]),
    "vs_refine_dev": (INT, [P, P, P, P, C.c_double, INT, INT, C.c_double, INT, INT, C.POINTER(C.c_int), D,
                            C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vs_dda_farfield": (INT, [P, C.c_int64, D, D, C.c_double, C.c_double, C.c_int64, D, D]),
    "vs_dda_cross_sections": (INT, [P, C.c_int64, D, D, D, C.c_int64, C.c_double, C.c_double, C.c_double, D]),
    "vs_dda_grad_eps": (INT, [P, C.c_int64, D, D, D, D]),
    "vs_plane_wave_batch_dev": (INT, [P, C.c_int64, D, INT, D, INT, D, C.c_double, P]),
    "vs_bicgstab_batch_dev": (INT, [P, P, INT, P, P, INT, C.c_double, INT, C.POINTER(C.c_int), D, D]),
    "vs_refine_batch_dev": (INT, [P, P, INT, P, P, C.c_double, INT, C.c_double, INT, C.POINTER(C.c_int), D,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vs_bicgstab_batch_diag_dev": (INT, [P, P, INT, P, P, D, INT, C.c_double, INT, C.POINTER(C.c_int), D, D]),
    "vs_refine_batch_diag_dev": (INT, [P, P, INT, P, P, D, C.c_double, INT, C.c_double, INT, C.POINTER(C.c_int), D,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vs_bench_mvp": (INT, [P, P, P, INT, INT, INT, INT, D]),
    "vs_bench_pc": (INT, [P, P, P, INT, INT, INT, D]),
    "vs_bench_pc_blocks": (INT, [P, INT, INT, INT, D]),
    "vs_pgmres": (INT, [P, P, INT, D, D, INT, C.c_double, INT, D, C.POINTER(C.c_int), D, C.POINTER(C.c_int), D,
                        C.POINTER(C.c_int)]),
    "vs_pgmresdr": (INT, [P, P, INT, D, D, INT, C.c_double, INT, INT, D, C.POINTER(C.c_int), D,
                          C.POINTER(C.c_int), D, C.POINTER(C.c_int)]),
    "vs_recycle_create": (INT, [P, C.POINTER(P)]),
    "vs_recycle_destroy": (None, [P]),
    "vs_recycle_size": (INT, [P]),
    "vs_recycle_clear": (INT, [P]),
    "vs_pgmresdr_rec": (INT, [P, P, INT, D, D, INT, C.c_double, INT, INT, D, P, C.POINTER(C.c_int), D,
                              C.POINTER(C.c_int), D, C.POINTER(C.c_int)]),
}

_lib = None


def load(path: os.PathLike | None = None):
    """Load the shared library (once) and declare every entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise VoxScatterError(VS_ERR_NODEV,
                              f"HIP library {p} not built (run `python -m voxscatter_amd.build`); "
                              "there is no CPU fallback")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int):
    if rc != VS_OK:
        msg = _lib.vs_last_error().decode() if _lib is not None else ""
        raise VoxScatterError(rc, msg)


def dptr(a: np.ndarray):
    """Pointer to the data of a contiguous float64/complex128 array."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    assert a.dtype in (np.float64, np.complex128), a.dtype
    return a.ctypes.data_as(D)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(I64)


def cplx_arr(x, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).ravel())
    if n is not None and a.size != n:
        raise VoxScatterError(VS_ERR_ARG, f"expected {n} complex values, got {a.size}")
    return a


def c2(z) -> np.ndarray:
    z = complex(z)
    return np.array([z.real, z.imag], dtype=np.float64)
