"""Host-side mirror of the reference's MATLAB interface for the solve loop, backed by the HIP library.

The MATLAB functions keep their names, argument order and meaning:

* ``getOPERATOR_DDA``      src/src_vie/src_operators/getOPERATOR_DDA.m:1-2
* ``MVP`` / ``mv_AN_const`` src/src_vie/src_mvp/MVP.m:1 (``GPU_flag`` is accepted and ignored: the
                           product always runs on the GPU)
* ``circ_1_level``         src/src_preconditioning/circ_1_level.m:1
* ``circ_2_level``         src/src_preconditioning/circ_2_level.m:1
* ``mvp_circ_2_level``     src/src_preconditioning/mvp_circ_2_level.m:2
* ``chan_mvp_idx``         src/src_preconditioning/chan_mvp_idx.m:2
* ``bicgstab`` / ``gmres`` MATLAB built-ins as called at main_DDA.m:82,89,120,128
* ``pgmres``               src/src_numeric/src_iterative/pgmres.m:70

Differences forced by the language boundary: indices are 0-based; an operator "function handle" is an
:class:`Operator` whose domain has been set (``A.set_domain``) and a preconditioner handle is a
:class:`Preconditioner`, because the Krylov loops run on the device and cannot call back into Python.
"""
from __future__ import annotations

import atexit
import ctypes as C
import hashlib
import weakref
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import VS_FP32, VS_FP64, VoxScatterError, c2, check, cplx_arr, dptr, i64ptr

# --------------------------------------------------------------------------------------------------
# handle lifetime
# --------------------------------------------------------------------------------------------------
# The library keeps a context alive while any handle made on it lives (reference counted), so handles
# may be released in any order.  Every live handle is also registered here, and an exit hook releases
# them children first (buffers, checkpoints, preconditioners, operators, then contexts) while the
# interpreter and the HIP runtime are still whole, instead of leaving the order to finalisers.
_LIVE = {kind: weakref.WeakSet() for kind in ("array", "recycle", "pc", "op", "ctx")}


def _track(kind: str, obj):
    _LIVE[kind].add(obj)


def close_all():
    """Release every live handle of this process, children first."""
    for kind in ("array", "recycle", "pc", "op", "ctx"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(close_all)

# --------------------------------------------------------------------------------------------------
# context
# --------------------------------------------------------------------------------------------------


class Context:
    """One HIP stream on one device."""

    def __init__(self, device: int = 0):
        self.lib = capi.load()
        h = C.c_void_p()
        check(self.lib.vs_ctx_create(int(device), C.byref(h)))
        self.h = h
        self.device = device
        _track("ctx", self)

    def synchronize(self):
        check(self.lib.vs_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.vs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def get_context() -> Context:
    global _default_ctx
    if _default_ctx is None or not getattr(_default_ctx, "h", None):
        _default_ctx = Context(0)
    return _default_ctx


def device_count() -> int:
    lib = capi.load()
    n = C.c_int(0)
    check(lib.vs_device_count(C.byref(n)))
    return n.value


def _fortran_flat(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1, order="F"))


def _coef(v, V):
    a = np.asarray(v, dtype=np.complex128)
    if a.size == 1:
        return np.ascontiguousarray(a.reshape(1)), 1
    if a.size != V:
        raise VoxScatterError(capi.VS_ERR_ARG, f"coefficient must be scalar or L*M*N={V} values")
    return _fortran_flat(a, np.complex128), V


def _digest(*arrays) -> str:
    h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _full_digest(a: np.ndarray) -> str:
    """Digest of every element of a (in memory order, no copy when a is C- or F-contiguous): the key under
    which a device copy made from a host array is reused, so that an in-place edit anywhere in the array,
    or another array at a reused address, makes a new copy."""
    a = np.asarray(a)
    flat = a.reshape(-1, order="A") if (a.flags.c_contiguous or a.flags.f_contiguous) else np.ascontiguousarray(a).ravel()
    buf = memoryview(flat.view(np.uint8))
    try:
        import xxhash
        h = xxhash.xxh3_128()
        h.update(buf)
        body = h.hexdigest()
    except ImportError:
        body = hashlib.blake2b(buf, digest_size=16).hexdigest()
    return f"{a.dtype.str}{a.shape}{'F' if a.flags.f_contiguous and not a.flags.c_contiguous else 'C'}{body}"


class DeviceArray:
    """A complex vector (or batch of vectors) in HBM, allocated through the library's own runtime."""

    def __init__(self, shape, dtype=np.complex128, ctx: Context | None = None):
        self.ctx = ctx or get_context()
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        assert self.dtype in (np.complex128, np.complex64)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(self.ctx.lib.vs_dev_alloc(self.ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value
        self._ctxh = self.ctx.h  # the buffer keeps its context alive in the library
        _track("array", self)

    @property
    def precision(self) -> int:
        return VS_FP64 if self.dtype == np.complex128 else VS_FP32

    @classmethod
    def from_numpy(cls, a, dtype=None, ctx: Context | None = None):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype or np.asarray(a).dtype))
        d = cls(a.shape, a.dtype, ctx)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype))
        assert a.nbytes == self.nbytes
        check(self.ctx.lib.vs_dev_copy(self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 0))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.lib.vs_dev_copy(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                       self.nbytes, 1))
        return out

    def zero(self):
        self.upload(np.zeros(self.shape, dtype=self.dtype))

    def close(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.vs_dev_free(self._ctxh, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------------------------------
# operator
# --------------------------------------------------------------------------------------------------


class Operator:
    """Device-resident circulant operator fN (and Toeplitz tensor), plus the unknown set of MVP."""

    def __init__(self, handle, ctx: Context, shape, precisions: int):
        self.h = handle
        self.ctx = ctx
        self.L, self.M, self.N = shape
        self.precisions = precisions
        self._domain_key = None
        self.n = 0
        _track("op", self)

    # construction ---------------------------------------------------------------------------------
    @classmethod
    def dda(cls, r, ko: float, nearby_quad: str = "off", precisions: int = VS_FP64, ctx: Context | None = None):
        ctx = ctx or get_context()
        r = np.asarray(r, dtype=np.float64)
        L, M, N, three = r.shape
        assert three == 3
        h = C.c_void_p()
        rf = _fortran_flat(r, np.float64)
        check(ctx.lib.vs_op_create_dda(ctx.h, L, M, N, dptr(rf), float(ko), 1 if nearby_quad == "on" else 0,
                                       int(precisions), C.byref(h)))
        return cls(h, ctx, (L, M, N), precisions)

    @classmethod
    def galerkin(cls, shape, dx: float, ko: float, singular=None, precisions: int = VS_FP64,
                 ctx: Context | None = None):
        """Galerkin N operator (getOPERATORS(r, f, 'N', 'FFTofC')) on an L x M x N grid of cubes of side dx.
        singular: the five singular face-pair integrals [ST, EAc, EAo, VAc, VAo]; None takes the reference's
        own values (the direct evaluation method at Np = 20, ``galerkin_singular_directfn``)."""
        ctx = ctx or get_context()
        L, M, N = (int(v) for v in shape)
        sg = None if singular is None else cplx_arr(np.asarray(singular, dtype=np.complex128))
        h = C.c_void_p()
        check(ctx.lib.vs_op_create_galerkin(ctx.h, L, M, N, float(dx), float(ko), dptr(sg), int(precisions),
                                            C.byref(h)))
        return cls(h, ctx, (L, M, N), precisions)

    @classmethod
    def dense(cls, A, precisions: int = VS_FP64, ctx: Context | None = None):
        """Matrix operand (iterapp.m:41-58): y = A x for an n x n complex matrix; ``transp`` gives A^H x."""
        ctx = ctx or get_context()
        A = np.asarray(A, dtype=np.complex128)
        n = A.shape[0]
        assert A.shape == (n, n)
        a = _fortran_flat(A, np.complex128)
        h = C.c_void_p()
        check(ctx.lib.vs_op_create_dense(ctx.h, n, dptr(a), int(precisions), C.byref(h)))
        op = cls(h, ctx, (0, 0, 0), precisions)
        op.n = n
        return op

    @classmethod
    def from_fN(cls, fN, Toep=None, precisions: int = VS_FP64, ctx: Context | None = None):
        ctx = ctx or get_context()
        fN = np.asarray(fN)
        L2, M2, N2, six = fN.shape
        assert six == 6 and L2 % 2 == 0 and M2 % 2 == 0 and N2 % 2 == 0
        L, M, N = L2 // 2, M2 // 2, N2 // 2
        f = _fortran_flat(fN, np.complex128)
        t = None if Toep is None else _fortran_flat(Toep, np.complex128)
        h = C.c_void_p()
        check(ctx.lib.vs_op_create_fN(ctx.h, L, M, N, dptr(f), dptr(t), int(precisions), C.byref(h)))
        return cls(h, ctx, (L, M, N), precisions)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vs_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.L, self.M, self.N)

    @property
    def V(self):
        return self.L * self.M * self.N

    # data -----------------------------------------------------------------------------------------
    def toeplitz(self) -> np.ndarray:
        out = np.zeros(self.V * 6, dtype=np.complex128)
        check(self.ctx.lib.vs_op_get_toeplitz(self.h, dptr(out)))
        return out.reshape((self.L, self.M, self.N, 6), order="F")

    def fN(self) -> np.ndarray:
        out = np.zeros(self.V * 48, dtype=np.complex128)
        check(self.ctx.lib.vs_op_get_fN(self.h, dptr(out)))
        return out.reshape((2 * self.L, 2 * self.M, 2 * self.N, 6), order="F")

    def set_domain(self, idx, Gram, Mr, Mc):
        """Unknown set (0-based indices into 3 L M N) and the diagonal Gram .* Mr, Mc of MVP.m."""
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64).ravel())
        g, gl = _coef(Gram, self.V)
        mr, ml = _coef(Mr, self.V)
        mc, cl = _coef(Mc, self.V)
        key = _digest(idx, g, mr, mc)
        if key == self._domain_key:
            return self
        check(self.ctx.lib.vs_op_set_domain(self.h, i64ptr(idx), idx.size, dptr(g), gl, dptr(mr), ml, dptr(mc), cl))
        self._domain_key = key
        self._domain_args = (idx, Gram, Mr, Mc)
        self.n = int(idx.size)
        self.idx = idx
        return self

    def mvp(self, x, transp: bool = False) -> np.ndarray:
        x = cplx_arr(x, self.n)
        y = np.empty_like(x)
        check(self.ctx.lib.vs_mvp(self.h, dptr(x), dptr(y), 1 if transp else 0))
        return y

    def __call__(self, x):
        return self.mvp(x)

    def matvec_variant(self, precision: int = VS_FP64) -> str:
        """Matvec variant in use for a precision: 'pass5', 'slab' or 'split' (autotuned at the first domain)."""
        v = C.c_int()
        check(self.ctx.lib.vs_op_matvec_variant(self.h, int(precision), C.byref(v)))
        return ("pass5", "slab", "split", "split2")[v.value]

    def mvp_dev(self, x_ptr: int, y_ptr: int, transp: bool = False, precision: int = VS_FP64, nrhs: int = 1):
        """Device matvec on raw device pointers (e.g. torch ``data_ptr()``), asynchronous on the context stream."""
        check(self.ctx.lib.vs_mvp_dev(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), 1 if transp else 0,
                                      int(precision), int(nrhs)))

    def mvp_dev_diag(self, x_ptr: int, y_ptr: int, diag_ptr: int, transp: bool = False, precision: int = VS_FP64,
                     nrhs: int = 1):
        """Device matvec with one diagonal (Gram Mr) per right-hand side, diag_ptr: nrhs complex128 on the device."""
        check(self.ctx.lib.vs_mvp_dev_diag(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(diag_ptr),
                                           1 if transp else 0, int(precision), int(nrhs)))


# --------------------------------------------------------------------------------------------------
# preconditioner
# --------------------------------------------------------------------------------------------------


class Preconditioner:
    """Explicit inverse blocks of a Chan 1- or 2-level circulant approximation, resident on the device."""

    def __init__(self, handle, ctx: Context, levels: int, n: int):
        self.h = handle
        self.ctx = ctx
        self.levels = levels
        self.n = n
        _track("pc", self)

    @classmethod
    def from_operator(cls, op: Operator, levels: int, a, w=None, wscalar=None, precisions: int = VS_FP64):
        """Blocks inv(a I - diag(w) C) built on the device from the operator's Toeplitz tensor."""
        if w is None and wscalar is None:
            raise VoxScatterError(capi.VS_ERR_ARG, "w or wscalar required")
        h = C.c_void_p()
        wv = None if w is None else cplx_arr(w)
        ws = None if wscalar is None else c2(wscalar)
        check(op.ctx.lib.vs_pc_create_circ(op.h, int(levels), dptr(c2(a)), dptr(wv), dptr(ws), int(precisions),
                                           C.byref(h)))
        return cls(h, op.ctx, levels, op.n)

    @classmethod
    def of_operator(cls, mop: "Operator"):
        """Preconditioner function handle whose apply is ``mop``'s matvec (M^-1 x = mop x), e.g. the material
        preconditioner mv_AN_const(J, fN, 1./Mr, 1./Mr - 1, Gram, ...) of JVIE_Solver_const.m:232-234 as an
        operator on the same fN with that diagonal.  The preconditioner keeps ``mop`` alive."""
        h = C.c_void_p()
        check(mop.ctx.lib.vs_pc_create_op(mop.h, C.byref(h)))
        pc = cls(h, mop.ctx, 0, mop.n)
        pc._mop = mop
        return pc

    @classmethod
    def from_blocks(cls, levels: int, inv_blocks, L, M, N, idx, precisions: int = VS_FP64, ctx: Context | None = None):
        ctx = ctx or get_context()
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64).ravel())
        raw = blocks_to_abi(np.asarray(inv_blocks, dtype=np.complex128), levels)
        h = C.c_void_p()
        check(ctx.lib.vs_pc_create_blocks(ctx.h, int(levels), int(L), int(M), int(N), dptr(raw), i64ptr(idx),
                                          idx.size, int(precisions), C.byref(h)))
        return cls(h, ctx, levels, int(idx.size))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vs_pc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x) -> np.ndarray:
        x = cplx_arr(x, self.n)
        y = np.empty_like(x)
        check(self.ctx.lib.vs_pc_apply(self.h, dptr(x), dptr(y)))
        return y

    def __call__(self, x):
        return self.apply(x)

    def apply_dev(self, x_ptr: int, y_ptr: int, precision: int = VS_FP64, nrhs: int = 1):
        check(self.ctx.lib.vs_pc_apply_dev(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(precision), int(nrhs)))


def blocks_to_abi(blocks: np.ndarray, levels: int) -> np.ndarray:
    """(L, M, n, n) or (L, n, n) numpy blocks -> the ABI's flat column-major block sequence."""
    if levels == 2:
        L, M, n, _ = blocks.shape
        # block (i, j) at i + L j, entry (r, s) at r + n s
        return np.ascontiguousarray(blocks.transpose(1, 0, 3, 2)).ravel()
    L, n, _ = blocks.shape
    return np.ascontiguousarray(blocks.transpose(0, 2, 1)).ravel()


def blocks_from_abi(raw: np.ndarray, levels: int, L: int, M: int, n: int) -> np.ndarray:
    if levels == 2:
        return raw.reshape(M, L, n, n).transpose(1, 0, 3, 2).copy()
    return raw.reshape(L, n, n).transpose(0, 2, 1).copy()


# --------------------------------------------------------------------------------------------------
# MATLAB-named entry points
# --------------------------------------------------------------------------------------------------


def alpha_LDR(refInd, ko, d, kvec, Eo) -> complex:
    lib = capi.load()
    out = np.zeros(2)
    check(lib.vs_alpha_ldr(dptr(c2(refInd)), float(ko), float(d), dptr(np.asarray(kvec, dtype=np.float64)),
                           dptr(np.asarray(Eo, dtype=np.float64)), dptr(out)))
    return complex(out[0], out[1])


def getOPERATOR_DDA(r, ko, refInd, kvec, Eo, nearby_quad="off", precisions: int = VS_FP64, ctx=None):
    """``[op_out, Toep, alpha_LDR] = getOPERATOR_DDA(r, ko, refInd, kvec, Eo, nearby_quad)``.

    ``op_out`` is the device-resident :class:`Operator` (its ``fN()`` downloads the circulant FFT).
    """
    r = np.asarray(r, dtype=np.float64)
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    op = Operator.dda(r, ko, nearby_quad, precisions, ctx)
    return op, op.toeplitz(), alpha_LDR(refInd, ko, dx, kvec, Eo)


_fn_cache: dict = {}


def _as_operator(fN) -> Operator:
    if isinstance(fN, Operator):
        return fN
    fN = np.asarray(fN)
    key = _full_digest(fN)
    op = _fn_cache.get(key)
    if op is None:
        op = Operator.from_fN(fN)
        _fn_cache.clear()
        _fn_cache[key] = op
    return op


def MVP(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag=1):
    """``JOut = MVP(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag)`` (0-based ``idx``)."""
    if transp_flag not in ("transp", "notransp"):
        raise VoxScatterError(capi.VS_ERR_ARG, "transp_flag must be 'transp' or 'notransp'")
    op = _as_operator(fN)
    Mr = np.asarray(Mr)
    if Mr.ndim == 3 and Mr.shape != op.shape:
        raise VoxScatterError(capi.VS_ERR_ARG, "Mr shape does not match the operator grid")
    op.set_domain(idx, Gram, Mr, Mc)
    return op.mvp(JIn0, transp=(transp_flag == "transp"))


mv_AN_const = MVP


def circ_1_level(opToeplitz, L, M, N, parallel_prec=0, on_off="off", ctx=None):
    """``[circ, circ_L_opToep] = circ_1_level(opToeplitz, L, M, N, parallel_prec, on_off)``."""
    ctx = ctx or get_context()
    t = _fortran_flat(opToeplitz, np.complex128)
    cL = np.zeros(L * M * N * 6, dtype=np.complex128)
    blocks = None
    if on_off == "on":
        n = 3 * M * N
        blocks = np.zeros(L * n * n, dtype=np.complex128)
    check(ctx.lib.vs_circ_1_level(ctx.h, L, M, N, dptr(t), dptr(cL), dptr(blocks)))
    circ = None if blocks is None else blocks_from_abi(blocks, 1, L, M, 3 * M * N)
    return circ, cL.reshape((L, M, N, 6), order="F")


def circ_2_level(circ_L_opToep, L, M, N, ctx=None):
    """``[circ_M_opToep, circ] = circ_2_level(circ_L_opToep, L, M, N)``; circ is (L, M, 3N, 3N)."""
    ctx = ctx or get_context()
    cL = _fortran_flat(circ_L_opToep, np.complex128)
    cM = np.zeros(L * M * N * 6, dtype=np.complex128)
    n = 3 * N
    blocks = np.zeros(L * M * n * n, dtype=np.complex128)
    check(ctx.lib.vs_circ_2_level(ctx.h, L, M, N, dptr(cL), dptr(cM), dptr(blocks)))
    return cM.reshape((L, M, N, 6), order="F"), blocks_from_abi(blocks, 2, L, M, n)


def invert_blocks(blocks, a, w=None, ctx=None):
    """``inv(a*eye(n) - diag(w)*B)`` for every block of (..., n, n) ``blocks`` (``w`` scalar or n-vector)."""
    ctx = ctx or get_context()
    B = np.asarray(blocks, dtype=np.complex128)
    n = B.shape[-1]
    flat = B.reshape(-1, n, n)
    raw = np.ascontiguousarray(flat.transpose(0, 2, 1)).ravel()
    out = np.zeros_like(raw)
    wv = np.asarray(w if w is not None else 1.0, dtype=np.complex128)
    wvec = cplx_arr(wv) if wv.size > 1 else None
    ws = c2(complex(wv.ravel()[0])) if wv.size == 1 else None
    check(ctx.lib.vs_invert_blocks(ctx.h, flat.shape[0], n, dptr(raw), dptr(c2(a)), dptr(wvec), dptr(ws), dptr(out)))
    return out.reshape(-1, n, n).transpose(0, 2, 1).reshape(B.shape).copy()


_pc_cache: dict = {}


def _pc_from_blocks(levels, inv_blocks, L, M, N, idx):
    idx = np.asarray(idx, dtype=np.int64)
    ib = np.asarray(inv_blocks)
    key = (levels, L, M, N, _full_digest(ib), _digest(idx))
    pc = _pc_cache.get(key)
    if pc is None:
        pc = Preconditioner.from_blocks(levels, ib, L, M, N, idx)
        _pc_cache.clear()
        _pc_cache[key] = pc
    return pc


def mvp_circ_2_level(inv_blocks, JIn, L, M, N, idx):
    """``mvp = mvp_circ_2_level(inv_blocks, JIn, L, M, N, idx)``; inv_blocks (L, M, 3N, 3N)."""
    if isinstance(inv_blocks, Preconditioner):
        return inv_blocks.apply(JIn)
    return _pc_from_blocks(2, inv_blocks, L, M, N, idx).apply(JIn)


def chan_mvp_idx(inv_blocks, JIn, L, M, N, idx):
    """``mvp = chan_mvp_idx(inv_blocks, JIn, L, M, N, idx)``; inv_blocks (L, 3MN, 3MN)."""
    if isinstance(inv_blocks, Preconditioner):
        return inv_blocks.apply(JIn)
    return _pc_from_blocks(1, inv_blocks, L, M, N, idx).apply(JIn)


@dataclass
class SolveInfo:
    flag: int
    relres: float
    iter: object
    resvec: np.ndarray


def _check_A(A) -> Operator:
    if not isinstance(A, Operator) or A.n == 0:
        raise VoxScatterError(capi.VS_ERR_ARG, "A must be an Operator with its domain set (A.set_domain)")
    return A


def bicgstab(A, b, tol=1e-6, maxit=None, M1=None, M2=None, x0=None):
    """``[x, flag, relres, iter, resvec] = bicgstab(A, b, tol, maxit, M1, [], x0)`` on the device."""
    A = _check_A(A)
    if M2 is not None:
        raise VoxScatterError(capi.VS_ERR_UNSUPPORTED, "a second preconditioner factor is not supported")
    b = cplx_arr(b, A.n)
    maxit = min(A.n, 20) if maxit is None else int(maxit)
    x = np.zeros_like(b)
    x0a = None if x0 is None else cplx_arr(x0, A.n)
    flag = C.c_int()
    relres = C.c_double()
    it = C.c_double()
    nres = C.c_int()
    resvec = np.zeros(2 * maxit + 1)
    check(A.ctx.lib.vs_bicgstab(A.h, None if M1 is None else M1.h, dptr(b), dptr(x), float(tol), maxit, dptr(x0a),
                                C.byref(flag), C.byref(relres), C.byref(it), dptr(resvec), C.byref(nres)))
    return x, flag.value, relres.value, it.value, resvec[:nres.value].copy()


def gmres(A, b, restart=None, tol=1e-6, maxit=None, M1=None, M2=None, x0=None):
    """``[x, flag, relres, iter, resvec] = gmres(A, b, restart, tol, maxit, M1, [], x0)`` on the device."""
    A = _check_A(A)
    if M2 is not None:
        raise VoxScatterError(capi.VS_ERR_UNSUPPORTED, "a second preconditioner factor is not supported")
    n = A.n
    restart = n if restart is None or restart > n else int(restart)
    restart = min(restart, 511)
    maxit = min(int(np.ceil(n / restart)), 10) if maxit is None else int(maxit)
    b = cplx_arr(b, n)
    x = np.zeros_like(b)
    x0a = None if x0 is None else cplx_arr(x0, n)
    flag = C.c_int()
    relres = C.c_double()
    it = (C.c_int * 2)()
    nres = C.c_int()
    resvec = np.zeros(restart * maxit + 1)
    check(A.ctx.lib.vs_gmres(A.h, None if M1 is None else M1.h, dptr(b), dptr(x), restart, float(tol), maxit,
                             dptr(x0a), C.byref(flag), C.byref(relres), it, dptr(resvec), C.byref(nres)))
    return x, flag.value, relres.value, (it[0], it[1]), resvec[:nres.value].copy()


def pgmres(A, b, restart=50, tol=1e-3, maxit=10, L1=None, L2=None, R1=None, R2=None, x0=None):
    """``[x, flag, relres, iter, resvec] = pgmres(A, b, restart, tol, maxit, L1, L2, R1, R2, x0)``.

    One preconditioner factor is supported, either as L1 (left) or R1 (right).
    """
    A = _check_A(A)
    if L2 is not None or R2 is not None or (L1 is not None and R1 is not None):
        raise VoxScatterError(capi.VS_ERR_UNSUPPORTED, "only one of L1 / R1 is supported")
    pc, side = (L1, 0) if L1 is not None else (R1, 1)
    n = A.n
    b = cplx_arr(b, n)
    x = np.zeros_like(b)
    x0a = None if x0 is None else cplx_arr(x0, n)
    flag = C.c_int()
    relres = C.c_double()
    it = (C.c_int * 2)()
    nres = C.c_int()
    resvec = np.zeros(int(restart) * int(maxit) + 1)
    check(A.ctx.lib.vs_pgmres(A.h, None if pc is None else pc.h, side, dptr(b), dptr(x), int(restart), float(tol),
                              int(maxit), dptr(x0a), C.byref(flag), C.byref(relres), it, dptr(resvec),
                              C.byref(nres)))
    return x, flag.value, relres.value, (it[0], it[1]), resvec[:nres.value].copy()


class Recycle:
    """Krylov-basis checkpoint of GMRES-DR: the deflation space a ``pgmresDR`` solve ended with, kept in
    device memory for the next solve with the same operator, preconditioner and side (for example the
    adjoint solve of a DDA gradient after the forward solve, A^T = A).  No counterpart in the reference."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or get_context()
        h = C.c_void_p()
        check(self.ctx.lib.vs_recycle_create(self.ctx.h, C.byref(h)))
        self.h = h
        _track("recycle", self)

    @property
    def size(self) -> int:
        """Vectors held (0 when empty)."""
        return int(self.ctx.lib.vs_recycle_size(self.h))

    def clear(self):
        check(self.ctx.lib.vs_recycle_clear(self.h))

    def close(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.ctx.lib.vs_recycle_destroy(h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pgmresDR(A, b, restart=50, tol=1e-3, maxit=10, ritz=None, L1=None, L2=None, R1=None, R2=None, x0=None,
             recycle: Recycle | None = None):
    """``[x, flag, relres, iter, resvec] = pgmresDR(A, b, restart, tol, maxit, ritz, L1, L2, R1, R2, x0)``:
    GMRES with deflated restarting (ritz harmonic Ritz vectors kept between cycles).  One preconditioner
    factor is supported, either as L1 (left) or R1 (right).  ``recycle``: a ``Recycle`` checkpoint; when it
    holds a space from an earlier solve, the first cycle is replaced by the projection onto it, and on
    return it holds this solve's final deflation space."""
    A = _check_A(A)
    if L2 is not None or R2 is not None or (L1 is not None and R1 is not None):
        raise VoxScatterError(capi.VS_ERR_UNSUPPORTED, "only one of L1 / R1 is supported")
    pc, side = (L1, 0) if L1 is not None else (R1, 1)
    n = A.n
    b = cplx_arr(b, n)
    x = np.zeros_like(b)
    x0a = None if x0 is None else cplx_arr(x0, n)
    flag = C.c_int()
    relres = C.c_double()
    it = (C.c_int * 2)()
    nres = C.c_int()
    resvec = np.zeros(int(restart) * (int(maxit) + 1) + 2)
    if recycle is None:
        check(A.ctx.lib.vs_pgmresdr(A.h, None if pc is None else pc.h, side, dptr(b), dptr(x), int(restart),
                                    float(tol), int(maxit), 0 if ritz is None else int(ritz), dptr(x0a),
                                    C.byref(flag), C.byref(relres), it, dptr(resvec), C.byref(nres)))
    else:
        check(A.ctx.lib.vs_pgmresdr_rec(A.h, None if pc is None else pc.h, side, dptr(b), dptr(x), int(restart),
                                        float(tol), int(maxit), 0 if ritz is None else int(ritz), dptr(x0a),
                                        recycle.h, C.byref(flag), C.byref(relres), it, dptr(resvec),
                                        C.byref(nres)))
    return x, flag.value, relres.value, (it[0], it[1]), resvec[:nres.value].copy()


# --------------------------------------------------------------------------------------------------
# precision form and device-pointer entry points
# --------------------------------------------------------------------------------------------------

def refine(A, b, M1=None, tol=1e-6, max_outer=10, inner="bicgstab", inner_tol=1e-4, inner_maxit=2000, restart=50,
           x0=None):
    """FP32 inner solve + FP64 iterative refinement; returns ``(x, flag, relres, outer, inner_total, resvec)``.

    ``A`` needs FP64 and FP32 tensors (``precisions=VS_FP64|VS_FP32``), ``M1`` FP32 blocks.
    ``resvec`` is the FP64 relative residual before each correction.
    """
    A = _check_A(A)
    b = cplx_arr(b, A.n)
    x = np.zeros_like(b)
    x0a = None if x0 is None else cplx_arr(x0, A.n)
    flag = C.c_int()
    relres = C.c_double()
    outer = C.c_int()
    tot = C.c_int()
    nres = C.c_int()
    resvec = np.zeros(max_outer + 1)
    check(A.ctx.lib.vs_refine(A.h, None if M1 is None else M1.h, dptr(b), dptr(x), float(tol), int(max_outer),
                              0 if inner == "bicgstab" else 1, float(inner_tol), int(inner_maxit), int(restart),
                              dptr(x0a), C.byref(flag), C.byref(relres), C.byref(outer), C.byref(tot), dptr(resvec),
                              C.byref(nres)))
    return x, flag.value, relres.value, outer.value, tot.value, resvec[:nres.value].copy()


def set_fused_krylov(on: bool, ctx=None) -> None:
    """Turn the fused BiCGStab halves (Krylov updates and inner products inside the FFT passes) on or off for
    the whole process; they are on by default."""
    lib = (ctx or get_context()).lib
    check(lib.vs_set_fused_krylov(1 if on else 0))


def fused_krylov_available(A: Operator, M1: Preconditioner, precision: int) -> bool:
    """Whether single-right-hand-side BiCGStab solves with A and M1 in `precision` take the fused halves."""
    on = C.c_int()
    check(A.ctx.lib.vs_fused_krylov_available(A.h, M1.h, int(precision), C.byref(on)))
    return bool(on.value)


def bicgstab_dev(A: Operator, b_ptr: int, x_ptr: int, tol=1e-6, maxit=2000, M1=None):
    flag, relres, it, nres = C.c_int(), C.c_double(), C.c_double(), C.c_int()
    check(A.ctx.lib.vs_bicgstab_dev(A.h, None if M1 is None else M1.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                    float(tol), int(maxit), C.byref(flag), C.byref(relres), C.byref(it),
                                    C.byref(nres)))
    return flag.value, relres.value, it.value, nres.value


def gmres_dev(A: Operator, b_ptr: int, x_ptr: int, restart=50, tol=1e-6, maxit=40, M1=None):
    flag, relres, nres = C.c_int(), C.c_double(), C.c_int()
    it = (C.c_int * 2)()
    check(A.ctx.lib.vs_gmres_dev(A.h, None if M1 is None else M1.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                 int(restart), float(tol), int(maxit), C.byref(flag), C.byref(relres), it,
                                 C.byref(nres)))
    return flag.value, relres.value, (it[0], it[1]), nres.value


def refine_dev(A: Operator, b_ptr: int, x_ptr: int, M1=None, tol=1e-6, max_outer=10, inner="bicgstab",
               inner_tol=1e-4, inner_maxit=2000, restart=50):
    flag, relres, outer, tot = C.c_int(), C.c_double(), C.c_int(), C.c_int()
    check(A.ctx.lib.vs_refine_dev(A.h, None if M1 is None else M1.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                  float(tol), int(max_outer), 0 if inner == "bicgstab" else 1, float(inner_tol),
                                  int(inner_maxit), int(restart), C.byref(flag), C.byref(relres), C.byref(outer),
                                  C.byref(tot)))
    return flag.value, relres.value, outer.value, tot.value


def galerkin_singular(ko: float, dx: float, ctx: Context | None = None) -> np.ndarray:
    """[ST, EAc, EAo, VAc, VAo]: int_S int_S' exp(-i ko R)/R over pairs of squares of side dx (coincident,
    edge-adjacent coplanar/orthogonal, vertex-adjacent coplanar/orthogonal), evaluated on the device."""
    ctx = ctx or get_context()
    out = np.zeros(10)
    check(ctx.lib.vs_galerkin_singular(ctx.h, float(ko), float(dx), dptr(out)))
    return out[0::2] + 1j * out[1::2]


def galerkin_singular_directfn(ko: float, dx: float, np_1d: int = 20) -> np.ndarray:
    """[ST, EAc, EAo, VAc, VAo] by the direct evaluation method (DIRECTFN) with np_1d-point rules, as the
    reference's singular_*_const.m compute them (np_1d = 20 is its default).  Host computation."""
    lib = capi.load()
    out = np.zeros(10)
    check(lib.vs_galerkin_singular_directfn(float(ko), float(dx), int(np_1d), dptr(out)))
    return out[0::2] + 1j * out[1::2]


def bench_mvp(A: Operator, x_ptr: int, y_ptr: int, transp=False, precision=VS_FP64, nrhs=1, nrep=20) -> float:
    """Mean milliseconds per matvec, HIP events on the library stream."""
    ms = C.c_double()
    check(A.ctx.lib.vs_bench_mvp(A.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), 1 if transp else 0, int(precision),
                                 int(nrhs), int(nrep), C.byref(ms)))
    return ms.value


def bench_pc(M: Preconditioner, x_ptr: int, y_ptr: int, precision=VS_FP64, nrhs=1, nrep=20) -> float:
    ms = C.c_double()
    check(M.ctx.lib.vs_bench_pc(M.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(precision), int(nrhs), int(nrep),
                                C.byref(ms)))
    return ms.value


def _diag_arg(diag, k):
    if diag is None:
        return None
    d = cplx_arr(diag)
    if d.size != k:
        raise VoxScatterError(capi.VS_ERR_ARG, f"diag needs one value per right-hand side ({k}), got {d.size}")
    return d


def bench_pc_blocks(M: Preconditioner, precision=VS_FP32, nrhs=1, nrep=5) -> float:
    """Mean milliseconds of the preconditioner's block-product stage alone (HIP events, library stream)."""
    ms = C.c_double()
    check(M.ctx.lib.vs_bench_pc_blocks(M.h, int(precision), int(nrhs), int(nrep), C.byref(ms)))
    return ms.value


def bicgstab_batch_dev(A: Operator, B: "DeviceArray", X: "DeviceArray", tol=1e-6, maxit=2000, M1=None, diag=None):
    """Batched BiCGStab over the rows of B (nrhs x n device array, FP64 or FP32); returns (flags, relres, iters).
    ``diag``: one operator diagonal (Gram Mr, e.g. 1/alpha_LDR of each incidence) per right-hand side."""
    k = B.shape[0]
    flags = (C.c_int * k)()
    relres = np.zeros(k)
    iters = np.zeros(k)
    d = _diag_arg(diag, k)
    check(A.ctx.lib.vs_bicgstab_batch_diag_dev(A.h, None if M1 is None else M1.h, k, C.c_void_p(B.ptr),
                                               C.c_void_p(X.ptr), dptr(d), B.precision, float(tol), int(maxit), flags,
                                               dptr(relres), dptr(iters)))
    return np.array(list(flags)), relres, iters


def refine_batch_dev(A: Operator, B: "DeviceArray", X: "DeviceArray", M1=None, tol=1e-6, max_outer=10, inner_tol=1e-3,
                     inner_maxit=2000, diag=None):
    """Batched FP32 + FP64-refinement solve; returns (flags, relres, outer, inner_iterations).
    ``diag``: one operator diagonal (Gram Mr) per right-hand side, as in ``bicgstab_batch_dev``."""
    k = B.shape[0]
    flags = (C.c_int * k)()
    relres = np.zeros(k)
    outer = C.c_int()
    tot = C.c_int()
    d = _diag_arg(diag, k)
    check(A.ctx.lib.vs_refine_batch_diag_dev(A.h, None if M1 is None else M1.h, k, C.c_void_p(B.ptr),
                                             C.c_void_p(X.ptr), dptr(d), float(tol), int(max_outer), float(inner_tol),
                                             int(inner_maxit), flags, dptr(relres), C.byref(outer), C.byref(tot)))
    return np.array(list(flags)), relres, outer.value, tot.value
