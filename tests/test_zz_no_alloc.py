"""vs_mvp_dev_noalloc (the matvec that allocates nothing, its FFT scratch a fixed 64 MiB device array per
GPU) against vs_mvp_dev, through the C ABI: two calls in a row, the second on a smaller operator, then a
request beyond the scratch, which must be refused with VS_ERR_ARG.  FP64 within 1e-12 relative, FP32
within 2e-5 (the no-allocation call may run another matvec variant when the tuned one's scratch does not
fit)."""
import ctypes as C

import numpy as np
import pytest

from voxscatter_amd import api, capi, geometry as G
from voxscatter_amd.dda import DDASystem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _compare(ctx, op, prec, nrhs, seed):
    lib = capi.load()
    dt = np.complex128 if prec == capi.VS_FP64 else np.complex64
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(nrhs, op.n)) + 1j * rng.normal(size=(nrhs, op.n))).astype(dt)
    xd = api.DeviceArray.from_numpy(X, dt, ctx)
    y_ref = api.DeviceArray(X.shape, dt, ctx)
    y_na = api.DeviceArray(X.shape, dt, ctx)
    for transp in (0, 1):
        capi.check(lib.vs_mvp_dev(op.h, C.c_void_p(xd.ptr), C.c_void_p(y_ref.ptr), transp, prec, nrhs))
        capi.check(lib.vs_mvp_dev_noalloc(op.h, C.c_void_p(xd.ptr), C.c_void_p(y_na.ptr), transp, prec, nrhs))
        ctx.synchronize()
        a = y_na.numpy().astype(np.complex128)
        b = y_ref.numpy().astype(np.complex128)
        err = _rel(a, b)
        print(f"[no-alloc] n={op.n} prec={prec} nrhs={nrhs} transp={transp}: rel diff {err:.2e}", flush=True)
        assert err <= (1e-12 if prec == capi.VS_FP64 else 2e-5)


def test_noalloc_matvec_matches_allocating_matvec(ctx):
    koch = DDASystem.build(G.koch_prism(), precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    print(f"[no-alloc] koch prism {koch.op.n} unknowns", flush=True)
    _compare(ctx, koch.op, capi.VS_FP32, 1, 1)
    _compare(ctx, koch.op, capi.VS_FP64, 1, 2)
    plate = DDASystem.build(G.hex_plate(), precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    print(f"[no-alloc] hex plate {plate.op.n} unknowns", flush=True)
    _compare(ctx, plate.op, capi.VS_FP64, 2, 3)
    _compare(ctx, plate.op, capi.VS_FP32, 3, 4)
    # beyond the scratch: eight FP64 right-hand sides of the prism need 8 x 25 MB
    lib = capi.load()
    n = koch.op.n
    xd = api.DeviceArray((8, n), np.complex128, ctx)
    yd = api.DeviceArray((8, n), np.complex128, ctx)
    rc = lib.vs_mvp_dev_noalloc(koch.op.h, C.c_void_p(xd.ptr), C.c_void_p(yd.ptr), 0, capi.VS_FP64, 8)
    assert rc == capi.VS_ERR_ARG
    assert b"scratch" in lib.vs_last_error()
    # the allocating entry point is unaffected and the no-allocation one still works afterwards
    _compare(ctx, plate.op, capi.VS_FP64, 1, 5)
