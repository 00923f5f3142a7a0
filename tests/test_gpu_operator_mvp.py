"""GPU parity of operator assembly (getOPERATOR_DDA) and the FFT matvec (MVP, both branches) against
the FP64 oracle.  Tolerances: FP64 1e-12 relative (round-off of a different FFT/summation order),
FP32 2e-5 relative."""
import math

import numpy as np
import pytest

from oracle import mvp as OM
from oracle import operator as O
from voxscatter_amd import api, capi, geometry as G

pytestmark = pytest.mark.gpu

SHAPES = [(3, 4, 5), (7, 6, 5), (32, 32, 8), (11, 13, 3), (83, 4, 3), (16, 24, 6)]


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _setup(shape, seed=0, frac=0.6, d=0.07):
    rng = np.random.default_rng(seed)
    L, M, N = shape
    r = G.centred_grid(L, M, N, d)
    nD = L * M * N
    vox = np.sort(rng.choice(nD, size=max(1, int(frac * nD)), replace=False))
    idx3 = np.concatenate([vox, nD + vox, 2 * nD + vox])
    x = rng.normal(size=idx3.size) + 1j * rng.normal(size=idx3.size)
    return r, vox, idx3, x


@pytest.mark.parametrize("shape", SHAPES)
def test_operator_assembly_matches_oracle(ctx, shape):
    print(f"[operator] shape={shape}", flush=True)
    k = 2 * math.pi
    r, _, _, _ = _setup(shape)
    op = api.Operator.dda(r, k, ctx=ctx)
    Toep = op.toeplitz()
    Tref = O.toeplitz_dda(r, k)
    assert _rel(Toep, Tref) <= 1e-14
    fN = op.fN()
    fref = O.fft_operator(O.circulant_nop_const(Tref))
    assert _rel(fN, fref) <= 1e-13


def test_operator_nearby_quadrature(ctx):
    k = 2 * math.pi
    r = G.centred_grid(7, 3, 2, 0.05)
    op = api.Operator.dda(r, k, nearby_quad="on", ctx=ctx)
    assert _rel(op.toeplitz(), O.toeplitz_dda_nearby(r, k)) <= 1e-13


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("transp", ["notransp", "transp"])
def test_mvp_matches_oracle(ctx, shape, transp):
    print(f"[mvp] shape={shape} {transp}", flush=True)
    k = 2 * math.pi
    r, vox, idx3, x = _setup(shape, seed=1)
    L, M, N = shape
    fN_op, Toep, alpha = api.getOPERATOR_DDA(r, k, 1.31 + 0.02j, [0, 0, k], [1, 0, 0], ctx=ctx)
    fref = O.fft_operator(O.circulant_nop_const(O.toeplitz_dda(r, k)))
    dV = 0.07 ** 3
    chi = np.zeros(L * M * N)
    chi[vox] = 1
    chi = chi.reshape((L, M, N), order="F")
    y = api.MVP(x, fN_op, np.ones((L, M, N)), dV * chi, 1 / alpha, transp, idx3)
    yref = OM.MVP(x, fref, np.ones((L, M, N)), dV * chi, 1 / alpha, transp, idx3)
    assert _rel(y, yref) <= 1e-12


def test_mvp_per_voxel_coefficients_and_external_fN(ctx):
    """Galerkin-style call: per-voxel Mr, Mc and Gram arrays, operator passed as a host fN array."""
    k = 2 * math.pi
    L, M, N = 9, 8, 6
    r, vox, idx3, x = _setup((L, M, N), seed=2, frac=1.0)
    rng = np.random.default_rng(3)
    fref = O.fft_operator(O.circulant_nop_const(O.toeplitz_dda(r, k)))
    Mr = 1 + rng.uniform(size=(L, M, N)) + 0.1j * rng.uniform(size=(L, M, N))
    Mc = (Mr - 1) / Mr
    Gram = 0.07 ** 3
    for flag in ("notransp", "transp"):
        y = api.MVP(x, fref, Mr, Mc, Gram, flag, idx3)
        yref = OM.MVP(x, fref, Mr, Mc, Gram, flag, idx3)
        assert _rel(y, yref) <= 1e-12
    # an in-place edit of the same host array (same address, head and tail unchanged) is a new operator
    f2 = fref.copy()
    y0 = api.MVP(x, f2, Mr, Mc, Gram, "notransp", idx3)
    mid = (L, M, N, 0)
    f2[mid] *= 1.5  # the Nyquist entry (its own mirror in x, y and z), away from the array's ends
    y1 = api.MVP(x, f2, Mr, Mc, Gram, "notransp", idx3)
    assert _rel(y1, OM.MVP(x, f2, Mr, Mc, Gram, "notransp", idx3)) <= 1e-12 and _rel(y1, y0) > 1e-6
    # an fN without the mirror parity the operator rebuilds its octants from is refused
    f3 = fref.copy()
    f3[1, 2, 3, 1] += 0.5 * np.abs(fref[..., 1]).max()
    with pytest.raises(capi.VoxScatterError) as ei:
        api.Operator.from_fN(f3, ctx=ctx)
    assert ei.value.code == capi.VS_ERR_ARG and "parity" in str(ei.value)


def test_mvp_fp32_and_multi_rhs(ctx):
    k = 2 * math.pi
    s = G.hex_plate()
    from voxscatter_amd.dda import DDASystem
    sysd = DDASystem.build(s, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    n = sysd.op.n
    rng = np.random.default_rng(4)
    X = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    yref = np.stack([sysd.op.mvp(X[i]) for i in range(3)])
    for prec, dt, tol in ((capi.VS_FP64, np.complex128, 1e-14), (capi.VS_FP32, np.complex64, 2e-5)):
        xt = api.DeviceArray.from_numpy(X, dt, ctx)
        yt = api.DeviceArray(X.shape, dt, ctx)
        sysd.op.mvp_dev(xt.ptr, yt.ptr, False, prec, 3)
        ctx.synchronize()
        assert _rel(yt.numpy().astype(np.complex128), yref) <= tol


def test_mvp_full_size_linearity_and_adjoint(ctx):
    """128x128x16 Koch prism: size-independent properties (linearity, <Ax,y> = <x,A^H y> for real Gram)."""
    s = G.koch_prism()
    from voxscatter_amd.dda import DDASystem
    sysd = DDASystem.build(s, ctx=ctx)
    op = sysd.op
    rng = np.random.default_rng(5)
    n = op.n
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = rng.normal(size=n) + 1j * rng.normal(size=n)
    a = 0.3 - 1.1j
    lhs = op.mvp(x + a * y)
    rhs = op.mvp(x) + a * op.mvp(y)
    assert _rel(lhs, rhs) <= 1e-13
    # with a real Gram the transp branch is the Hermitian adjoint
    op.set_domain(sysd.idx3, 2.5, 1.0, sysd.dV)
    Ax = op.mvp(x)
    AHy = op.mvp(y, transp=True)
    assert abs(np.vdot(y, Ax) - np.vdot(AHy, x)) <= 1e-12 * abs(np.vdot(y, Ax))


@pytest.mark.parametrize("shape", [(24, 24, 24), (32, 32, 32), (64, 64, 16), (32, 32, 8), (5, 16, 4), (64, 64, 64)])
@pytest.mark.parametrize("transp", ["notransp", "transp"])
def test_fused_slab_mvp_matches_oracle_and_unfused(ctx, shape, transp, monkeypatch):
    """Grids whose shape has fused slab kernels: every matvec variant (five passes, padded slab, parity
    split, parity split over two workgroups per slab; VS_MATVEC pins one, an unavailable one falls back
    to the autotuned choice) against the oracle
    in FP64, and in FP32 against the oracle and the five-pass FP32 result."""
    print(f"[fused mvp] shape={shape} {transp}", flush=True)
    k = 2 * math.pi
    r, vox, idx3, x = _setup(shape, seed=6)
    L, M, N = shape
    fref = O.fft_operator(O.circulant_nop_const(O.toeplitz_dda(r, k)))
    dV = 0.07 ** 3
    chi = np.zeros(L * M * N)
    chi[vox] = 1
    chi = chi.reshape((L, M, N), order="F")
    alpha = O.alpha_ldr(1.31 + 0.02j, k, 0.07, [0, 0, 1], [1, 0, 0])
    yref = OM.MVP(x, fref, np.ones((L, M, N)), dV * chi, 1 / alpha, transp, idx3)
    op = api.Operator.dda(r, k, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    op.set_domain(idx3, 1 / alpha, 1.0, dV * chi.reshape(-1, order="F"))
    out = {}
    for variant in ("pass5", "slab", "split", "split2"):
        monkeypatch.setenv("VS_MATVEC", variant)
        out[("fp64", variant)] = op.mvp(x, transp=transp == "transp")
        xt = api.DeviceArray.from_numpy(x[None], np.complex64, ctx)
        yt = api.DeviceArray((1, x.size), np.complex64, ctx)
        op.mvp_dev(xt.ptr, yt.ptr, transp == "transp", capi.VS_FP32, 1)
        out[("fp32", variant)] = yt.numpy()[0].astype(np.complex128)
    for variant in ("pass5", "slab", "split", "split2"):
        assert _rel(out[("fp64", variant)], yref) <= 1e-12, variant
        assert _rel(out[("fp32", variant)], yref) <= 2e-5, variant
        assert _rel(out[("fp32", variant)], out[("fp32", "pass5")]) <= 2e-5, variant


def test_fused_slab_mvp_koch_full_size_fp32(ctx, monkeypatch):
    """128x128x16 Koch prism in FP32 (fused slab kernels for this shape) against the FP64 matvec, for
    every variant."""
    s = G.koch_prism()
    from voxscatter_amd.dda import DDASystem
    sysd = DDASystem.build(s, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    op = sysd.op
    n = op.n
    rng = np.random.default_rng(7)
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    yref = np.stack([op.mvp(X[i]) for i in range(2)])
    xt = api.DeviceArray.from_numpy(X, np.complex64, ctx)
    yt = api.DeviceArray(X.shape, np.complex64, ctx)
    res = {}
    for variant in ("pass5", "slab", "split", "split2"):
        monkeypatch.setenv("VS_MATVEC", variant)
        op.mvp_dev(xt.ptr, yt.ptr, False, capi.VS_FP32, 2)
        res[variant] = yt.numpy().astype(np.complex128)
        assert _rel(res[variant], yref) <= 2e-5, variant
        assert _rel(res[variant], res["pass5"]) <= 2e-5, variant
