"""GPU parity at the full Koch-prism size (128 x 128 x 16, level-3 Koch snowflake, kR = 20,
m = 1.31 + 0.01i, incidence +x, polarisation +z: the solve-loop benchmark workload) against the FP64
oracle.

The fused matvec kernels of this shape (the parity-split slab kernels built for 128-point y lines) and
the mirrored, packed 2-level preconditioner run here exactly as in the benchmark.  Bars:
* every matvec variant (five passes, padded slab, parity split, parity split over two workgroups),
  FP64, both MVP.m branches (MVP.m:65-133): <= 1e-12 relative to oracle.mvp.MVP; FP32 <= 2e-5;
* the FP64 2-level apply (mvp_circ_2_level.m:2-22) with the full oracle block set: <= 1e-11;
* FP64 BiCGStab + 2-level preconditioner (main_koch.m:127): identical length(resvec) and flag, and
  the solution within 1e-10 relative of the oracle's.
"""
import numpy as np
import pytest

from oracle import circulant as Ch
from oracle import krylov as OK
from oracle import mvp as OM
from oracle import operator as O
from voxscatter_amd import api, capi, geometry as G
from voxscatter_amd.dda import DDASystem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.fixture(scope="module")
def koch(ctx):
    s = G.koch_prism()
    L, M, N = s.shape
    print(f"[koch full size] grid {s.shape}, {s.idx.size} occupied voxels", flush=True)
    sysd = DDASystem.build(s, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    Tref = O.toeplitz_dda(s.r, s.k0)
    fref = O.fft_operator(O.circulant_nop_const(Tref))
    chi = np.zeros(L * M * N)
    chi[s.idx] = 1
    chi = chi.reshape((L, M, N), order="F")
    return dict(s=s, sysd=sysd, Tref=Tref, fref=fref, chi=chi, shape=(L, M, N))


@pytest.mark.parametrize("transp", ["notransp", "transp"])
def test_every_matvec_variant_matches_oracle(ctx, koch, transp, monkeypatch):
    sysd = koch["sysd"]
    L, M, N = koch["shape"]
    op = sysd.op
    n = op.n
    rng = np.random.default_rng(11)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    yref = OM.MVP(x, koch["fref"], np.ones((L, M, N)), sysd.dV * koch["chi"], 1 / sysd.alpha, transp, sysd.idx3)
    xt = api.DeviceArray.from_numpy(x[None], np.complex64, ctx)
    yt = api.DeviceArray((1, n), np.complex64, ctx)
    for variant in ("pass5", "slab", "split", "split2"):
        print(f"[koch full size] matvec {variant} {transp}", flush=True)
        monkeypatch.setenv("VS_MATVEC", variant)
        y64 = op.mvp(x, transp=transp == "transp")
        e64 = _rel(y64, yref)
        op.mvp_dev(xt.ptr, yt.ptr, transp == "transp", capi.VS_FP32, 1)
        e32 = _rel(yt.numpy()[0].astype(np.complex128), yref)
        print(f"[koch full size]   fp64 rel err {e64:.2e}, fp32 rel err {e32:.2e}", flush=True)
        assert e64 <= 1e-12, variant
        assert e32 <= 2e-5, variant


def test_two_level_apply_matches_oracle(ctx, koch):
    sysd = koch["sysd"]
    L, M, N = koch["shape"]
    _, cL = Ch.circ_1_level(koch["Tref"], L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    inv = Ch.invert_blocks_dda(blk, sysd.alpha, sysd.dV)
    koch["inv"] = inv
    pc = sysd.preconditioner(2)
    rng = np.random.default_rng(12)
    x = rng.normal(size=sysd.op.n) + 1j * rng.normal(size=sysd.op.n)
    y = pc.apply(x)
    yref = Ch.mvp_circ_2_level(inv, x, L, M, N, sysd.idx3)
    err = _rel(y, yref)
    print(f"[koch full size] 2-level apply rel err {err:.2e}", flush=True)
    assert err <= 1e-11


def test_fp64_bicgstab_two_level_matches_oracle(ctx, koch):
    sysd = koch["sysd"]
    L, M, N = koch["shape"]
    inv = koch.get("inv")
    if inv is None:
        _, cL = Ch.circ_1_level(koch["Tref"], L, M, N)
        _, blk = Ch.circ_2_level(cL, L, M, N)
        inv = Ch.invert_blocks_dda(blk, sysd.alpha, sysd.dV)
    ones = np.ones((L, M, N))
    Mc = sysd.dV * koch["chi"]
    fref = koch["fref"]

    def afun(v):
        return OM.MVP(v, fref, ones, Mc, 1 / sysd.alpha, "notransp", sysd.idx3)

    def mfun(v):
        return Ch.mvp_circ_2_level(inv, v, L, M, N, sysd.idx3)

    pc = sysd.preconditioner(2)
    xg, fg, rg, ig, resg = api.bicgstab(sysd.op, sysd.b, 1e-6, 2000, pc)
    xr, fr, rr, ir, resr = OK.bicgstab(afun, sysd.b, 1e-6, 2000, mfun)
    sol = _rel(xg, xr)
    print(f"[koch full size] bicgstab gpu iter={ig} len(resvec)={len(resg)}; oracle iter={ir} "
          f"len(resvec)={len(resr)}; solution rel diff {sol:.2e}", flush=True)
    assert fg == fr == 0
    assert ig == ir and len(resg) == len(resr)
    assert sol <= 1e-10
    # the true FP64 residual of the GPU solution, with the oracle's operator
    assert np.linalg.norm(sysd.b - afun(xg)) / np.linalg.norm(sysd.b) <= 1e-6
