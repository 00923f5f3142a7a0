"""The particle workload at its stated sizes with the reference's per-incidence operators.

The reference builds one DDA operator per incidence (main_DDA.m:62,75), its LDR polarisability depending
on the direction khat and polarisation e through S = sum_j (khat_j e_j)^2 (getOPERATOR_DDA.m:15-26).  One
particle each of 48^3, 64^3 and 96^3 (seeded, randomly oriented plate / column / ellipsoid), 64 plane
waves each (32 Fibonacci directions x 2 polarisations), solved together by the batched FP32 solve inside
FP64 refinement with one diagonal 1/alpha_LDR(khat, e) per right-hand side:
* every right-hand side reaches true FP64 relative residual <= 1e-6 under its own incidence's operator,
  checked one incidence at a time with the plain FP64 matvec of an operator whose diagonal is that
  incidence's (vs_op_set_domain, vs_mvp);
* three right-hand sides equal single FP64 BiCGStab solves (2-level preconditioner built with their own
  alpha, tol 1e-10) to 5e-5 relative (a residual of 1e-6 bounds the error by the condition number);
* the same solutions do not satisfy the single-alpha operator of the scatterer's default incidence
  (the system the per-incidence diagonal replaces) to 1e-6, for at least one of those three.
"""
import numpy as np
import pytest

from voxscatter_amd import api, capi, geometry as G
from voxscatter_amd.dda import DDASystem
from voxscatter_amd.particles import incidence_diagonals

pytestmark = pytest.mark.gpu

CASES = [(48, "plate", 1.38 + 0.005j, 101), (64, "column", 1.52 + 0.005j, 102), (96, "ellipsoid", 1.6 + 0.005j, 103)]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n,shape,m,seed", CASES)
def test_particle_per_incidence_refined_batch(ctx, n, shape, m, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = G.random_particle(rng, n, shape, m)
    print(f"[particle] {p.name}: {p.idx.size} occupied voxels", flush=True)
    sysd = DDASystem.build(p, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    op = sysd.op
    B, meta = G.plane_wave_batch(p, ndir=32)
    diag = incidence_diagonals(p, meta)
    assert B.shape[0] == 64 and np.ptp(diag.real) > 0
    pc = api.Preconditioner.from_operator(op, 2, complex(diag.mean()), wscalar=sysd.dV, precisions=capi.VS_FP32)
    Bd = api.DeviceArray.from_numpy(B, np.complex128, ctx)
    Xd = api.DeviceArray(B.shape, np.complex128, ctx)
    flags, relres, outer, inner = api.refine_batch_dev(op, Bd, Xd, pc, tol=1e-6, inner_tol=1e-3, diag=diag)
    X = Xd.numpy()
    print(f"[particle] {p.name}: outer={outer} inner={inner} max reported relres={relres.max():.2e}", flush=True)
    assert np.all(flags == 0)
    # every right-hand side under its own incidence's FP64 operator
    worst = 0.0
    for j in range(64):
        op.set_domain(sysd.idx3, diag[j], 1.0, sysd.dV)
        rr = np.linalg.norm(B[j] - op.mvp(X[j])) / np.linalg.norm(B[j])
        worst = max(worst, rr)
        assert rr <= 1e-6, (j, rr)
    print(f"[particle] {p.name}: worst true relres over 64 incidences {worst:.2e}", flush=True)
    # three single FP64 solves with their own alpha
    wrong = []
    for j in (5, 37, 63):
        op.set_domain(sysd.idx3, diag[j], 1.0, sysd.dV)
        pcj = api.Preconditioner.from_operator(op, 2, diag[j], wscalar=sysd.dV, precisions=capi.VS_FP64)
        xj, fj, rj, ij, _ = api.bicgstab(op, B[j], 1e-10, 2000, pcj)
        pcj.close()
        err = _rel(X[j], xj)
        op.set_domain(sysd.idx3, 1 / sysd.alpha, 1.0, sysd.dV)
        rr_single = np.linalg.norm(B[j] - op.mvp(X[j])) / np.linalg.norm(B[j])
        wrong.append(rr_single)
        print(f"[particle] {p.name} rhs {j}: single FP64 solve flag={fj} iter={ij}, batch vs single {err:.2e}; "
              f"residual under the single-alpha operator {rr_single:.2e}", flush=True)
        assert fj == 0 and err <= 5e-5
    assert max(wrong) > 1e-6
    for h in (Bd, Xd, pc, op):
        h.close()


def test_matvec_per_rhs_diagonal_matches_per_incidence_operators(ctx):
    """vs_mvp_dev_diag on 4 right-hand sides equals 4 matvecs of operators whose diagonal is that vector's,
    in both branches: FP64 <= 1e-12, FP32 <= 2e-5 relative."""
    s = G.hex_plate()
    sysd = DDASystem.build(s, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    op = sysd.op
    rng = np.random.default_rng(9)
    X = rng.normal(size=(4, op.n)) + 1j * rng.normal(size=(4, op.n))
    diag = (1 / sysd.alpha) * (1 + 0.02 * rng.normal(size=4) + 0.01j * rng.normal(size=4))
    dd = api.DeviceArray.from_numpy(diag, np.complex128, ctx)
    for transp in (False, True):
        ref = []
        for j in range(4):
            op.set_domain(sysd.idx3, diag[j], 1.0, sysd.dV)
            ref.append(op.mvp(X[j], transp=transp))
        ref = np.stack(ref)
        for prec, dt, tol in ((capi.VS_FP64, np.complex128, 1e-12), (capi.VS_FP32, np.complex64, 2e-5)):
            xd = api.DeviceArray.from_numpy(X, dt, ctx)
            yd = api.DeviceArray(X.shape, dt, ctx)
            op.mvp_dev_diag(xd.ptr, yd.ptr, dd.ptr, transp, prec, 4)
            ctx.synchronize()
            err = _rel(yd.numpy().astype(np.complex128), ref)
            print(f"[diag matvec] transp={transp} prec={prec}: {err:.2e}", flush=True)
            assert err <= tol


def test_device_plane_wave_batch_matches_host(ctx):
    """vs_plane_wave_batch_dev against the host evaluation of PlaneWavePointWise.m:7-19 for 32 Fibonacci
    directions x 2 polarisations at the occupied voxels of a rotated particle: 1e-13 relative."""
    rng = np.random.Generator(np.random.PCG64(7))
    p = G.random_particle(rng, 32, "column", 1.45 + 0.005j)
    B, meta = G.plane_wave_batch(p, ndir=32)
    Bd, meta_d = G.plane_wave_batch_dev(p, ndir=32, ctx=ctx)
    assert len(meta_d) == len(meta) == 64
    for (k1, e1), (k2, e2) in zip(meta, meta_d):
        assert np.allclose(k1, k2, rtol=0, atol=0) and np.allclose(e1, e2, rtol=0, atol=0)
    err = np.abs(Bd.numpy() - B).max() / np.abs(B).max()
    print(f"[plane waves] device vs host max rel err {err:.2e}", flush=True)
    assert err <= 1e-13
