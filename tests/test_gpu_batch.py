"""Batched multi-right-hand-side solves: every system of the batch against its own single solve, and
the FP32 + FP64-refinement batch against the FP64 true residual."""
import numpy as np
import pytest

from voxscatter_amd import api, geometry as G
from voxscatter_amd.capi import VS_FP32, VS_FP64
from voxscatter_amd.dda import DDASystem

pytestmark = pytest.mark.gpu


def _setup(ctx, precisions=VS_FP64):
    s = G.hex_plate()
    sysd = DDASystem.build(s, precisions=precisions, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=precisions)
    B, meta = G.plane_wave_batch(s, ndir=4)
    return s, sysd, pc, B


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_batched_bicgstab_matches_single_solves(ctx):
    s, sysd, pc, B = _setup(ctx)
    k, n = B.shape
    Bd = api.DeviceArray.from_numpy(B, np.complex128, ctx)
    Xd = api.DeviceArray((k, n), np.complex128, ctx)
    flags, relres, iters = api.bicgstab_batch_dev(sysd.op, Bd, Xd, 1e-8, 2000, pc)
    X = Xd.numpy()
    print(f"  batch: flags={flags} iters={iters}")
    assert np.all(flags == 0) and np.all(relres <= 1e-8)
    for i in range(k):
        # the alpha used in the system depends on the polarisation only through S (0 for every pair here)
        r = np.linalg.norm(sysd.op.mvp(X[i]) - B[i]) / np.linalg.norm(B[i])
        assert r <= 1e-8
        x1, f1, rr1, it1, rv1 = api.bicgstab(sysd.op, B[i], 1e-8, 2000, pc)
        assert f1 == 0 and _rel(X[i], x1) <= 1e-6


def test_batched_refinement(ctx):
    s, sysd, pc, B = _setup(ctx, VS_FP64 | VS_FP32)
    k, n = B.shape
    Bd = api.DeviceArray.from_numpy(B, np.complex128, ctx)
    Xd = api.DeviceArray((k, n), np.complex128, ctx)
    flags, relres, outer, inner = api.refine_batch_dev(sysd.op, Bd, Xd, pc, tol=1e-6, inner_tol=1e-3)
    X = Xd.numpy()
    print(f"  refine batch: flags={flags} outer={outer} inner iterations={inner} relres={relres.max():.2e}")
    assert np.all(flags == 0)
    for i in range(k):
        assert np.linalg.norm(sysd.op.mvp(X[i]) - B[i]) / np.linalg.norm(B[i]) <= 1e-6


def test_fp32_batch_and_particle_generator(ctx):
    parts = G.particle_batch(count=3, seed=0, sizes=(24,))
    for p in parts:
        assert p.idx.size > 0.1 * 24 ** 3
    p = parts[0]
    sysd = DDASystem.build(p, precisions=VS_FP64 | VS_FP32, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=VS_FP32)
    B, _ = G.plane_wave_batch(p, ndir=2)
    Bd = api.DeviceArray.from_numpy(B, np.complex64, ctx)
    Xd = api.DeviceArray(B.shape, np.complex64, ctx)
    flags, relres, iters = api.bicgstab_batch_dev(sysd.op, Bd, Xd, 1e-4, 500, pc)
    assert np.all(flags == 0) and np.all(relres <= 2e-4)


@pytest.mark.parametrize("nrhs", [3, 4, 9])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_multi_rhs_matvec_and_preconditioner(ctx, nrhs, prec):
    """The nrhs-wide matvec and preconditioner applies (right-hand-side chunks of 1, 4 and 8, ragged
    last chunk) give each column's single apply."""
    precision = VS_FP64 if prec == "fp64" else VS_FP32
    dt = np.complex128 if prec == "fp64" else np.complex64
    tol = 1e-13 if prec == "fp64" else 1e-6
    s = G.hex_plate()
    sysd = DDASystem.build(s, precisions=precision, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=precision)
    n = sysd.op.n
    rng = np.random.default_rng(nrhs)
    X = (rng.normal(size=(nrhs, n)) + 1j * rng.normal(size=(nrhs, n))).astype(dt)
    Xd = api.DeviceArray.from_numpy(X, dt, ctx)
    Yd = api.DeviceArray((nrhs, n), dt, ctx)
    x1 = api.DeviceArray((1, n), dt, ctx)
    y1 = api.DeviceArray((1, n), dt, ctx)
    for name, fn in (("matvec", lambda a, b, k: sysd.op.mvp_dev(a, b, precision=precision, nrhs=k)),
                     ("precond", lambda a, b, k: pc.apply_dev(a, b, precision=precision, nrhs=k))):
        fn(Xd.ptr, Yd.ptr, nrhs)
        Y = Yd.numpy()
        for i in range(nrhs):
            x1.upload(X[i:i + 1])
            fn(x1.ptr, y1.ptr, 1)
            y = y1.numpy()[0]
            err = np.linalg.norm(Y[i] - y) / np.linalg.norm(y)
            assert err <= tol, (name, i, err)


def test_particle_workload_refined_batches(ctx):
    """Orientation-averaging workload: seeded particles of different grids (24^3, 32^3), 64 plane waves
    each (32 Fibonacci directions x 2 polarisations), one batched FP32 + FP64-refinement solve per
    particle, each plane wave with its own incidence's diagonal 1/alpha_LDR(khat, e); every right-hand side
    reaches FP64 relative residual 1e-6, checked again with a fresh FP64 operator of that incidence, and
    the first one equals the plain FP64 BiCGStab solution of its own operator to the tolerance."""
    from voxscatter_amd.particles import incidence_diagonals, solve_particles
    parts = [p for p in G.particle_batch(count=5, seed=0) if p.r.shape[0] <= 32]
    assert sorted({p.r.shape[0] for p in parts}) == [24, 32]
    res = solve_particles(parts, ndir=32, tol=1e-6, ctx=ctx, keep_solutions=True)
    for p, r in zip(parts, res):
        print(f"[particles] {r.name} grid={r.grid} unknowns={r.unknowns} nrhs={r.nrhs} outer={r.outer} "
              f"inner={r.inner} max relres={r.relres.max():.2e} solve={r.solve_s * 1e3:.1f} ms", flush=True)
        assert r.nrhs == 64 and np.all(r.flags == 0) and np.all(r.relres <= 1e-6)
        sysd = DDASystem.build(p, ctx=ctx)
        B, meta = G.plane_wave_batch(p, ndir=32)
        diag = incidence_diagonals(p, meta)
        for j in (0, 37, 63):
            sysd.op.set_domain(sysd.idx3, diag[j], 1.0, sysd.dV)
            assert _rel(sysd.op.mvp(r.X[j]), B[j]) <= 1.01e-6
        sysd.op.set_domain(sysd.idx3, diag[0], 1.0, sysd.dV)
        pc0 = api.Preconditioner.from_operator(sysd.op, 2, diag[0], wscalar=sysd.dV)
        x, flag, relres, it, _ = api.bicgstab(sysd.op, B[0], 1e-6, 2000, pc0)
        assert flag == 0
        # both solutions are within 1e-6 (relative residual) of the exact one
        assert _rel(r.X[0], x) <= 1e-4


def test_batch_edge_cases(ctx):
    """A zero right-hand side (x = 0, flag 0), a duplicated one (the same result from another slot, also
    after finished systems have been dropped from the batch), a scaled one, and a batch beyond the
    per-call limit (an error before any work)."""
    s, sysd, pc, B = _setup(ctx)
    n = B.shape[1]
    Bx = np.stack([B[0], np.zeros(n), B[1], B[0], 2 * B[1], B[2], B[3]])
    k = Bx.shape[0]
    Bd = api.DeviceArray.from_numpy(Bx, np.complex128, ctx)
    Xd = api.DeviceArray((k, n), np.complex128, ctx)
    flags, relres, iters = api.bicgstab_batch_dev(sysd.op, Bd, Xd, 1e-8, 2000, pc)
    X = Xd.numpy()
    print(f"  edge batch: flags={flags} iters={iters}")
    assert np.all(flags == 0)
    assert np.all(X[1] == 0) and relres[1] == 0
    assert _rel(X[3], X[0]) <= 1e-12 and iters[3] == iters[0]
    assert _rel(X[4], 2 * X[2]) <= 1e-6
    big_b = api.DeviceArray((200, n), np.complex128, ctx)
    big_x = api.DeviceArray((200, n), np.complex128, ctx)
    with pytest.raises(api.VoxScatterError, match="too many right-hand sides"):
        api.bicgstab_batch_dev(sysd.op, big_b, big_x, 1e-8, 10, pc)
