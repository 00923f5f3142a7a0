"""GPU adjoint gradients of C_ext and of the forward-scattered intensity with respect to per-voxel
permittivity: against the oracle's dense FP64 adjoint, against central finite differences (h = 1e-6) of
device forward solves, and through the torch.autograd harness."""
import math

import numpy as np
import pytest

from oracle import adjoint as AD
from voxscatter_amd import geometry as G
from voxscatter_amd.adjoint import PerVoxelDDA

pytestmark = pytest.mark.gpu


def _eps(s, seed=0, rel=0.05):
    rng = np.random.default_rng(seed)
    return (s.refInd ** 2) * (1 + rel * rng.uniform(-1, 1, s.idx.size))


def test_gradient_matches_dense_oracle(ctx):
    s = G.hex_plate(6, 6, 2, m=1.31 + 0.01j, a_vox=3.0)
    eps = _eps(s)
    sysd = PerVoxelDDA.build(s, eps)
    c, gre, gim = sysd.cext_and_grad(1e-12)
    rv = s.r.reshape(-1, 3, order="F")[s.idx]
    cr, gre_r, gim_r = AD.cext_and_grad(rv, eps, sysd.Einc, s.k0, s.d, 0.0)
    assert abs(c - cr) <= 1e-10 * abs(cr)
    scale = max(np.abs(gre_r).max(), np.abs(gim_r).max())
    assert np.abs(gre - gre_r).max() <= 1e-8 * scale
    assert np.abs(gim - gim_r).max() <= 1e-8 * scale


@pytest.mark.parametrize("objective", ["cext", "fwd"])
def test_gradient_matches_central_differences(ctx, objective):
    """Perturbed hex plate 96 x 96 x 24 (eps = m^2 (1 + 0.05 U(-1,1)), m = 1.31 + 0.01i, seed 0); 64 voxels
    chosen with seed 1; central differences with h = 1e-6 on Re and Im eps."""
    s = G.perturbed_hex_plate(96, 96, 24, seed=0)
    L, M, N = s.shape
    eps = s.eps_r.reshape(-1, order="F")[s.idx]
    sysd = PerVoxelDDA.build(s, eps)
    tol = 1e-12
    f = sysd.cext_and_grad if objective == "cext" else sysd.fwd_intensity_and_grad
    val, gre, gim = f(tol)
    print(f"[adjoint] {objective}: value {val:.6e}, {s.idx.size} voxels", flush=True)

    def value(e):
        sy = PerVoxelDDA.build(s, e, op=sysd.op, pc=sysd.pc)
        P = sy.solve(sy.Einc, tol)
        if objective == "cext":
            return sy.cext(P)
        z = sy.forward_rows() @ P
        return float(np.sum(np.abs(z) ** 2))

    rng = np.random.default_rng(1)
    h = 1e-6
    scale = max(np.abs(gre).max(), np.abs(gim).max())
    worst = 0.0
    for j in rng.choice(s.idx.size, size=64, replace=False):
        for part, g in ((1.0, gre), (1j, gim)):
            e1 = eps.copy()
            e1[j] += h * part
            e2 = eps.copy()
            e2[j] -= h * part
            fd = (value(e1) - value(e2)) / (2 * h)
            worst = max(worst, abs(fd - g[j]) / scale)
    print(f"  worst |fd - adjoint| / max|grad| = {worst:.2e}")
    assert worst <= 1e-4


def test_torch_autograd_harness(ctx):
    torch = pytest.importorskip("torch")
    from voxscatter_amd.autograd import Extinction
    s = G.hex_plate(8, 8, 2, m=1.31 + 0.01j, a_vox=4.0)
    eps = _eps(s, seed=3)
    f = Extinction(s, tol=1e-12)
    er = torch.tensor(eps.real, dtype=torch.float64, requires_grad=True)
    ei = torch.tensor(eps.imag, dtype=torch.float64, requires_grad=True)
    c = f(er, ei)
    c.backward()
    rv = s.r.reshape(-1, 3, order="F")[s.idx]
    sysd = PerVoxelDDA.build(s, eps)
    cr, gre_r, gim_r = AD.cext_and_grad(rv, eps, sysd.Einc, s.k0, s.d, 0.0)
    assert abs(c.item() - cr) <= 1e-10 * abs(cr)
    assert np.allclose(er.grad.numpy(), gre_r, rtol=1e-7, atol=1e-9 * np.abs(gre_r).max())
    assert np.allclose(ei.grad.numpy(), gim_r, rtol=1e-7, atol=1e-9 * np.abs(gim_r).max())


def test_checkpointed_gmres_gradient(ctx):
    """GMRES-DR(10), 5 harmonic Ritz vectors, with the forward solve's deflation space starting the adjoint
    solve (Krylov-basis checkpoint), unpreconditioned high-index plate (m = 2.5 + 0.01i) so that the solves
    restart many times: the gradient equals the BiCGStab one, and the checkpointed adjoint solve needs
    fewer iterations than a cold one (the oracle's restatement gives 80 vs 91 residual norms)."""
    s = G.hex_plate(16, 16, 4, m=2.5 + 0.01j, a_vox=10.0)
    eps = _eps(s, seed=2)
    sysd = PerVoxelDDA.build(s, eps, precond_levels=0)
    c0, gre0, gim0 = sysd.cext_and_grad(1e-11)
    c1, gre1, gim1 = sysd.cext_and_grad(1e-11, solver="pgmresdr", restart=10, ritz=5)
    warm = sysd.last_nres
    scale = max(np.abs(gre0).max(), np.abs(gim0).max())
    assert abs(c1 - c0) <= 1e-8 * abs(c0)
    assert np.abs(gre1 - gre0).max() <= 1e-7 * scale and np.abs(gim1 - gim0).max() <= 1e-7 * scale
    sysd.solve(sysd.cext_vector(), 1e-11, solver="pgmresdr", restart=10, ritz=5)
    cold = sysd.last_nres
    print(f"[adjoint checkpoint] adjoint residual history: {warm} checkpointed vs {cold} cold", flush=True)
    assert warm < cold


def test_checkpointed_gmres_gradient_perturbed_plate(ctx):
    """The differentiation workload's own system (perturbed hex plate 96 x 96 x 24, 2-level preconditioner
    on the right) through GMRES-DR(10) with 5 harmonic Ritz vectors: the forward solve's deflation space
    (Krylov-basis checkpoint, vs_recycle) starts the adjoint solve.  Both gradients (C_ext and the forward
    intensity) equal the BiCGStab ones, and the checkpointed adjoint solve needs fewer residual norms than
    the same adjoint solve started cold."""
    s = G.perturbed_hex_plate(96, 96, 24, seed=0)
    eps = s.eps_r.reshape(-1, order="F")[s.idx]
    sysd = PerVoxelDDA.build(s, eps)
    tol = 1e-10
    for name, f in (("cext", sysd.cext_and_grad), ("fwd", sysd.fwd_intensity_and_grad)):
        v0, gre0, gim0 = f(tol)
        v1, gre1, gim1 = f(tol, solver="pgmresdr", restart=10, ritz=5)
        warm = sysd.last_nres
        scale = max(np.abs(gre0).max(), np.abs(gim0).max())
        e = max(np.abs(gre1 - gre0).max(), np.abs(gim1 - gim0).max()) / scale
        print(f"[adjoint checkpoint 96x96x24] {name}: value {v1:.6e} vs {v0:.6e}, gradient diff {e:.2e}",
              flush=True)
        assert abs(v1 - v0) <= 1e-8 * abs(v0)
        assert e <= 1e-6
    # the adjoint of the intensity objective again, cold (no checkpoint)
    P = sysd.solve(sysd.Einc, tol, solver="pgmresdr", restart=10, ritz=5)
    Bm = sysd.forward_rows()
    w = np.conj(Bm @ P) @ Bm
    sysd.solve(w, tol, solver="pgmresdr", restart=10, ritz=5)
    cold = sysd.last_nres
    print(f"[adjoint checkpoint 96x96x24] adjoint residual norms: {warm} checkpointed vs {cold} cold", flush=True)
    assert warm < cold
