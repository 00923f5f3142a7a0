"""Adjoint gradients of DDA objectives with respect to the per-voxel relative permittivity eps_j.

The DDA system with per-voxel polarisability is A(eps) P = Einc with A = diag(1/alpha_j) - dV G, G the
symmetric point-dyadic interaction, so A^T = A: the adjoint solve A lam = w reuses the forward operator
(its 'notransp' matvec) and preconditioner.  For a holomorphic objective z = w^T P,
    dz/deps_j = g_j = -(d(1/alpha)/deps)_j sum_c lam_{c,j} P_{c,j}      (device kernel vs_dda_grad_eps)
(the forward solve's GMRES-DR deflation space can start the adjoint solve: a Krylov-basis checkpoint),
and for real objectives built on it:
    C_ext = Im(c^T P), c = 4 pi k dV conj(Einc)/|E0|^2:   dC/dRe eps = Im g,   dC/dIm eps = Re g
    I_fwd = sum_c |b_c^T P|^2 (forward far field):         w = sum_c conj(b_c^T P) b_c,
                                                           dI/dRe eps = 2 Re g, dI/dIm eps = -2 Im g
1/alpha_LDR(eps) = 1/alpha_CM + (b1 + eps (b2 + b3 S)) (kd)^2 - 2/3 i (kd)^3, alpha_CM = 3/(4 pi) (eps-1)/(eps+2)
(getOPERATOR_DDA.m:7-26 with m^2 = eps), so d(1/alpha)/deps = -4 pi / (eps - 1)^2 + (b2 + b3 S) (kd)^2.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import api
from .api import Operator, Preconditioner, get_context
from .capi import VS_FP64, check, cplx_arr, dptr
from .excitation import PlaneWavePointWise
from .geometry import Scatterer

B1, B2, B3 = -1.8915316, 0.1648469, -1.7700004


def _S(kvec, Eo):
    a = np.asarray(kvec, float) / np.linalg.norm(kvec)
    e = np.asarray(Eo, float) / np.linalg.norm(Eo)
    return float(np.sum((a * e) ** 2))


def alpha_inv_ldr(eps, k, d, kvec, Eo):
    """(1/alpha_LDR, d(1/alpha_LDR)/deps) for relative permittivity eps (scalar or array)."""
    eps = np.asarray(eps, dtype=np.complex128)
    S = _S(kvec, Eo)
    kd = k * d
    ainv = (4 * math.pi / 3) * (eps + 2) / (eps - 1) + (B1 + eps * (B2 + B3 * S)) * kd ** 2 - 2j / 3 * kd ** 3
    dainv = -4 * math.pi / (eps - 1) ** 2 + (B2 + B3 * S) * kd ** 2
    return ainv, dainv


@dataclass
class PerVoxelDDA:
    """DDA system with per-voxel permittivity on the occupied voxels of a scatterer."""
    scat: Scatterer
    op: Operator
    pc: Preconditioner | None
    eps: np.ndarray          # eps at the occupied voxels
    ainv: np.ndarray         # 1/alpha_j
    dainv: np.ndarray        # d(1/alpha)/deps at the voxels
    dV: float
    idx3: np.ndarray
    Einc: np.ndarray         # [Ex; Ey; Ez] at the occupied voxels (= b)
    kvec: np.ndarray

    @classmethod
    def build(cls, scat: Scatterer, eps_vox, precond_levels: int = 2, op: Operator | None = None,
              pc: Preconditioner | None = None):
        """``op`` / ``pc`` may be reused from an earlier build on the same grid (``pc`` is only a
        preconditioner: a slightly different mean polarisability does not change the solution)."""
        r = scat.r
        L, M, N, _ = r.shape
        nD = L * M * N
        d = r[1, 0, 0, 0] - r[0, 0, 0, 0]
        kvec = scat.k0 * np.asarray(scat.dInc, float)
        eps_vox = np.asarray(eps_vox, dtype=np.complex128)
        ainv, dainv = alpha_inv_ldr(eps_vox, scat.k0, d, kvec, scat.Eo)
        if op is None:
            op = Operator.dda(r, scat.k0)
        idx = np.asarray(scat.idx, np.int64)
        idx3 = np.concatenate([idx, nD + idx, 2 * nD + idx])
        gram = np.ones(nD, dtype=np.complex128) * np.mean(ainv)
        gram[idx] = ainv
        dV = d ** 3
        op.set_domain(idx3, gram.reshape((L, M, N), order="F"), 1.0, dV)
        if pc is None and precond_levels:
            # circulant approximation with the mean polarisability on the diagonal
            pc = Preconditioner.from_operator(op, precond_levels, np.mean(ainv), wscalar=dV)
        Einc = PlaneWavePointWise(scat.Eo, kvec, r).reshape(-1, order="F")[idx3]
        return cls(scat, op, pc, eps_vox, ainv, dainv, dV, idx3, Einc, kvec)

    def solve(self, rhs, tol=1e-10, maxit=2000, solver="bicgstab", recycle: api.Recycle | None = None,
              restart=30, ritz=10):
        """A x = rhs by BiCGStab (main_DDA.m) or by GMRES-DR with the preconditioner on the right; ``recycle``
        (GMRES-DR only) is a Krylov-basis checkpoint that this solve starts from when it holds a space and
        leaves its own deflation space in.  The residual-history length is kept in ``last_nres``."""
        if solver == "bicgstab":
            x, flag, relres, it, resvec = api.bicgstab(self.op, rhs, tol, maxit, self.pc)
        elif solver == "pgmresdr":
            x, flag, relres, it, resvec = api.pgmresDR(self.op, rhs, restart, tol, max(1, maxit // restart), ritz,
                                                       R1=self.pc, recycle=recycle)
        else:
            raise ValueError(f"unknown solver {solver!r}")
        self.last_nres = len(resvec)
        if flag != 0:
            raise RuntimeError(f"{solver} did not converge (flag {flag}, relres {relres:.2e})")
        return x

    def _checkpoint(self, solver, checkpoint):
        return api.Recycle(self.op.ctx) if solver == "pgmresdr" and checkpoint else None

    def grad_holomorphic(self, P, lam):
        ctx = get_context()
        nv = self.eps.size
        g = np.zeros(nv, dtype=np.complex128)
        Pv, lv, da = cplx_arr(P, 3 * nv), cplx_arr(lam, 3 * nv), cplx_arr(self.dainv, nv)
        check(ctx.lib.vs_dda_grad_eps(ctx.h, nv, dptr(Pv), dptr(lv), dptr(da), dptr(g)))
        return g

    # ---- objectives -----------------------------------------------------------------------------
    def cext_vector(self, E0sq=1.0):
        return 4 * math.pi * self.scat.k0 * self.dV / E0sq * np.conj(self.Einc)

    def cext(self, P, E0sq=1.0):
        return float(np.imag(self.cext_vector(E0sq) @ P))

    def cext_and_grad(self, tol=1e-10, solver="bicgstab", checkpoint=True, restart=30, ritz=10):
        """C_ext and (dC/dRe eps, dC/dIm eps) at every occupied voxel (one forward + one adjoint solve).
        With ``solver="pgmresdr"`` and ``checkpoint`` the forward solve's deflation space starts the
        adjoint solve (same operator: A^T = A)."""
        rec = self._checkpoint(solver, checkpoint)
        P = self.solve(self.Einc, tol, solver=solver, recycle=rec, restart=restart, ritz=ritz)
        c = self.cext_vector()
        lam = self.solve(c, tol, solver=solver, recycle=rec, restart=restart, ritz=ritz)
        g = self.grad_holomorphic(P, lam)
        return self.cext(P), g.imag.copy(), g.real.copy()

    def forward_rows(self):
        """b_c (3 x n): F_c(k_hat) = b_c^T P, F = k^2 dV sum_j (I - k k) P_j exp(-i k k.r_j)."""
        k = self.scat.k0
        khat = self.kvec / np.linalg.norm(self.kvec)
        rv = self.scat.r.reshape(-1, 3, order="F")[self.scat.idx]
        ph = np.exp(-1j * k * rv @ khat)
        nv = rv.shape[0]
        Bm = np.zeros((3, 3 * nv), dtype=np.complex128)
        proj = np.eye(3) - np.outer(khat, khat)
        for c in range(3):
            for cc in range(3):
                Bm[c, cc * nv:(cc + 1) * nv] = k * k * self.dV * proj[c, cc] * ph
        return Bm

    def fwd_intensity_and_grad(self, tol=1e-10, solver="bicgstab", checkpoint=True, restart=30, ritz=10):
        """|F(k_hat)|^2 and its gradients (dI/dRe eps, dI/dIm eps); ``solver`` / ``checkpoint`` as in
        ``cext_and_grad``."""
        rec = self._checkpoint(solver, checkpoint)
        P = self.solve(self.Einc, tol, solver=solver, recycle=rec, restart=restart, ritz=ritz)
        Bm = self.forward_rows()
        z = Bm @ P
        w = np.conj(z) @ Bm
        lam = self.solve(w, tol, solver=solver, recycle=rec, restart=restart, ritz=ritz)
        g = self.grad_holomorphic(P, lam)
        return float(np.sum(np.abs(z) ** 2)), 2 * g.real, -2 * g.imag
