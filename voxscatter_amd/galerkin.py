"""The Galerkin (JVIE, piecewise-constant basis) solve sequence on the device.

``GalerkinSystem.build`` performs JVIE_Solver_const.m:133-157,243-251 with form 2 (the unknowns are
the equivalent currents J on the non-air voxels, Gram = dx^3, Mr = 1, Mc = (eps - 1)/eps, right-hand
side Gram j omega eps0 (eps - 1)/eps E_inc with the voxel-averaged plane wave of
PlaneWave_Excitation.m), with the N operator of getOPERATORS.m assembled on the GPU.
``GalerkinSystem.preconditioner`` builds the 2-level Chan preconditioner of scatteringHumanHead.m:
155-193: inv(Gram I - diag(face3_2) C) with the contrast averaged over x and y.
``GalerkinSystem.total_field`` is E_field_Nop_const.m (scatteringHumanHead.m:116-120).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import Context, Operator, Preconditioner
from .capi import VS_FP64
from .excitation import PlaneWave_Excitation

ETA0 = 3.767303134617706e2  # free-space impedance (EMconstants.m)


def averaged_contrast(eps):
    """face3_2 of scatteringHumanHead.m:155-176: (e - 1)/e of the permittivity averaged over x then y, per z,
    repeated for the three components (length 3N)."""
    eps = np.asarray(eps)
    L, M, N = eps.shape
    e = eps.sum(axis=0) / L
    e = e.sum(axis=0) / M
    f = (e - 1) / e
    return np.concatenate([f, f, f])


@dataclass
class GalerkinSystem:
    op: Operator
    eps: np.ndarray
    dx: float
    ko: float
    idx3: np.ndarray
    b: np.ndarray
    Gram: float
    Einc: np.ndarray  # voxel-averaged incident field, L x M x N x 3

    @classmethod
    def build(cls, eps, dx: float, ko: float, kdir=(1.0, 0.0, 0.0), Eo=(0.0, 0.0, 1.0), singular=None,
              precisions: int = VS_FP64, ctx: Context | None = None) -> "GalerkinSystem":
        eps = np.asarray(eps, dtype=np.complex128)
        L, M, N = eps.shape
        nD = L * M * N
        xs = [(np.arange(n) - (n - 1) / 2) * dx for n in (L, M, N)]
        X, Y, Z = np.meshgrid(*xs, indexing="ij")
        r = np.stack([X, Y, Z], axis=-1)
        kvec = ko * np.asarray(kdir, dtype=np.float64)
        Einc = PlaneWave_Excitation(r, kvec, Eo)
        op = Operator.galerkin((L, M, N), dx, ko, singular, precisions, ctx)
        Mc = eps - 1.0
        idx = np.flatnonzero(np.abs(Mc.reshape(-1, order="F")) > 1e-12)
        idx3 = np.concatenate([idx, nD + idx, 2 * nD + idx])
        Gram = dx ** 3
        omega_eps0 = ko / ETA0
        tau = 1j * omega_eps0 * Mc
        rhs = Gram * np.concatenate([tau.reshape(-1, order="F")] * 3) * Einc.reshape(-1, order="F")
        # form 2: Mc <- Mc/Mr, Vrhs <- Vrhs/Mr, Mr <- 1
        er3 = np.concatenate([eps.reshape(-1, order="F")] * 3)
        b = rhs[idx3] / er3[idx3]
        op.set_domain(idx3, Gram, 1.0, (Mc / eps).reshape(-1, order="F"))
        return cls(op, eps, dx, ko, idx3, b, Gram, Einc)

    def preconditioner(self, levels: int = 2, precisions: int = VS_FP64) -> Preconditioner:
        """inv(Gram I - diag(face3_2) C) with the x/y-averaged contrast (scatteringHumanHead.m:178-184)."""
        if levels != 2:
            raise ValueError("the averaged-contrast Galerkin preconditioner is 2-level")
        w = averaged_contrast(self.eps)
        return Preconditioner.from_operator(self.op, 2, self.Gram, w=w, precisions=precisions)

    def total_field(self, J) -> np.ndarray:
        """Total field on the whole box (E_field_Nop_const.m) for the solution J on the non-air unknowns."""
        from .fields import total_field_galerkin
        return total_field_galerkin(self.op, J, self.Einc, self.Gram, self.ko, self.idx3)
