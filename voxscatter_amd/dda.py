"""The DDA solve sequence of main_DDA.m:49-133 on the device.

``DDASystem.build`` performs main_DDA.m:49-75 (incident field, operator, MVP handle), and
``DDASystem.preconditioner`` main_DDA.m:95-116 (Chan approximation and inverse blocks), every step on the
GPU.  The unknowns are [Px(idx); Py(idx); Pz(idx)] (main_DDA.m:72).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import Context, Operator, Preconditioner, alpha_LDR
from .capi import VS_FP64
from .excitation import PlaneWavePointWise
from .geometry import Scatterer


@dataclass
class DDASystem:
    scat: Scatterer
    op: Operator
    alpha: complex
    dV: float
    idx3: np.ndarray
    b: np.ndarray          # Vrhs = Einc(idx3)
    kvec: np.ndarray

    @classmethod
    def build(cls, scat: Scatterer, precisions: int = VS_FP64, ctx: Context | None = None) -> "DDASystem":
        r = scat.r
        L, M, N, _ = r.shape
        nD = L * M * N
        kvec = scat.k0 * np.asarray(scat.dInc, dtype=np.float64)
        Einc = PlaneWavePointWise(scat.Eo, kvec, r)
        dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
        alpha = alpha_LDR(scat.refInd, scat.k0, dx, kvec, scat.Eo)
        op = Operator.dda(r, scat.k0, "off", precisions, ctx)
        dV = dx ** 3
        idx = np.asarray(scat.idx, dtype=np.int64)
        idx3 = np.concatenate([idx, nD + idx, 2 * nD + idx])
        # Mc = dV*chi equals dV on every retained unknown (chi(idx) = 1), so it is passed as a scalar
        op.set_domain(idx3, 1.0 / alpha, 1.0, dV)
        b = Einc.reshape(-1, order="F")[idx3]
        return cls(scat, op, alpha, dV, idx3, b, kvec)

    def preconditioner(self, levels: int = 2, precisions: int = VS_FP64) -> Preconditioner:
        """Inverse blocks inv(1/alpha I - dV circ) of the 1- or 2-level Chan approximation."""
        return Preconditioner.from_operator(self.op, levels, 1.0 / self.alpha, wscalar=self.dV, precisions=precisions)
