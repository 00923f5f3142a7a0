"""Many particles, many incident fields: the orientation-averaging workload of the solve loop.

Each particle gets its own DDA operator (main_DDA.m:61-75) and 2-level Chan preconditioner
(main_DDA.m:98-111); its 2 * ndir right-hand sides (PlaneWavePointWise.m:7-19 for ndir Fibonacci-sphere
directions, two polarisations each, evaluated on the device) are solved together by one batched FP32 solve inside FP64
iterative refinement (``vs_refine_batch_dev``), so every matvec and preconditioner launch of a particle
serves all of its right-hand sides.  Particles run one after another on the context's stream; a
particle's set-up (operator, inverse blocks) is timed apart from its solve.

The reference builds one DDA operator per incidence (main_DDA.m:62,75): the LDR polarisability depends
on the incidence direction and polarisation through S = sum_j (khat_j e_j)^2 (getOPERATOR_DDA.m:15-26).
Here the tensor is shared and each right-hand side carries its own diagonal 1/alpha_LDR(khat, e)
(``vs_refine_batch_diag_dev``), so every solution is that of its own incidence's operator; the shared
preconditioner uses the mean of those diagonals.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .api import Context, DeviceArray, Preconditioner, alpha_LDR, get_context, refine_batch_dev
from .capi import VS_FP32, VS_FP64
from .dda import DDASystem
from .geometry import Scatterer, plane_wave_batch_dev


def incidence_diagonals(scat: Scatterer, meta) -> np.ndarray:
    """1/alpha_LDR of each (khat, e) of a plane-wave batch: the operator diagonal of each incidence."""
    r = scat.r
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    return np.array([1.0 / alpha_LDR(scat.refInd, scat.k0, dx, scat.k0 * np.asarray(kh), e) for kh, e in meta])


@dataclass
class ParticleResult:
    name: str
    grid: tuple
    unknowns: int
    nrhs: int
    flags: np.ndarray
    relres: np.ndarray          # true FP64 relative residual of each right-hand side
    outer: int                  # refinement steps
    inner: int                  # FP32 BiCGStab iterations over all refinement steps
    setup_s: float              # operator + preconditioner build
    solve_s: float              # the batched refined solve, inputs resident on the device
    X: np.ndarray | None = field(default=None, repr=False)


def solve_particles(parts: list[Scatterer], ndir: int = 32, tol: float = 1e-6, inner_tol: float = 1e-3,
                    max_outer: int = 10, inner_maxit: int = 2000, ctx: Context | None = None,
                    keep_solutions: bool = False, on_done=None) -> list[ParticleResult]:
    """Solve every particle for its 2 * ndir plane waves to FP64 relative residual ``tol``.

    ``on_done(i, result)`` is called after each particle (progress reporting); ``keep_solutions`` returns
    the (2 ndir, unknowns) solutions on the host.
    """
    ctx = ctx or get_context()
    out = []
    for i, p in enumerate(parts):
        t0 = time.perf_counter()
        sysd = DDASystem.build(p, precisions=VS_FP64 | VS_FP32, ctx=ctx)
        Bd, meta = plane_wave_batch_dev(p, ndir=ndir, ctx=ctx)
        diag = incidence_diagonals(p, meta)
        pc = Preconditioner.from_operator(sysd.op, 2, complex(diag.mean()), wscalar=sysd.dV, precisions=VS_FP32)
        Xd = DeviceArray(Bd.shape, np.complex128, ctx)
        ctx.synchronize()
        t1 = time.perf_counter()
        flags, relres, outer, inner = refine_batch_dev(sysd.op, Bd, Xd, pc, tol=tol, max_outer=max_outer,
                                                       inner_tol=inner_tol, inner_maxit=inner_maxit, diag=diag)
        ctx.synchronize()
        t2 = time.perf_counter()
        res = ParticleResult(p.name, tuple(p.r.shape[:3]), int(Bd.shape[1]), int(Bd.shape[0]), flags, relres,
                             outer, inner, t1 - t0, t2 - t1, Xd.numpy() if keep_solutions else None)
        Bd.close()
        Xd.close()
        pc.close()
        sysd.op.close()
        out.append(res)
        if on_done is not None:
            on_done(i, res)
    return out
