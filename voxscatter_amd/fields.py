"""Fields of a DDA or Galerkin solution, far field and cross sections of a DDA solution.

* ``total_field_dda``   main_DDA.m:136-150: E = Einc - MVP_all(P) + P/chi, evaluated with the device
                        matvec over every voxel of the box (idxAll, Mc = dV, Gram = 1/alpha).
* ``total_field_galerkin`` E_field_Nop_const.m:62-84 (mv_L_const.m): Einc + (N J - Gram J)/(Gram j w eps0),
                        the same full-box device matvec with Gram -> -Gram, Mc = -1.
* ``farfield`` / ``cross_sections``: device kernels (no reference counterpart, README.md:35).
* ``sphere_quadrature`` Gauss-Legendre in cos(theta) x trapezoid in phi, for C_sca = int |F|^2 dOmega.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .api import Context, Operator, get_context
from .capi import check, cplx_arr, dptr


def total_field_dda(op: Operator, P_sol, Einc, refInd, alpha, dV, idx3):
    """Total field on the whole L x M x N x 3 box; restores the operator's domain afterwards."""
    L, M, N = op.shape
    nD = L * M * N
    saved = getattr(op, "_domain_args", None)
    Pout = np.zeros(3 * nD, dtype=np.complex128)
    Pout[idx3] = P_sol
    CHI = (refInd ** 2 - 1) / (4 * np.pi)
    Eout = np.zeros(3 * nD, dtype=np.complex128)
    Eout[idx3] = np.asarray(P_sol) / CHI
    op.set_domain(np.arange(3 * nD), 1.0 / alpha, 1.0, dV)
    temp = op.mvp(Pout)
    E = np.asarray(Einc).reshape(-1, order="F") - temp + Eout
    if saved is not None:
        op.set_domain(*saved)
    return E.reshape((L, M, N, 3), order="F")


def total_field_galerkin(op: Operator, J, Einc, Gram, ko, idx3):
    """E_field_Nop_const.m:62-84 with mv_L_const.m: Einc + (N J - Gram J) / (Gram j omega eps0) on the whole
    L x M x N x 3 box, omega eps0 = ko / eta0, for J given on the unknowns idx3.  N J - Gram J is one
    full-grid matvec with the operator's diagonal set to -Gram and its contrast to -1; the operator's
    domain is restored afterwards."""
    from .galerkin import ETA0
    L, M, N = op.shape
    nD = L * M * N
    Jf = np.zeros(3 * nD, dtype=np.complex128)
    Jf[idx3] = np.asarray(J, dtype=np.complex128).ravel()
    Ei = np.asarray(Einc, dtype=np.complex128).reshape(-1, order="F")
    if not np.any(Jf):
        return Ei.reshape((L, M, N, 3), order="F")
    saved = getattr(op, "_domain_args", None)
    op.set_domain(np.arange(3 * nD), -Gram, 1.0, -1.0)
    y = op.mvp(Jf)
    if saved is not None:
        op.set_domain(*saved)
    E = Ei + y / (Gram * 1j * ko / ETA0)
    return E.reshape((L, M, N, 3), order="F")


def farfield(r_vox, P, dV, k, dirs, ctx: Context | None = None) -> np.ndarray:
    """F(n) for unit vectors ``dirs`` (ndir, 3); ``r_vox`` (nvox, 3), ``P`` = [Px; Py; Pz] (3 nvox)."""
    ctx = ctx or get_context()
    r = np.ascontiguousarray(np.asarray(r_vox, dtype=np.float64).reshape(-1, 3))
    nvox = r.shape[0]
    Pv = cplx_arr(P, 3 * nvox)
    d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float64).reshape(-1, 3))
    F = np.zeros(3 * d.shape[0], dtype=np.complex128)
    check(ctx.lib.vs_dda_farfield(ctx.h, nvox, dptr(r), dptr(Pv), float(dV), float(k), d.shape[0], dptr(d),
                                  dptr(F)))
    return F.reshape(-1, 3)


def cross_sections(P, Einc_vox, alpha_inv, dV, k, E0sq=1.0, ctx: Context | None = None):
    """``(C_ext, C_abs, C_sca)``; Einc_vox = [Ex; Ey; Ez] at the occupied voxels, alpha_inv scalar or per voxel."""
    ctx = ctx or get_context()
    Pv = cplx_arr(P)
    nvox = Pv.size // 3
    E = cplx_arr(Einc_vox, 3 * nvox)
    a = cplx_arr(alpha_inv)
    out = np.zeros(3)
    check(ctx.lib.vs_dda_cross_sections(ctx.h, nvox, dptr(Pv), dptr(E), dptr(a), a.size, float(dV), float(k),
                                        float(E0sq), dptr(out)))
    return float(out[0]), float(out[1]), float(out[2])


def sphere_quadrature(nth: int = 48, nph: int = 96):
    """Directions (nth*nph, 3) and weights summing to 4 pi."""
    x, w = np.polynomial.legendre.leggauss(nth)
    ph = 2 * np.pi * np.arange(nph) / nph
    ct = np.repeat(x, nph)
    st = np.sqrt(1 - ct ** 2)
    pp = np.tile(ph, nth)
    dirs = np.stack([st * np.cos(pp), st * np.sin(pp), ct], axis=-1)
    wt = np.repeat(w, nph) * (2 * np.pi / nph)
    return dirs, wt
