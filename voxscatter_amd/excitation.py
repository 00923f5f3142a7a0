"""Incident fields (host-side right-hand-side set-up).

* ``PlaneWavePointWise`` -- src/src_vie/src_fields/src_excitations/PlaneWavePointWise.m:7-19
  (E0 exp(+i k.r) sampled at the voxel centres, the DDA right-hand side).
* ``PlaneWave_Excitation`` -- src/src_vie/src_fields/src_excitations/PlaneWave_Excitation.m:39-99
  (E0 exp(-i k.r) averaged over each voxel, the Galerkin right-hand side).
"""
from __future__ import annotations

import numpy as np


def PlaneWavePointWise(Eo, kvec, r) -> np.ndarray:
    """(L, M, N, 3) incident field E0 exp(i k.r) at the voxel centres."""
    Eo = np.asarray(Eo, dtype=np.float64)
    kvec = np.asarray(kvec, dtype=np.float64)
    kr = kvec[0] * r[..., 0] + kvec[1] * r[..., 1] + kvec[2] * r[..., 2]
    e = np.exp(1j * kr)
    return np.stack([Eo[0] * e, Eo[1] * e, Eo[2] * e], axis=-1)


def PlaneWave_Excitation(r, kvec, Eo) -> np.ndarray:
    """(L, M, N, 3) voxel average of E0 exp(-i k.r): per axis exp(-i k_a r_a)(exp(-i k_a d/2) - exp(i k_a d/2))
    / (-i k_a), or d when k_a = 0, divided by d^3 (engineering time convention of the Galerkin kernel)."""
    r = np.asarray(r, dtype=np.float64)
    kvec = np.asarray(kvec, dtype=np.float64)
    Eo = np.asarray(Eo, dtype=np.complex128)
    d = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    fac = np.ones(r.shape[:3], dtype=np.complex128)
    for a in range(3):
        ka = kvec[a]
        if ka != 0.0:
            fac = fac * (np.exp(-1j * ka * r[..., a]) * (np.exp(-1j * ka * d / 2) - np.exp(1j * ka * d / 2)) / (-1j * ka))
        else:
            fac = fac * d
    fac /= d ** 3
    return np.stack([Eo[0] * fac, Eo[1] * fac, Eo[2] * fac], axis=-1)
