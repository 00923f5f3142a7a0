"""Incident fields (host-side right-hand-side set-up).

* ``PlaneWavePointWise`` -- src/src_vie/src_fields/src_excitations/PlaneWavePointWise.m:7-19
  (E0 exp(+i k.r) sampled at the voxel centres, the DDA right-hand side).
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
