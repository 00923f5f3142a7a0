"""The oracle's dense adjoint gradient of C_ext against central finite differences."""
import math

import numpy as np

from oracle import adjoint as AD
from voxscatter_amd import geometry as G
from voxscatter_amd.excitation import PlaneWavePointWise


def test_dense_adjoint_matches_central_differences():
    s = G.hex_plate(6, 6, 2, m=1.31 + 0.01j, a_vox=3.0)
    rng = np.random.default_rng(0)
    nv = s.idx.size
    eps = (s.refInd ** 2) * (1 + 0.05 * rng.uniform(-1, 1, nv))
    k = s.k0
    d = s.d
    kvec = k * s.dInc
    rv = s.r.reshape(-1, 3, order="F")[s.idx]
    Einc = PlaneWavePointWise(s.Eo, kvec, s.r).reshape(-1, 3, order="F")[s.idx].T.reshape(-1)
    S = 0.0
    c, gre, gim = AD.cext_and_grad(rv, eps, Einc, k, d, S)
    h = 1e-6
    for j in rng.choice(nv, size=6, replace=False):
        for part, g in ((1.0, gre), (1j, gim)):
            e1 = eps.copy()
            e1[j] += h * part
            e2 = eps.copy()
            e2[j] -= h * part
            fd = (AD.cext_of_eps(rv, e1, Einc, k, d, S) - AD.cext_of_eps(rv, e2, Einc, k, d, S)) / (2 * h)
            assert abs(fd - g[j]) <= 1e-6 * max(abs(g).max(), 1e-30), (j, part, fd, g[j])
