"""FFT matvec of the volume integral operator, FP64 (TEST INFRASTRUCTURE: see oracle/__init__.py).

Follows src/src_vie/src_mvp/MVP.m (identical to mv_AN_const.m): both the 'notransp' (MVP.m:101-133)
and 'transp' (MVP.m:65-99) branches, with the scatter (MVP.m:53-58) and gather (MVP.m:144) of the
non-air unknowns.  ``idx`` is 0-based into the 3 L M N column-major unknown vector.
"""
from __future__ import annotations

import numpy as np

from . import _fft

# symmetric 3x3 tensor from the six stored components [xx, xy, xz, yy, yz, zz] (MVP.m:108-123)
TENSOR_IDX = ((0, 1, 2), (1, 3, 4), (2, 4, 5))


def _field(a, shape):
    a = np.asarray(a)
    return a if a.ndim == 0 else a.reshape(shape, order="F")


def MVP(JIn0, fN, Mr, Mc, Gram, transp_flag, idx):
    """``JOut = MVP(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, 0)``.

    ``Mr`` fixes the domain shape (L, M, N) exactly as in the reference (MVP.m:42); ``Mc`` and ``Gram``
    may be scalars or L x M x N arrays.
    """
    LfN, MfN, NfN, _ = fN.shape
    Mr = np.asarray(Mr)
    L, M, N = Mr.shape
    nD = L * M * N
    JIn = np.zeros(3 * nD, dtype=np.complex128)
    JIn[np.asarray(idx)] = np.asarray(JIn0).ravel()
    JIn = JIn.reshape((L, M, N, 3), order="F")
    Mc = np.asarray(Mc)
    Gram = np.asarray(Gram)
    JOut = np.zeros((L, M, N, 3), dtype=np.complex128)
    s = (LfN, MfN, NfN)
    if transp_flag == "transp":
        Mcc = np.conj(Mc)
        acc = [0, 0, 0]
        for c in range(3):
            fJ = _fft.fftn(np.conj(Mcc * JIn[..., c]), s=s, axes=(0, 1, 2))
            for o in range(3):
                acc[o] = acc[o] + fN[..., TENSOR_IDX[c][o]] * fJ
        for o in range(3):
            J = _fft.ifftn(acc[o])
            JOut[..., o] = Gram * np.conj(Mr) * JIn[..., o] - np.conj(J[:L, :M, :N])
    elif transp_flag == "notransp":
        acc = [0, 0, 0]
        for c in range(3):
            fJ = _fft.fftn(JIn[..., c], s=s, axes=(0, 1, 2))
            for o in range(3):
                acc[o] = acc[o] + fN[..., TENSOR_IDX[c][o]] * fJ
        for o in range(3):
            J = _fft.ifftn(acc[o])
            JOut[..., o] = Gram * Mr * JIn[..., o] - Mc * J[:L, :M, :N]
    else:
        raise ValueError("transp_flag must be 'transp' or 'notransp'")
    return JOut.reshape(-1, order="F")[np.asarray(idx)]


def dense_dda_matrix(r, idx_vox, ko, alpha, dV):
    """Brute-force dense DDA matrix A = (1/alpha) I - dV G(r_p - r_q) on the occupied voxels.

    Independent of the FFT path: used to pin ``MVP`` on tiny grids.  Unknown ordering is
    [x(idx); y(idx); z(idx)] as in main_DDA.m:72.
    """
    from .operator import green_dyadic
    pts = r.reshape(-1, 3, order="F")[np.asarray(idx_vox)]
    n = pts.shape[0]
    G6 = green_dyadic(pts[:, None, :] - pts[None, :, :], ko)
    A = np.zeros((3 * n, 3 * n), dtype=np.complex128)
    for a in range(3):
        for b in range(3):
            A[a * n:(a + 1) * n, b * n:(b + 1) * n] = -dV * G6[..., TENSOR_IDX[a][b]]
    A += np.eye(3 * n) / alpha
    return A


def total_field_dda(P_sol, fN, alpha, dV, refInd, Einc, shape, idx3):
    """Total field everywhere in the box (main_DDA.m:136-150): E = Einc - MVP_all(P) + P/chi."""
    L, M, N = shape
    nD = L * M * N
    Pout = np.zeros(3 * nD, dtype=np.complex128)
    Pout[idx3] = P_sol
    CHI = (refInd ** 2 - 1) / (4 * np.pi)
    Eout = np.zeros(3 * nD, dtype=np.complex128)
    Eout[idx3] = P_sol / CHI
    idxAll = np.arange(3 * nD)
    temp = MVP(Pout, fN, np.ones((L, M, N)), dV * np.ones((L, M, N)), 1 / alpha, "notransp", idxAll)
    return np.asarray(Einc).reshape(-1, order="F") - temp + Eout
