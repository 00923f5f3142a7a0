"""DDA operator assembly, FP64 (TEST INFRASTRUCTURE: see oracle/__init__.py).

Follows src/src_vie/src_operators/getOPERATOR_DDA.m, circulant_nop_const.m, gperiodic_coeff_nop.m and
fft_operator.m.  Arrays are MATLAB-ordered (x fastest); component order is [xx, xy, xz, yy, yz, zz].
"""
from __future__ import annotations

import numpy as np

# LDR coefficients, getOPERATOR_DDA.m:7-9
B1 = -1.8915316
B2 = 0.1648469
B3 = -1.7700004

# parity signs of the six components under reflection of each axis set (gperiodic_coeff_nop.m:4-31)
GP_COEFF = {
    "L": np.array([+1, -1, -1, +1, +1, +1], dtype=np.float64),
    "M": np.array([+1, -1, +1, +1, -1, +1], dtype=np.float64),
    "N": np.array([+1, +1, -1, +1, -1, +1], dtype=np.float64),
    "LM": np.array([+1, +1, -1, +1, -1, +1], dtype=np.float64),
    "LN": np.array([+1, -1, +1, +1, -1, +1], dtype=np.float64),
    "MN": np.array([+1, -1, -1, +1, +1, +1], dtype=np.float64),
    "LMN": np.array([+1, +1, +1, +1, +1, +1], dtype=np.float64),
}


def alpha_ldr(refInd, ko: float, d: float, kvec, Eo):
    """Lattice-dispersion-relation polarisability (getOPERATOR_DDA.m:4-26).

    ``refInd`` may be a scalar or an array (per-voxel index); the result has its shape.
    """
    kvec = np.asarray(kvec, dtype=np.float64)
    Eo = np.asarray(Eo, dtype=np.float64)
    a_hat = kvec / np.linalg.norm(kvec)
    e_hat = Eo / np.linalg.norm(Eo)
    S = 0.0
    for j in range(3):
        S = S + (a_hat[j] * e_hat[j]) ** 2
    msqr = np.asarray(refInd, dtype=np.complex128) ** 2
    dcube = d ** 3
    alpha_CM = 3 / (4 * np.pi) * (msqr - 1) / (msqr + 2)
    return alpha_CM / (1 + alpha_CM * ((B1 + msqr * B2 + msqr * B3 * S) * (ko * d) ** 2
                                       - 2 / 3 * 1j * ko ** 3 * dcube))


def green_dyadic(dr: np.ndarray, ko: float) -> np.ndarray:
    """Point dyadic of getOPERATOR_DDA.m:99-113 at offsets ``dr`` (..., 3) -> (..., 6); G(0) = 0."""
    dr = np.asarray(dr, dtype=np.float64)
    rjk = np.sqrt(np.sum(dr * dr, axis=-1))
    out = np.zeros(dr.shape[:-1] + (6,), dtype=np.complex128)
    nz = np.abs(rjk) > 1e-15
    rr = rjk[nz]
    rh = dr[nz] / rr[:, None]
    pre = np.exp(1j * ko * rr) / rr
    c1 = ko ** 2
    c2 = (1j * ko * rr - 1) / rr ** 2
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for p, (a, b) in enumerate(pairs):
        dab = 1.0 if a == b else 0.0
        rr_ab = rh[:, a] * rh[:, b]
        out[nz, p] = pre * (c1 * (dab - rr_ab) + c2 * (dab - 3 * rr_ab))
    return out


def toeplitz_dda(r: np.ndarray, ko: float) -> np.ndarray:
    """Toeplitz tensor Toep(i,j,k,:) = G(r(i,j,k) - r(1,1,1)) (getOPERATOR_DDA.m:28-119, quad 'off')."""
    R0 = r[0, 0, 0, :]
    return green_dyadic(r - R0, ko)


def toeplitz_dda_nearby(r: np.ndarray, ko: float, nQuad: int = 10) -> np.ndarray:
    """``nearby_quad='on'``: 10^3-point Gauss average of the dyadic for 0 < r < 5 dx (getOPERATOR_DDA.m:47-79)."""
    Toep = toeplitz_dda(r, ko)
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    xG, wG = np.polynomial.legendre.leggauss(nQuad)
    R0 = r[0, 0, 0, :]
    rjk = np.linalg.norm(r - R0, axis=-1)
    near = (rjk < 5 * dx) & (rjk > 1e-15)
    XG, YG, ZG = np.meshgrid(xG, xG, xG, indexing="xy")
    XW, YW, ZW = np.meshgrid(wG * 0.5, wG * 0.5, wG * 0.5, indexing="xy")
    W = (XW * YW * ZW).ravel()
    for (i, j, k) in zip(*np.nonzero(near)):
        R1 = r[i, j, k, :]
        pts = np.stack([R1[0] + dx / 2 * XG.ravel(), R1[1] + dx / 2 * YG.ravel(),
                        R1[2] + dx / 2 * ZG.ravel()], axis=-1)
        Toep[i, j, k, :] = (green_dyadic(pts - R0, ko) * W[:, None]).sum(axis=0)
    return Toep


def circulant_nop_const(G: np.ndarray) -> np.ndarray:
    """2L x 2M x 2N circulant embedding with parity signs (circulant_nop_const.m:5-40)."""
    L, M, N, _ = G.shape
    Gp = np.zeros((2 * L, 2 * M, 2 * N, 6), dtype=np.complex128)
    Gp[:L, :M, :N] = G
    rL = slice(L - 1, 0, -1)  # L:-1:2
    rM = slice(M - 1, 0, -1)
    rN = slice(N - 1, 0, -1)
    Gp[L + 1:, :M, :N] = G[rL, :, :] * GP_COEFF["L"]
    Gp[:L, M + 1:, :N] = G[:, rM, :] * GP_COEFF["M"]
    Gp[:L, :M, N + 1:] = G[:, :, rN] * GP_COEFF["N"]
    Gp[L + 1:, M + 1:, :N] = G[rL, rM, :] * GP_COEFF["LM"]
    Gp[L + 1:, :M, N + 1:] = G[rL, :, rN] * GP_COEFF["LN"]
    Gp[:L, M + 1:, N + 1:] = G[:, rM, rN] * GP_COEFF["MN"]
    Gp[L + 1:, M + 1:, N + 1:] = G[rL, rM, rN]
    return Gp


def fft_operator(A: np.ndarray) -> np.ndarray:
    """Unnormalised forward 3-D FFT of each component (fft_operator.m:1-8)."""
    return np.fft.fftn(A, axes=(0, 1, 2))


def getOPERATOR_DDA(r, ko, refInd, kvec, Eo, nearby_quad="off"):
    """``[op_out, Toep, alpha_LDR]`` of getOPERATOR_DDA.m:1-122."""
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    alpha = alpha_ldr(refInd, ko, dx, kvec, Eo)
    Toep = toeplitz_dda_nearby(r, ko) if nearby_quad == "on" else toeplitz_dda(r, ko)
    return fft_operator(circulant_nop_const(Toep)), Toep, alpha
