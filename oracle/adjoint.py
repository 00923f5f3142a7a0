"""Dense FP64 adjoint of the DDA system with per-voxel permittivity (TEST INFRASTRUCTURE: see
oracle/__init__.py).  No reference counterpart (the reference has no cross sections or gradients);
pinned by central finite differences in tests/test_oracle_adjoint.py."""
from __future__ import annotations

import math

import numpy as np

from .mvp import TENSOR_IDX
from .operator import B1, B2, B3, green_dyadic


def alpha_inv_ldr(eps, k, d, S):
    eps = np.asarray(eps, dtype=np.complex128)
    kd = k * d
    ainv = (4 * math.pi / 3) * (eps + 2) / (eps - 1) + (B1 + eps * (B2 + B3 * S)) * kd ** 2 - 2j / 3 * kd ** 3
    dainv = -4 * math.pi / (eps - 1) ** 2 + (B2 + B3 * S) * kd ** 2
    return ainv, dainv


def dense_matrix(r_vox, k, ainv, dV):
    """A = diag(1/alpha_j) (3 components) - dV G(r_p - r_q), unknowns [x(idx); y(idx); z(idx)]."""
    pts = np.asarray(r_vox)
    n = pts.shape[0]
    G6 = green_dyadic(pts[:, None, :] - pts[None, :, :], k)
    A = np.zeros((3 * n, 3 * n), dtype=np.complex128)
    for a in range(3):
        for b in range(3):
            A[a * n:(a + 1) * n, b * n:(b + 1) * n] = -dV * G6[..., TENSOR_IDX[a][b]]
    A += np.diag(np.tile(np.asarray(ainv, dtype=np.complex128), 3))
    return A


def cext(P, Einc, k, dV):
    return float(np.imag(4 * math.pi * k * dV * np.conj(Einc) @ P))


def cext_and_grad(r_vox, eps, Einc, k, d, S):
    dV = d ** 3
    ainv, dainv = alpha_inv_ldr(eps, k, d, S)
    A = dense_matrix(r_vox, k, ainv, dV)
    P = np.linalg.solve(A, Einc)
    c = 4 * math.pi * k * dV * np.conj(Einc)
    lam = np.linalg.solve(A.T, c)
    n = len(eps)
    g = -dainv * (lam.reshape(3, n) * P.reshape(3, n)).sum(axis=0)
    return cext(P, Einc, k, dV), g.imag, g.real


def cext_of_eps(r_vox, eps, Einc, k, d, S):
    ainv, _ = alpha_inv_ldr(eps, k, d, S)
    P = np.linalg.solve(dense_matrix(r_vox, k, ainv, d ** 3), Einc)
    return cext(P, Einc, k, d ** 3)
