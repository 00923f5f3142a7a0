"""Far field, cross sections and the Mie series (TEST INFRASTRUCTURE: see oracle/__init__.py).

The reference has no far-field or cross-section routine (README.md:35).  These follow Draine (1988),
Draine & Flatau (1994) for the DDA quantities and Bohren & Huffman (1983, App. A) for the Mie series.
"""
from __future__ import annotations

import numpy as np
from scipy import special


def farfield(r_vox, P, dV, k, dirs):
    r = np.asarray(r_vox, dtype=np.float64).reshape(-1, 3)
    nv = r.shape[0]
    Pm = np.asarray(P).reshape(3, nv)
    d = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
    ph = np.exp(-1j * k * d @ r.T)                   # ndir x nv
    S = ph @ Pm.T * dV                                # ndir x 3: sum_j p_j exp(-i k n.r_j)
    return k * k * (S - d * np.sum(d * S, axis=1, keepdims=True))


def cross_sections(P, Einc, alpha_inv, dV, k, E0sq=1.0):
    P = np.asarray(P).reshape(3, -1)
    E = np.asarray(Einc).reshape(3, -1)
    ai = np.broadcast_to(np.asarray(alpha_inv), (P.shape[1],))
    p2 = np.sum(np.abs(P) ** 2, axis=0)
    cext = 4 * np.pi * k / E0sq * np.sum(np.imag(np.conj(E) * P)) * dV
    cabs = 4 * np.pi * k / E0sq * (np.sum(-ai.imag * p2) * dV - 2 / 3 * k ** 3 * np.sum(p2) * dV * dV)
    return cext, cabs, cext - cabs


def mie_efficiencies(x, m):
    """Q_ext, Q_sca of a homogeneous sphere of size parameter x and relative index m (BHMIE recurrences)."""
    nstop = int(np.ceil(x + 4 * x ** (1 / 3) + 2))
    n = np.arange(1, nstop + 1)
    mx = m * x
    # Riccati-Bessel functions via scipy (complex argument through the logarithmic derivative D_n)
    nmx = int(max(nstop, abs(mx)) + 16)
    D = np.zeros(nmx + 1, dtype=complex)
    for j in range(nmx, 0, -1):
        D[j - 1] = j / mx - 1 / (D[j] + j / mx)
    psi = x * special.spherical_jn(n, x)
    psi_m1 = x * special.spherical_jn(n - 1, x)
    chi = -x * special.spherical_yn(n, x)
    chi_m1 = -x * special.spherical_yn(n - 1, x)
    xi = psi - 1j * chi
    xi_m1 = psi_m1 - 1j * chi_m1
    Dn = D[1:nstop + 1]
    a = ((Dn / m + n / x) * psi - psi_m1) / ((Dn / m + n / x) * xi - xi_m1)
    b = ((Dn * m + n / x) * psi - psi_m1) / ((Dn * m + n / x) * xi - xi_m1)
    qext = 2 / x ** 2 * np.sum((2 * n + 1) * np.real(a + b))
    qsca = 2 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return qext, qsca
