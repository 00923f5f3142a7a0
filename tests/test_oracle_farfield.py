"""Pins of the oracle's far field / cross sections / Mie series: closed forms and limits."""
import math

import numpy as np

from oracle import farfield as FF


def test_single_dipole_far_field_closed_form():
    k = 2 * math.pi
    r0 = np.array([[0.1, -0.2, 0.3]])
    p = np.array([1.0 + 0.5j, -0.3j, 0.2])
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0], [0.6, 0.0, 0.8]])
    F = FF.farfield(r0, p, 1.0, k, dirs)
    for n, f in zip(dirs, F):
        ref = k * k * (p - n * np.dot(n, p)) * np.exp(-1j * k * np.dot(n, r0[0]))
        assert np.allclose(f, ref, rtol=1e-13)


def test_optical_theorem_identity():
    """Sum-form C_ext equals (4 pi/k) Im(e* . F(k_hat)) for any polarisation (E0 = 1)."""
    rng = np.random.default_rng(0)
    k = 2 * math.pi
    nv = 50
    r = rng.uniform(-0.5, 0.5, size=(nv, 3))
    P = rng.normal(size=3 * nv) + 1j * rng.normal(size=3 * nv)
    khat = np.array([0.0, 0.0, 1.0])
    e = np.array([1.0, 0.0, 0.0])
    Einc = (e[:, None] * np.exp(1j * k * r @ khat)[None, :]).reshape(-1)
    cext, _, _ = FF.cross_sections(P, Einc, 1.0, 0.01, k)
    F = FF.farfield(r, P, 0.01, k, khat[None, :])[0]
    assert abs(cext - 4 * math.pi / k * np.imag(np.conj(e) @ F)) <= 1e-10 * abs(cext)


def test_mie_limits():
    m = 1.5
    # Rayleigh limit
    x = 1e-2
    qe, qs = FF.mie_efficiencies(x, m)
    K = (m * m - 1) / (m * m + 2)
    assert abs(qs - 8 / 3 * x ** 4 * abs(K) ** 2) <= 1e-3 * qs
    # lossless: extinction equals scattering
    for x in (0.5, 1.0, 3.0, 10.0):
        qe, qs = FF.mie_efficiencies(x, m)
        assert abs(qe - qs) <= 1e-10 * qe
    # absorbing Rayleigh sphere: Q_abs = 4 x Im(K)
    mc = 1.5 + 0.1j
    qe, qs = FF.mie_efficiencies(1e-3, mc)
    Kc = (mc * mc - 1) / (mc * mc + 2)
    assert abs((qe - qs) - 4e-3 * Kc.imag) <= 1e-3 * 4e-3 * Kc.imag
    # extinction paradox: Q_ext -> 2 for large spheres
    qe, _ = FF.mie_efficiencies(300.0, 1.5 + 0.01j)
    assert abs(qe - 2.0) < 0.05
