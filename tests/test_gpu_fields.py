"""GPU far field and cross sections: parity with the oracle, energy balance of a lossless plate, and
the DDA sphere against the Mie series (the reference has no such routine, README.md:35)."""
import math

import numpy as np
import pytest

from oracle import farfield as FF
from voxscatter_amd import api, fields, geometry as G
from voxscatter_amd.dda import DDASystem
from voxscatter_amd.excitation import PlaneWavePointWise

pytestmark = pytest.mark.gpu


def test_farfield_and_cross_sections_match_oracle(ctx):
    rng = np.random.default_rng(1)
    k = 2 * math.pi
    nv = 300
    r = rng.uniform(-0.5, 0.5, size=(nv, 3))
    P = rng.normal(size=3 * nv) + 1j * rng.normal(size=3 * nv)
    E = rng.normal(size=3 * nv) + 1j * rng.normal(size=3 * nv)
    dirs, _ = fields.sphere_quadrature(6, 8)
    F = fields.farfield(r, P, 0.001, k, dirs)
    Fr = FF.farfield(r, P, 0.001, k, dirs)
    assert np.abs(F - Fr).max() <= 1e-12 * np.abs(Fr).max()
    ai = rng.normal(size=nv) + 1j * rng.normal(size=nv)
    got = fields.cross_sections(P, E, ai, 0.001, k)
    ref = FF.cross_sections(P, E, ai, 0.001, k)
    assert np.allclose(got, ref, rtol=1e-12)


def _solve(s, tol=1e-9, levels=2):
    sysd = DDASystem.build(s)
    pc = sysd.preconditioner(levels)
    x, flag, relres, it, resvec = api.bicgstab(sysd.op, sysd.b, tol, 2000, pc)
    assert flag == 0
    return sysd, x


def test_lossless_plate_energy_balance(ctx):
    """m real: C_abs vanishes (LDR radiative reaction), and C_ext = integral of |F|^2 (optical theorem)."""
    s = G.hex_plate()
    sysd, P = _solve(s)
    nv = s.idx.size
    rv = s.r.reshape(-1, 3, order="F")[s.idx]
    Einc = sysd.b
    cext, cabs, csca = fields.cross_sections(P, Einc, 1 / sysd.alpha, sysd.dV, s.k0)
    assert abs(cabs) <= 1e-9 * cext
    dirs, w = fields.sphere_quadrature(64, 128)
    F = fields.farfield(rv, P, sysd.dV, s.k0, dirs)
    csca_int = float(np.sum(w * np.sum(np.abs(F) ** 2, axis=1)))
    print(f"  C_ext={cext:.6e} C_sca(int)={csca_int:.6e} C_abs={cabs:.2e}")
    assert abs(csca_int - cext) <= 1e-5 * cext


def test_sphere_against_mie(ctx):
    m = 1.5 + 0.0j
    n = 16
    k = 2 * math.pi
    d = 1.0 / (8 * k)            # sphere radius 8 d: x = 1
    s = G.sphere(n, m, d)
    sysd, P = _solve(s)
    nv = s.idx.size
    a_eq = (3 * nv * d ** 3 / (4 * math.pi)) ** (1 / 3)
    cext, cabs, csca = fields.cross_sections(P, sysd.b, 1 / sysd.alpha, sysd.dV, k)
    qe_dda = cext / (math.pi * a_eq ** 2)
    qe_mie, qs_mie = FF.mie_efficiencies(k * a_eq, m)
    print(f"  Q_ext DDA={qe_dda:.5f} Mie={qe_mie:.5f} (x={k * a_eq:.3f}, {nv} dipoles)")
    assert abs(qe_dda - qe_mie) <= 0.03 * qe_mie


def test_total_field_matches_oracle(ctx):
    from oracle import mvp as OM
    from oracle import operator as O
    s = G.hex_plate(12, 12, 4, a_vox=6.0)
    sysd, P = _solve(s)
    L, M, N = s.shape
    Einc = PlaneWavePointWise(s.Eo, sysd.kvec, s.r)
    E = fields.total_field_dda(sysd.op, P, Einc, s.refInd, sysd.alpha, sysd.dV, sysd.idx3)
    fref, _, _ = O.getOPERATOR_DDA(s.r, s.k0, s.refInd, sysd.kvec, s.Eo)
    Eref = OM.total_field_dda(P, fref, sysd.alpha, sysd.dV, s.refInd, Einc, (L, M, N), sysd.idx3)
    assert np.linalg.norm(E.reshape(-1, order="F") - Eref) <= 1e-12 * np.linalg.norm(Eref)
    # the operator's solve domain is restored
    assert sysd.op.n == sysd.idx3.size
