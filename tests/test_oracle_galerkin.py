"""Oracle pins for the Galerkin N operator: the face-face cubature against the reference's own C++
(cubature_nop_const.cpp compiled in oracle/_ref), the DIRECTFN fixtures against the same build, the
projection table and the voxel-averaged plane wave against closed forms, the total field against the
constitutive relation of an exact solution."""
import json
import math
import os

import numpy as np
import pytest

from oracle import galerkin as OG

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not OG.ref_available(), reason="oracle/_ref not built (oracle/build_ref.sh)")


@needs_ref
@pytest.mark.parametrize("n1d,m", [(4, (3, 1, 2)), (4, (7, 0, 0)), (10, (2, 3, 4)), (10, (0, 2, 0)), (20, (1, 1, 0))])
def test_cubature_matches_reference_cpp(n1d, m):
    ko, dx = 2 * math.pi, 1 / 35
    with np.errstate(divide="ignore", invalid="ignore"):
        ours = OG.cubature_nop(n1d, m, ko, dx)
        ref = OG.reference_cubature(n1d, m, ko, dx)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(ours))
    assert np.allclose(ours[fin], ref[fin], rtol=1e-13, atol=0)


@needs_ref
def test_singular_fixtures_are_the_reference_values():
    fx = json.load(open(os.path.join(HERE, "golden", "galerkin_singular.json")))
    for rec in fx[:2]:
        vals = OG.reference_singular(rec["ko"], rec["dx"], 20)
        ref = np.array([complex(*v) for v in rec["directfn_20"]])
        assert np.allclose(vals, ref, rtol=1e-14, atol=0)


def test_projection_table_closed_forms():
    P = OG.vector_projection()
    # face normals -x, +x, -y, +y, -z, +z; components xx, xy, xz, yy, yz, zz
    assert np.all(P[0:2, :, 0] == 0) and np.all(P[:, 0:2, 0] == 0)       # n x x = 0 on the x faces
    assert P[3, 3, 0] == 1 and P[2, 3, 0] == -1                           # (+y x x).(+y x x), (-y x x).(+y x x)
    assert P[5, 5, 3] == 1 and P[4, 5, 3] == -1                           # yy on the z faces
    for c in (0, 3, 5):                                                   # diagonal components symmetric in faces
        assert np.array_equal(P[:, :, c], P[:, :, c].T)


def test_plane_wave_excitation_limits():
    L, M, N, d = 4, 3, 2, 0.1
    xs = [(np.arange(n) - (n - 1) / 2) * d for n in (L, M, N)]
    r = np.stack(np.meshgrid(*xs, indexing="ij"), axis=-1)
    Eo = np.array([0.0, 0.0, 1.0])
    E = OG.plane_wave_excitation(r, [0.0, 0.0, 0.0], Eo)
    assert np.allclose(E[..., 2], 1.0) and np.allclose(E[..., :2], 0.0)
    k = np.array([2 * math.pi, 0.0, 0.0])
    E = OG.plane_wave_excitation(r, k, Eo)
    # average of exp(-ikx) over [x - d/2, x + d/2] = exp(-ikx) sinc(kd/2)
    ref = np.exp(-1j * k[0] * r[..., 0]) * np.sin(k[0] * d / 2) / (k[0] * d / 2)
    assert np.allclose(E[..., 2], ref, rtol=1e-13)


def test_static_self_term_closed_form_in_fixture_scale():
    # the exact static self term of a square of side a: a^3 (4/3 (1 - sqrt 2) + 4 ln(1 + sqrt 2));
    # the reference's DIRECTFN at Np = 64 is within 1e-6 of it (its Np = 20 value within 5e-5)
    fx = json.load(open(os.path.join(HERE, "golden", "galerkin_singular.json")))
    exact = 4 / 3 * (1 - math.sqrt(2)) + 4 * math.log(1 + math.sqrt(2))
    rec = [r for r in fx if r["ko"] < 1e-6][0]
    assert abs(rec["directfn_64"][0][0] - exact) / exact < 1e-6
    assert abs(rec["directfn_20"][0][0] - exact) / exact < 5e-5


def test_total_field_constitutive_relation():
    """E_field_Nop_const of an exact JVIE solution equals J / (j omega eps0 (eps - 1)) in the scatterer,
    and mv_L_const agrees with MVP.m on the whole box with Gram -> -Gram, Mc = -1."""
    from oracle import mvp as OM
    from oracle import operator as O
    L, M, N = 3, 3, 2
    ko, dx = 2 * math.pi, 1 / 35
    fx = json.load(open(os.path.join(HERE, "golden", "galerkin_singular.json")))
    sing = np.array([complex(*v) for v in fx[0]["directfn_20"]])
    T = OG.assembly_nop_const(L, M, N, dx, ko, sing)
    fN = O.fft_operator(O.circulant_nop_const(T))
    rng = np.random.default_rng(5)
    eps = (2.0 + rng.uniform(size=(L, M, N))) * (1 + 0.01j)
    eps[0, 0, 0] = 1.0  # one air voxel
    nD = L * M * N
    Mc = eps - 1
    idx = np.flatnonzero(np.abs(Mc.reshape(-1, order="F")) > 1e-12)
    idx3 = np.concatenate([idx, nD + idx, 2 * nD + idx])
    Gram = dx ** 3
    xs = [(np.arange(n) - (n - 1) / 2) * dx for n in (L, M, N)]
    r = np.stack(np.meshgrid(*xs, indexing="ij"), axis=-1)
    Einc = OG.plane_wave_excitation(r, ko * np.array([1.0, 0, 0]), np.array([0, 0, 1.0]))
    we0 = ko / OG.ETA0
    b = (Gram * 1j * we0 * np.concatenate([Mc.reshape(-1, order="F")] * 3) * Einc.reshape(-1, order="F"))[idx3]
    er3 = np.concatenate([eps.reshape(-1, order="F")] * 3)[idx3]
    ones = np.ones((L, M, N))

    def afun(v):  # form 2: Gram J - (eps - 1)/eps N J = b / eps
        return OM.MVP(v, fN, ones, (Mc / eps), Gram, "notransp", idx3)

    A = np.stack([afun(e) for e in np.eye(idx3.size)], axis=1)
    J = np.linalg.solve(A, b / er3)
    Jf = np.zeros(3 * nD, dtype=complex)
    Jf[idx3] = J
    Je = Jf.reshape((L, M, N, 3), order="F")
    E = OG.E_field_Nop_const(Je, fN, Gram, ko, Einc)
    inside = (np.abs(Mc) > 1e-12)
    expect = Je / (1j * we0 * Mc[..., None])
    assert np.max(np.abs(E[inside] - expect[inside])) <= 1e-10 * np.max(np.abs(expect[inside]))
    # the air voxel carries the incident plus the scattered field, which is not small here
    assert np.linalg.norm(E[0, 0, 0] - Einc[0, 0, 0]) > 1e-3 * np.linalg.norm(Einc[0, 0, 0])
    mvl = OG.mv_L_const(Je, fN, Gram).reshape(-1, order="F")
    alt = OM.MVP(Jf, fN, ones, -np.ones((L, M, N)), -Gram, "notransp", np.arange(3 * nD))
    assert np.allclose(mvl, alt, rtol=0, atol=1e-12 * np.abs(alt).max())
    # zero current: the incident field unchanged
    assert np.array_equal(OG.E_field_Nop_const(0 * Je, fN, Gram, ko, Einc), Einc)
