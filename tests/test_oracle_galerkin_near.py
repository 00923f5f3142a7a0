"""Independent pins of the oracle's Galerkin near box (nearby_correct.m:1-143, assembly_nop_const.m:96-146).

1. The (code, keep) table of every cell of the 2 x 2 x 2 near box is derived from the face geometry of
   cube_faces.m alone: two closed unit squares (observation face on the cell at offset m, source face on
   the cell at the origin, as the cubature pairs them) that coincide take ST, that share an edge take EAc
   (coplanar) or EAo (orthogonal), that share only a vertex take VAc / VAo, and disjoint ones keep the
   cubature.  The derived table must equal the oracle's transcribed one, and the copy the HIP library
   holds (exported by vs_galerkin_near_table), and -- where the reference checkout is present -- the
   matrices written in nearby_correct.m itself.
2. Every near-box face-pair integral the oracle assembles (DIRECTFN values at Np = 20 substituted for the
   touching pairs, the 20-point cubature kept for the others) is compared with a converged integral of
   that pair's actual geometry: the static part 1/R with the inner integral over the source square in
   closed form and the outer one by a geometrically graded Gauss rule, the bounded part
   (exp(-ikR) - 1)/R by tensor Gauss.  DIRECTFN at Np = 20 is accurate to a few 1e-5 (the static self
   term against its closed form), so the bar is 1e-4 relative per entry.
"""
import json
import math
import os
import re

import numpy as np
import pytest

from oracle import galerkin as OG

HERE = os.path.dirname(os.path.abspath(__file__))
REF_NEARBY = "/root/reference/src/src_vie/src_operators/nearby_correct.m"
CELLS = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1)]


def _square(face, offset):
    c, a, b = OG.FACES[face]
    c = c + np.asarray(offset, dtype=float)
    lo, hi = np.minimum(c, c + a + b), np.maximum(c, c + a + b)
    normal = int(np.flatnonzero((a + b) == 0)[0])
    return lo, hi, normal


def derived_table(m0):
    code = np.zeros((6, 6), dtype=int)
    keep = np.zeros((6, 6), dtype=int)
    for fo in range(6):
        for fs in range(6):
            lo1, hi1, n1 = _square(fo, m0)
            lo2, hi2, n2 = _square(fs, (0, 0, 0))
            lo, hi = np.maximum(lo1, lo2), np.minimum(hi1, hi2)
            if np.any(lo > hi):
                keep[fo, fs] = 1
                continue
            dim = int(np.sum(hi > lo))
            coplanar = n1 == n2
            code[fo, fs] = {2: 1, 1: 2 if coplanar else 3, 0: 4 if coplanar else 5}[dim]
    return code, keep


@pytest.mark.parametrize("m0", CELLS)
def test_near_table_derived_from_face_geometry(m0):
    code, keep = derived_table(m0)
    tc, tk = (np.array(a) for a in OG._NEAR[m0])
    assert np.array_equal(code, tc), (m0, code, tc)
    assert np.array_equal(keep, tk), (m0, keep, tk)


def test_library_near_table_matches_geometry():
    """The HIP library's copy of the table (host data, no device needed)."""
    import ctypes as C
    from voxscatter_amd import capi
    try:
        lib = capi.load()
    except Exception as e:  # library not built: nothing to compare
        pytest.skip(f"library not loadable: {e}")
    code = (C.c_int * 288)()
    keep = (C.c_int * 288)()
    assert lib.vs_galerkin_near_table(code, keep) == 0
    for i, (mx, my, mz) in enumerate([(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]):
        dc, dk = derived_table((mx, my, mz))
        assert np.array_equal(np.array(code[36 * i:36 * i + 36]).reshape(6, 6), dc)
        assert np.array_equal(np.array(keep[36 * i:36 * i + 36]).reshape(6, 6), dk)


@pytest.mark.skipif(not os.path.exists(REF_NEARBY), reason="reference checkout absent")
def test_near_table_matches_reference_text():
    """Parse the C and P matrices of nearby_correct.m (1-based m) and compare with the derived table."""
    txt = open(REF_NEARBY).read()
    names = {"0.0": 0, "ST": 1, "EAc": 2, "EAo": 3, "VAc": 4, "VAo": 5}
    blocks = re.findall(r"isequal\(m,\[(\d),(\d),(\d)\]\).*?C = \[(.*?)\];.*?P = \[(.*?)\];", txt, re.S)
    assert len(blocks) == 8
    for mx, my, mz, ctext, ptext in blocks:
        m0 = (int(mx) - 1, int(my) - 1, int(mz) - 1)
        code = np.array([[names[t] for t in row.split()] for row in ctext.strip().splitlines() if row.strip()])
        keep = np.array([[int(float(t)) for t in row.split()] for row in ptext.strip().splitlines() if row.strip()])
        dc, dk = derived_table(m0)
        assert np.array_equal(code, dc), m0
        assert np.array_equal(keep, dk), m0


# ------------------------------------------------------------------------------------------------
# converged face-pair integrals
# ------------------------------------------------------------------------------------------------
def _graded_rule(levels=10, ratio=0.2, npts=8):
    """Gauss rule on [0, 1] graded geometrically towards both ends."""
    x, w = np.polynomial.legendre.leggauss(npts)
    edges = [0.0] + [0.5 * ratio ** k for k in range(levels, 0, -1)] + [0.5]
    edges = np.array(edges + [1 - e for e in edges[-2::-1]])
    a, b = edges[:-1, None], edges[1:, None]
    return ((b - a) / 2 * x + (a + b) / 2).ravel(), ((b - a) / 2 * w).ravel()


def _rect_potential(u, v, w, x1, x2, y1, y2):
    """int_{x1}^{x2} int_{y1}^{y2} dx dy / sqrt((x-u)^2 + (y-v)^2 + w^2), in closed form."""
    w = np.abs(w)

    def F(X, Y):
        rx = np.sqrt(X * X + w * w)
        ry = np.sqrt(Y * Y + w * w)
        R = np.sqrt(X * X + Y * Y + w * w)
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = np.where(rx > 0, X * np.arcsinh(Y / np.where(rx > 0, rx, 1)), 0.0)
            t2 = np.where(ry > 0, Y * np.arcsinh(X / np.where(ry > 0, ry, 1)), 0.0)
            t3 = np.where((w > 0) & (R > 0), w * np.arctan(X * Y / np.where((w > 0) & (R > 0), w * R, 1)), 0.0)
        return t1 + t2 - t3

    return F(x2 - u, y2 - v) - F(x1 - u, y2 - v) - F(x2 - u, y1 - v) + F(x1 - u, y1 - v)


def face_pair_integral(fo, m0, fs, kdx):
    """int over face fo of the cell at m0, int over face fs of the cell at 0, of exp(-i k R)/R, unit cells
    (multiply by dx^3 for cells of side dx with k = kdx/dx)."""
    co, ao, bo = OG.FACES[fo]
    co = co + np.asarray(m0, dtype=float)
    cs, as_, bs = OG.FACES[fs]
    # static part: inner integral in closed form, outer graded Gauss
    s, ws = _graded_rule()
    S, T = np.meshgrid(s, s, indexing="ij")
    W = np.outer(ws, ws)
    p = co + S[..., None] * ao + T[..., None] * bo
    lo, hi, nrm = _square(fs, (0, 0, 0))
    i, j = [a for a in range(3) if a != nrm]
    stat = np.sum(W * _rect_potential(p[..., i], p[..., j], p[..., nrm] - lo[nrm], lo[i], hi[i], lo[j], hi[j]))
    # bounded part (exp(-ikR) - 1)/R = -ik - k^2 R/2 + ...: tensor Gauss
    x, wg = np.polynomial.legendre.leggauss(24)
    x, wg = (x + 1) / 2, wg / 2
    U, V = np.meshgrid(x, x, indexing="ij")
    Wg = np.outer(wg, wg).ravel()
    po = (co + U[..., None] * ao + V[..., None] * bo).reshape(-1, 3)
    ps = (cs + U[..., None] * as_ + V[..., None] * bs).reshape(-1, 3)
    R = np.linalg.norm(po[:, None, :] - ps[None, :, :], axis=2)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.where(R > 0, np.expm1(-1j * kdx * R) / np.where(R > 0, R, 1), -1j * kdx)
    dyn = Wg @ h @ Wg
    return stat + dyn


def test_face_pair_integrator_closed_forms():
    """The integrator itself: the static self term of a unit square, 4/3 (1 - sqrt 2) + 4 ln(1 + sqrt 2), and
    a far pair against the point-source limit."""
    exact = 4 / 3 * (1 - math.sqrt(2)) + 4 * math.log(1 + math.sqrt(2))
    st = face_pair_integral(0, (1, 0, 0), 1, 0.0)   # -x face of the cell at +x is the +x face of cell 0
    assert abs(st - exact) <= 1e-9 * exact
    far = face_pair_integral(4, (40, 0, 0), 4, 1e-9)
    assert abs(far - 1 / 40.0) <= 1e-3 / 40


@pytest.mark.parametrize("case", [0, 1])
def test_near_box_entries_match_converged_integrals(case):
    fx = json.load(open(os.path.join(HERE, "golden", "galerkin_singular.json")))[case]
    ko, dx = fx["ko"], fx["dx"]
    sing = np.array([complex(*v) for v in fx["directfn_20"]])
    worst = 0.0
    for m0 in CELLS:
        J = OG.nearby_correct(m0, OG.cubature_nop(20, m0, ko, dx), sing)
        code = np.array(OG._NEAR[m0][0])
        for fo in range(6):
            for fs in range(6):
                ref = face_pair_integral(fo, m0, fs, ko * dx) * dx ** 3
                err = abs(J[fo, fs] - ref) / abs(ref)
                worst = max(worst, err)
                assert err <= 1e-4, (m0, fo, fs, code[fo, fs], J[fo, fs], ref)
    print(f"[near box] ko={ko} dx={dx}: worst relative entry error {worst:.2e}")
