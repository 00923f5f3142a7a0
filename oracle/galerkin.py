"""Oracle (test infrastructure only): Galerkin N-operator Toeplitz tensor, restated in NumPy.

Follows assembly_nop_const.m:1-174 (far 4-point, medium 10-point, near 20-point Gauss rules per face
dimension, the near 2x2x2 box corrected by nearby_correct.m:1-145), cube_faces.m:1-39,
vector_projection.m:1-25, gauss_2d.m:1-19, mexcubatures/cubature_nop_const.cpp:8-124 and
cubatures/cubature_nop_sym_near.m:1-88; the total field follows E_field_Nop_const.m:62-84 and
mv_L_const.m:19-85.  The five singular face-pair integrals are an input (the
reference obtains them from DIRECTFN through singular_*_const.m); ``reference_singular`` evaluates them
with the reference's own C++ when oracle/_ref has been built (oracle/build_ref.sh).

Parity: the face-face cubature is pinned against the reference's compiled cubature_nop_const.cpp; the
assembly (box selection, symmetry fill, nearby correction, projection) is pinned by the GPU/oracle
comparison only through this restatement and is otherwise parity unpinned against MATLAB.  The total
field is pinned by the constitutive relation it must satisfy inside the scatterer for an exact solution
of the JVIE, E = J / (j omega eps0 (eps - 1)).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ETA0 = 3.767303134617706e2  # free-space impedance: omega eps0 = ko / ETA0

# faces of the unit cube (cube_faces.m): corner, edge a (vertex 2 - vertex 1), edge b (vertex 4 - vertex 1)
FACES = np.array([
    [[0, 0, 0], [0, 1, 0], [0, 0, 1]],   # -x
    [[1, 0, 0], [0, 1, 0], [0, 0, 1]],   # +x
    [[0, 0, 0], [1, 0, 0], [0, 0, 1]],   # -y
    [[0, 1, 0], [1, 0, 0], [0, 0, 1]],   # +y
    [[0, 0, 0], [1, 0, 0], [0, 1, 0]],   # -z
    [[0, 0, 1], [1, 0, 0], [0, 1, 0]],   # +z
], dtype=float)

# the 21 computed face pairs (cubature_nop_const.cpp:28-32), 0-based
PAIR_OBS = [0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5]
PAIR_SRC = [0, 1, 2, 3, 4, 5, 0, 2, 3, 4, 5, 2, 3, 4, 5, 2, 4, 5, 4, 5, 4]


def gauss_2d(n):
    """gauss_2d.m: tensor Gauss-Legendre rule on [0,1]^2, weights summing to 1 (ordering I outer, J inner)."""
    x, w = np.polynomial.legendre.leggauss(n)
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    I, J = I.ravel(), J.ravel()
    return 0.25 * w[I] * w[J], (x[I] + 1) / 2, (x[J] + 1) / 2


def vector_projection():
    """vector_projection.m: P[k, l, c] = (n_k x e_a) . (n_l x e_b), c = xx, xy, xz, yy, yz, zz."""
    e = np.eye(3)
    KL = np.array([-e[0], e[0], -e[1], e[1], -e[2], e[2]])
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    P = np.zeros((6, 6, 6))
    for k in range(6):
        for l in range(6):
            for c, (a, b) in enumerate(comps):
                P[k, l, c] = np.dot(np.cross(KL[k], e[a]), np.cross(KL[l], e[b]))
    return P


def symmetric_fill(I):
    """The 15 face pairs obtained from the 21 computed ones (cubature_nop_const.cpp:104-120)."""
    I[1, 1] = I[0, 0]
    I[2, 0] = I[1, 3]; I[2, 1] = I[0, 3]
    I[3, 0] = I[1, 2]; I[3, 1] = I[0, 2]; I[3, 3] = I[2, 2]
    I[4, 0] = I[1, 5]; I[4, 1] = I[0, 5]; I[4, 2] = I[3, 5]; I[4, 3] = I[2, 5]
    I[5, 0] = I[1, 4]; I[5, 1] = I[0, 4]; I[5, 2] = I[3, 4]; I[5, 3] = I[2, 4]; I[5, 5] = I[4, 4]
    return I


def cubature_nop(n1d, m, ko, dx):
    """6 x 6 face-face integrals between the cell at integer offset m and cell 0 with the n1d-point rule
    (cubature_nop_const.cpp:8-124 / cubature_nop_sym_near.m), scaled by dx^4."""
    w, u, v = gauss_2d(n1d)
    I = np.zeros((6, 6), dtype=complex)
    m = np.asarray(m, dtype=float)
    for fo, fs in zip(PAIR_OBS, PAIR_SRC):
        co, ao, bo = FACES[fo] * dx
        cs, as_, bs = FACES[fs] * dx
        rp = co + np.outer(u, ao) + np.outer(v, bo) + m * dx
        rq = cs + np.outer(u, as_) + np.outer(v, bs)
        R = np.linalg.norm(rq[None, :, :] - rp[:, None, :], axis=2)
        I[fo, fs] = np.sum(np.outer(w, w) * np.exp(-1j * ko * R) / R)
    return dx ** 4 * symmetric_fill(I)


# nearby_correct.m:1-145 -- codes 0 zero, 1 ST, 2 EAc, 3 EAo, 4 VAc, 5 VAo; KEEP 1 = keep the cubature
_NEAR = {
    (0, 0, 0): ([[1, 0, 3, 3, 3, 3], [0, 1, 3, 3, 3, 3], [3, 3, 1, 0, 3, 3], [3, 3, 0, 1, 3, 3],
                 [3, 3, 3, 3, 1, 0], [3, 3, 3, 3, 0, 1]],
                [[0, 1, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0], [0, 0, 0, 1, 0, 0], [0, 0, 1, 0, 0, 0],
                 [0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 1, 0]]),
    (1, 0, 0): ([[0, 1, 3, 3, 3, 3], [0] * 6, [0, 3, 2, 0, 5, 5], [0, 3, 0, 2, 5, 5], [0, 3, 5, 5, 2, 0],
                 [0, 3, 5, 5, 0, 2]],
                [[1, 0, 0, 0, 0, 0], [1] * 6, [1, 0, 0, 1, 0, 0], [1, 0, 1, 0, 0, 0], [1, 0, 0, 0, 0, 1],
                 [1, 0, 0, 0, 1, 0]]),
    (1, 0, 1): ([[0, 2, 5, 5, 0, 3], [0] * 6, [0, 5, 4, 0, 0, 5], [0, 5, 0, 4, 0, 5], [0, 3, 5, 5, 0, 2],
                 [0] * 6],
                [[1, 0, 0, 0, 1, 0], [1] * 6, [1, 0, 0, 1, 1, 0], [1, 0, 1, 0, 1, 0], [1, 0, 0, 0, 1, 0],
                 [1] * 6]),
    (0, 0, 1): ([[2, 0, 5, 5, 0, 3], [0, 2, 5, 5, 0, 3], [5, 5, 2, 0, 0, 3], [5, 5, 0, 2, 0, 3],
                 [3, 3, 3, 3, 0, 1], [0] * 6],
                [[0, 1, 0, 0, 1, 0], [1, 0, 0, 0, 1, 0], [0, 0, 0, 1, 1, 0], [0, 0, 1, 0, 1, 0],
                 [0, 0, 0, 0, 1, 0], [1] * 6]),
    (0, 1, 0): ([[2, 0, 0, 3, 5, 5], [0, 2, 0, 3, 5, 5], [3, 3, 0, 1, 3, 3], [0] * 6, [5, 5, 0, 3, 2, 0],
                 [5, 5, 0, 3, 0, 2]],
                [[0, 1, 1, 0, 0, 0], [1, 0, 1, 0, 0, 0], [0, 0, 1, 0, 0, 0], [1] * 6, [0, 0, 1, 0, 0, 1],
                 [0, 0, 1, 0, 1, 0]]),
    (1, 1, 0): ([[0, 2, 0, 3, 5, 5], [0] * 6, [0, 3, 0, 2, 5, 5], [0] * 6, [0, 5, 0, 5, 4, 0],
                 [0, 5, 0, 5, 0, 4]],
                [[1, 0, 1, 0, 0, 0], [1] * 6, [1, 0, 1, 0, 0, 0], [1] * 6, [1, 0, 1, 0, 0, 1],
                 [1, 0, 1, 0, 1, 0]]),
    (0, 1, 1): ([[4, 0, 0, 5, 0, 5], [0, 4, 0, 5, 0, 5], [5, 5, 0, 2, 0, 3], [0] * 6, [5, 5, 0, 3, 0, 2],
                 [0] * 6],
                [[0, 1, 1, 0, 1, 0], [1, 0, 1, 0, 1, 0], [0, 0, 1, 0, 1, 0], [1] * 6, [0, 0, 1, 0, 1, 0],
                 [1] * 6]),
    (1, 1, 1): ([[0, 4, 0, 5, 0, 5], [0] * 6, [0, 5, 0, 4, 0, 5], [0] * 6, [0, 5, 0, 5, 0, 4], [0] * 6],
                [[1, 0, 1, 0, 1, 0], [1] * 6, [1, 0, 1, 0, 1, 0], [1] * 6, [1, 0, 1, 0, 1, 0], [1] * 6]),
}


def nearby_correct(m0, I, sing):
    """nearby_correct.m for the 0-based near cell m0 (each index 0 or 1); sing = [ST, EAc, EAo, VAc, VAo]."""
    code, keep = (np.array(a) for a in _NEAR[tuple(int(v) for v in m0)])
    vals = np.concatenate([[0.0], np.asarray(sing, dtype=complex)])
    return np.where(keep == 1, I, vals[code])


def assembly_nop_const(L, M, N, dx, ko, sing):
    """Toeplitz tensor G (L x M x N x 6) of the N operator (assembly_nop_const.m), given the singular
    integrals sing = [ST, EAc, EAo, VAc, VAo]."""
    J = np.zeros((L, M, N, 6, 6), dtype=complex)
    nmL, nmM, nmN = min(5, L), min(5, M), min(5, N)
    for mx in range(L):
        for my in range(M):
            for mz in range(N):
                if mx < 2 and my < 2 and mz < 2:
                    J[mx, my, mz] = nearby_correct((mx, my, mz), cubature_nop(20, (mx, my, mz), ko, dx), sing)
                elif mx < nmL and my < nmM and mz < nmN:
                    J[mx, my, mz] = cubature_nop(10, (mx, my, mz), ko, dx)
                else:
                    J[mx, my, mz] = cubature_nop(4, (mx, my, mz), ko, dx)
    P = vector_projection()
    G = np.einsum("xyzij,ijk->xyzk", J, P)
    return G / (4 * np.pi)


# ------------------------------------------------------------------------------------------------
# the reference's own C++ (oracle/_ref, built by oracle/build_ref.sh) -- this container only
# ------------------------------------------------------------------------------------------------
_REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")


def ref_available() -> bool:
    return all(os.path.exists(os.path.join(_REF, f"libvsref_{t}.so")) for t in ("st", "ea", "va", "cub"))


def _lib(tag):
    return C.CDLL(os.path.join(_REF, f"libvsref_{tag}.so"))


def reference_singular(ko, dx, n=20):
    """[ST, EAc, EAo, VAc, VAo] from the reference's DIRECTFN, combined exactly as singular_st_const.m,
    singular_ea_const.m and singular_va_const.m do (two triangles per square, Np_1D = n)."""
    st, ea, va = _lib("st"), _lib("ea"), _lib("va")
    D3 = C.c_double * 3

    def v3(a):
        return D3(*[float(x) for x in a])

    def ST(r1, r2, r3):
        o = (C.c_double * 2)()
        st.vsref_directfn_st(v3(r1), v3(r2), v3(r3), C.c_double(ko), n, o)
        return complex(o[0], o[1])

    def EA(r1, r2, r3, r4):  # mexdirectfn_ws_ea_const gateway: P = (r1, r2, r3), Q = (r4, r2, r1)
        o = (C.c_double * 2)()
        ea.vsref_directfn_ea(v3(r1), v3(r2), v3(r3), v3(r4), v3(r2), v3(r1), C.c_double(ko), n, o)
        return complex(o[0], o[1])

    def VA(r1, r2, r3, r4, r5):  # mexdirectfn_ws_va_const gateway: P = (r1, r2, r3), Q = (r4, r5, r1)
        o = (C.c_double * 2)()
        va.vsref_directfn_va(v3(r1), v3(r2), v3(r3), v3(r4), v3(r5), v3(r1), C.c_double(ko), n, o)
        return complex(o[0], o[1])

    d = dx
    r1, r2, r3, r4 = [0, 0, 0], [d, 0, 0], [d, d, 0], [0, d, 0]
    I_ST = ST(r1, r2, r3) + EA(r1, r3, r2, r4) + EA(r1, r3, r4, r2) + ST(r1, r3, r4)
    r5, r6, r5o, r6o = [2 * d, 0, 0], [2 * d, d, 0], [d, 0, d], [d, d, d]
    EAc = EA(r2, r3, r1, r5) + VA(r3, r1, r2, r5, r6) + VA(r3, r4, r1, r2, r5) + VA(r3, r4, r1, r5, r6)
    EAo = EA(r2, r3, r1, r5o) + VA(r3, r1, r2, r5o, r6o) + VA(r3, r4, r1, r2, r5o) + VA(r3, r4, r1, r5o, r6o)
    r5, r6, r7 = [2 * d, d, 0], [2 * d, 2 * d, 0], [d, 2 * d, 0]
    r5o, r6o, r7o = [d, d, d], [d, 2 * d, d], [d, 2 * d, 0]
    VAc = VA(r3, r1, r2, r5, r6) + VA(r3, r1, r2, r6, r7) + VA(r3, r4, r1, r5, r6) + VA(r3, r4, r1, r6, r7)
    VAo = VA(r3, r1, r2, r5o, r6o) + VA(r3, r1, r2, r6o, r7o) + VA(r3, r4, r1, r5o, r6o) + VA(r3, r4, r1, r6o, r7o)
    return [I_ST, EAc, EAo, VAc, VAo]


def reference_cubature(n1d, m, ko, dx):
    """The reference's cubature_nop_const.cpp on the cell at integer offset m (r_n = 0), 6 x 6."""
    lib = _lib("cub")
    w, u, v = gauss_2d(n1d)
    Rf = np.zeros((3, 4, 6))
    for f in range(6):
        c, a, b = FACES[f] * dx
        Rf[:, 0, f], Rf[:, 1, f], Rf[:, 3, f], Rf[:, 2, f] = c, c + a, c + b, c + a + b
    Rf = np.ascontiguousarray(Rf)
    D = C.POINTER(C.c_double)
    rm = np.asarray(m, dtype=float) * dx
    rn = np.zeros(3)
    out = np.zeros(72)
    arr = [np.ascontiguousarray(a, dtype=float) for a in (w, u, v, rm, rn)]
    lib.vsref_cubature_nop(n1d * n1d, *[a.ctypes.data_as(D) for a in arr], Rf.ctypes.data_as(D),
                           C.c_double(ko), C.c_double(dx), out.ctypes.data_as(D))
    return (out[0::2] + 1j * out[1::2]).reshape(6, 6)


def plane_wave_excitation(r, kvec, Eo):
    """PlaneWave_Excitation.m:39-99: voxel average of Eo exp(-i k.r) over each cube (L x M x N x 3)."""
    L, M, N, _ = r.shape
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    fac = np.ones((L, M, N), dtype=complex)
    for a in range(3):
        ka = kvec[a]
        if ka != 0:
            fac = fac * np.exp(-1j * ka * r[..., a]) * (np.exp(-1j * ka * dx / 2) - np.exp(1j * ka * dx / 2)) / (-1j * ka)
        else:
            fac = fac * dx
    E = np.stack([Eo[i] * fac for i in range(3)], axis=-1)
    return E / dx ** 3


# ------------------------------------------------------------------------------------------------
# total field (E_field_Nop_const.m)
# ------------------------------------------------------------------------------------------------
_COMP = ((0, 1, 2), (1, 3, 4), (2, 4, 5))  # symmetric 3 x 3 of the six tensor components


def mv_L_const(J, fN, Gram):
    """mv_L_const.m:19-85 over the whole box: N J - Gram J for J (L x M x N x 3), fN the FFT of the
    circulant embedding (2L x 2M x 2N x 6)."""
    L, M, N, _ = J.shape
    shp = fN.shape[:3]
    fJ = [np.fft.fftn(J[..., c], s=shp) for c in range(3)]
    out = np.empty(J.shape, dtype=complex)
    for a in range(3):
        acc = fN[..., _COMP[a][0]] * fJ[0] + fN[..., _COMP[a][1]] * fJ[1] + fN[..., _COMP[a][2]] * fJ[2]
        out[..., a] = np.fft.ifftn(acc)[:L, :M, :N] - Gram * J[..., a]
    return out


def E_field_Nop_const(J, fN, Gram, ko, Einc):
    """E_field_Nop_const.m:62-84: Einc + (N J - Gram J) / (Gram j omega eps0), omega eps0 = ko / ETA0."""
    J = np.asarray(J, dtype=complex)
    if not np.any(J):
        return np.asarray(Einc, dtype=complex)
    return Einc + mv_L_const(J, fN, Gram) / (Gram * 1j * ko / ETA0)
