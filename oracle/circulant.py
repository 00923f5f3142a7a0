"""Chan circulant preconditioners, FP64 (TEST INFRASTRUCTURE: see oracle/__init__.py).

Follows src/src_preconditioning/circ_1_level.m, circ_2_level.m, mvp_circ_2_level.m, chan_mvp_idx.m
and the block inversions of main_DDA.m:105-111 / scatteringHumanHead.m:155-184.
"""
from __future__ import annotations

import numpy as np

from . import _fft

# components whose Chan average negates the first term: x-direction (circ_1_level.m:29,42),
# y-direction (circ_2_level.m:27,63)
ODD_X = (1, 2)
ODD_Y = (1, 4)


def toeplitz(col, row=None):
    """MATLAB ``toeplitz(c, r)`` (column wins on the diagonal) along the last axis of ``col``/``row``."""
    col = np.asarray(col)
    row = col if row is None else np.asarray(row)
    n = col.shape[-1]
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    d = i - j
    out_lo = np.take(col, np.abs(d), axis=-1)
    out_hi = np.take(row, np.abs(d), axis=-1)
    return np.where(d >= 0, out_lo, out_hi)


def _chan_axis0(A, odd):
    """c(1) = A(1); c(i) = +-(n+1-i)/n A(i) + (i-1)/n A(n-i+2) along axis 0 (1-based i)."""
    n = A.shape[0]
    c = np.zeros_like(A)
    c[0] = A[0]
    s = -1.0 if odd else 1.0
    for i in range(1, n):
        c[i] = s * (n - i) / n * A[i] + i / n * A[n - i]
    return c


def circ_L_opToep(opToeplitz):
    """Chan circulant along x + FFT along x, per component (circ_1_level.m:5-90)."""
    out = np.zeros_like(opToeplitz, dtype=np.complex128)
    for p in range(6):
        out[..., p] = _fft.fft(_chan_axis0(opToeplitz[..., p].astype(np.complex128), p in ODD_X), axis=0)
    return out


def circ_1_level(opToeplitz, L, M, N, parallel_prec=0, on_off="off"):
    """``[circ, circ_L_opToep] = circ_1_level(...)`` (circ_1_level.m:1-259).

    With ``on_off='on'`` ``circ`` is an (L, 3MN, 3MN) array of the 1-level diagonal blocks; unknowns
    inside a block are ordered y fastest, then z, then component (reshape of the L x 3MN matrix).
    """
    cL = circ_L_opToep(opToeplitz)
    if on_off == "off":
        return None, cL
    MN = M * N
    blocks = np.zeros((L, 3 * MN, 3 * MN), dtype=np.complex128)
    zt = toeplitz(np.arange(N))  # block (a,b) uses chan{|a-b|+1}
    for i in range(L):
        def build(arr, neg_col):
            chan = []
            for k in range(N):
                c = arr[i, :, k]
                col = np.concatenate([c[:1], -c[1:]]) if neg_col else c
                chan.append(toeplitz(col, c))
            res = np.zeros((MN, MN), dtype=np.complex128)
            for a in range(N):
                for b in range(N):
                    res[a * M:(a + 1) * M, b * M:(b + 1) * M] = chan[zt[a, b]]
            return res
        xx = build(cL[..., 0], False)
        xy = build(cL[..., 1], True)
        xz = build(cL[..., 2], False)
        U = np.triu(xz)
        xz = U - U.T + np.diag(np.diag(U))
        yy = build(cL[..., 3], False)
        yz = build(cL[..., 4], True)
        U = np.triu(yz)
        yz = U + U.T - np.diag(np.diag(U))
        zz = build(cL[..., 5], False)
        blocks[i] = np.block([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
    return blocks, cL


def circ_2_level(circ_L, L, M, N):
    """``[circ_M_opToep, circ] = circ_2_level(circ_L_opToep, L, M, N)`` (circ_2_level.m:1-155).

    ``circ`` is an (L, M, 3N, 3N) array; unknowns inside a block are ordered z fastest, then component.
    """
    circ_M = np.zeros_like(circ_L, dtype=np.complex128)
    blocks = np.zeros((L, M, 3 * N, 3 * N), dtype=np.complex128)
    for i in range(L):
        F = [_fft.fft(_chan_axis0(circ_L[i, :, :, p], p in ODD_Y), axis=0) for p in range(6)]
        for p in range(6):
            circ_M[i, :, :, p] = -F[p] if p == 2 else F[p]   # circ_2_level.m:44 (stored, unused)
        for j in range(M):
            def t_sym(c):
                return toeplitz(c, c)

            def t_odd(c):
                return toeplitz(np.concatenate([c[:1], -c[1:]]), c)
            T1 = t_sym(F[0][j])
            T2 = toeplitz(np.concatenate([F[1][j][:1], F[1][j][1:]]), F[1][j])
            T3 = t_odd(F[2][j])
            T4 = t_sym(F[3][j])
            T5 = t_odd(F[4][j])
            T6 = t_sym(F[5][j])
            blocks[i, j] = np.block([[T1, T2, T3], [T2, T4, T5], [T3, T5, T6]])
    return circ_M, blocks


def invert_blocks_dda(blocks, alpha, dV):
    """``inv(1/alpha_LDR*eye(3N) - dV*circ{i,j})`` for every block (main_DDA.m:107-111)."""
    n = blocks.shape[-1]
    return np.linalg.inv(np.eye(n) / alpha - dV * blocks)


def averaged_contrast(epsilon_r):
    """``face3_2`` of scatteringHumanHead.m:155-176: (eps-1)/eps of eps averaged over x then y."""
    L, M, N = epsilon_r.shape
    hey = epsilon_r.sum(axis=0) / L                   # M x N
    kk = hey.sum(axis=0) / M                          # N (average of the x-averaged slab over y)
    face_2 = (kk - 1) / kk
    return np.concatenate([face_2, face_2, face_2])


def invert_blocks_galerkin(blocks, Gram, face3_2):
    """``inv(Gram*eye(3N) - diag(face3_2)*circ_2_N{i,j})`` (scatteringHumanHead.m:178-184)."""
    n = blocks.shape[-1]
    return np.linalg.inv(Gram * np.eye(n) - face3_2[:, None] * blocks)


def mvp_circ_2_level(inv_blocks, JIn, L, M, N, idx):
    """Apply the 2-level preconditioner (mvp_circ_2_level.m:2-22).  ``idx`` is 0-based."""
    V = np.zeros(3 * L * M * N, dtype=np.complex128)
    V[idx] = JIn
    temp = _fft.fft(V.reshape((L, 3 * M * N), order="F"), axis=0).T        # 3MN x L
    for i in range(L):
        TEMP = _fft.fft(temp[:, i].reshape((M, 3 * N), order="F"), axis=0).T   # 3N x M
        TEMP = np.einsum("jab,bj->aj", inv_blocks[i], TEMP)
        TEMP = _fft.ifft(TEMP.T, axis=0)                                      # M x 3N
        temp[:, i] = TEMP.reshape(-1, order="F")
    temp = _fft.ifft(temp.T, axis=0)
    return temp.reshape(-1, order="F")[idx]


def chan_mvp_idx(inv_blocks, JIn, L, M, N, idx):
    """Apply the 1-level preconditioner (chan_mvp_idx.m:2-16).  ``idx`` is 0-based."""
    V = np.zeros(3 * L * M * N, dtype=np.complex128)
    V[idx] = JIn
    temp = _fft.fft(V.reshape((L, 3 * M * N), order="F"), axis=0).T
    for i in range(L):
        temp[:, i] = inv_blocks[i] @ temp[:, i]
    temp = _fft.ifft(temp.T, axis=0)
    return temp.reshape(-1, order="F")[idx]
