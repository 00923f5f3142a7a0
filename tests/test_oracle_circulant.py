"""Pins of the oracle's Chan preconditioners (circ_1_level.m, circ_2_level.m, mvp_circ_2_level.m,
chan_mvp_idx.m): the restated construction, with all of the reference's sign manipulations, must equal
the textbook optimal-circulant (Chan 1988) projection built here independently from the dyadic at
signed offsets."""
import math

import numpy as np
import pytest

from oracle import circulant as Ch
from oracle import operator as O
from voxscatter_amd import geometry as G

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
COMP = {(a, b): p for p, (a, b) in enumerate(PAIRS)}
COMP.update({(b, a): p for p, (a, b) in enumerate(PAIRS)})


def _gen(d, k):
    """t(mx, my, mz) for signed offsets, straight from the dyadic."""
    def t(mx, my, mz):
        return O.green_dyadic(np.array([mx * d, my * d, mz * d]), k)
    return t


def _dense_chan(L, M, N, d, k, levels):
    t = _gen(d, k)
    V = L * M * N
    Cd = np.zeros((3 * V, 3 * V), dtype=complex)

    def cgen(kx, ly, mz):
        if levels == 2:
            acc = 0
            for sx, wx in ((kx, (L - kx)), (kx - L, kx)):
                for sy, wy in ((ly, (M - ly)), (ly - M, ly)):
                    if wx and wy:
                        acc = acc + wx * wy * t(sx, sy, mz)
            return acc / (L * M)
        acc = 0
        for sx, wx in ((kx, (L - kx)), (kx - L, kx)):
            if wx:
                acc = acc + wx * t(sx, ly, mz)
        return acc / L

    cache = {}
    for x in range(L):
        for y in range(M):
            for z in range(N):
                for x2 in range(L):
                    for y2 in range(M):
                        for z2 in range(N):
                            kx = (x - x2) % L
                            ly = (y - y2) % M if levels == 2 else (y - y2)
                            key = (kx, ly, z - z2)
                            if key not in cache:
                                cache[key] = cgen(*key)
                            g6 = cache[key]
                            i = x + L * (y + M * z)
                            j = x2 + L * (y2 + M * z2)
                            for a in range(3):
                                for b in range(3):
                                    Cd[a * V + i, b * V + j] = g6[COMP[(a, b)]]
    return Cd


@pytest.mark.parametrize("shape", [(4, 3, 3), (3, 4, 2), (5, 2, 3)])
def test_two_level_equals_textbook_chan(shape):
    L, M, N = shape
    d, k = 0.09, 2 * math.pi
    r = G.centred_grid(L, M, N, d)
    Toep = O.toeplitz_dda(r, k)
    _, cL = Ch.circ_1_level(Toep, L, M, N, 0, "off")
    _, blocks = Ch.circ_2_level(cL, L, M, N)
    Cd = _dense_chan(L, M, N, d, k, 2)
    V = L * M * N
    rng = np.random.default_rng(3)
    u = rng.normal(size=3 * V) + 1j * rng.normal(size=3 * V)
    idx = np.arange(3 * V)
    got = Ch.mvp_circ_2_level(blocks, u, L, M, N, idx)
    ref = Cd @ u
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    # inverse blocks invert a I - w C
    a, w = 7.0 + 1.0j, d ** 3
    inv = Ch.invert_blocks_dda(blocks, 1 / a, w)
    v = a * u - w * ref
    assert np.linalg.norm(Ch.mvp_circ_2_level(inv, v, L, M, N, idx) - u) <= 1e-12 * np.linalg.norm(u)


@pytest.mark.parametrize("shape", [(4, 3, 2), (3, 2, 3)])
def test_one_level_equals_textbook_chan(shape):
    L, M, N = shape
    d, k = 0.09, 2 * math.pi
    r = G.centred_grid(L, M, N, d)
    Toep = O.toeplitz_dda(r, k)
    blocks, _ = Ch.circ_1_level(Toep, L, M, N, 0, "on")
    Cd = _dense_chan(L, M, N, d, k, 1)
    V = L * M * N
    rng = np.random.default_rng(4)
    u = rng.normal(size=3 * V) + 1j * rng.normal(size=3 * V)
    idx = np.arange(3 * V)
    got = Ch.chan_mvp_idx(blocks, u, L, M, N, idx)
    ref = Cd @ u
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_chan_of_circulant_generator_is_exact():
    """Chan's projection is idempotent: a generator that is already L-periodic comes back unchanged."""
    L = 6
    rng = np.random.default_rng(5)
    c = rng.normal(size=L) + 1j * rng.normal(size=L)
    # symmetric-in-x Toeplitz data whose wrapped extension is circulant: t_k = c_k, t_{k-L} = c_k
    A = np.zeros((L, 1, 1, 6), dtype=complex)
    A[:, 0, 0, 0] = c
    # the Chan average of t_k (k>=0) and t_{-(L-k)}: for an even component t_{-(L-k)} = t_{L-k}
    cL = Ch.circ_L_opToep(A)
    chan = np.array([c[0]] + [((L - i) * c[i] + i * c[L - i]) / L for i in range(1, L)])
    assert np.allclose(cL[:, 0, 0, 0], np.fft.fft(chan))
    # if the data is symmetric circulant (c_k = c_{L-k}), the projection returns it unchanged
    cs = c.copy()
    for i in range(1, L):
        cs[L - i] = cs[i]
    A[:, 0, 0, 0] = cs
    assert np.allclose(Ch.circ_L_opToep(A)[:, 0, 0, 0], np.fft.fft(cs))


def test_averaged_contrast_recipe():
    rng = np.random.default_rng(6)
    eps = 1 + rng.uniform(0, 3, size=(3, 4, 5)) + 0.01j
    f = Ch.averaged_contrast(eps)
    mean = eps.mean(axis=(0, 1))
    assert np.allclose(f, np.concatenate([(mean - 1) / mean] * 3))
