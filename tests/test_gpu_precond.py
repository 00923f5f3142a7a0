"""GPU parity of the Chan preconditioners (circ_1_level, circ_2_level, block inversion, apply)."""
import math

import numpy as np
import pytest

from oracle import circulant as Ch
from oracle import operator as O
from voxscatter_amd import api, geometry as G

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.mark.parametrize("shape", [(4, 3, 2), (8, 6, 5), (5, 7, 3)])
def test_circ_levels_match_oracle(ctx, shape):
    print(f"[circ] shape={shape}", flush=True)
    L, M, N = shape
    k = 2 * math.pi
    r = G.centred_grid(L, M, N, 0.08)
    Toep = O.toeplitz_dda(r, k)
    b1, cL = api.circ_1_level(Toep, L, M, N, 0, "on")
    rb1, rcL = Ch.circ_1_level(Toep, L, M, N, 0, "on")
    assert _rel(cL, rcL) <= 1e-13
    assert _rel(b1, rb1) <= 1e-13
    cM, b2 = api.circ_2_level(cL, L, M, N)
    rcM, rb2 = Ch.circ_2_level(rcL, L, M, N)
    assert _rel(cM, rcM) <= 1e-13
    assert _rel(b2, rb2) <= 1e-13


@pytest.mark.parametrize("n,nblk", [(24, 50), (48, 20), (72, 3), (96, 2), (150, 2), (288, 3), (300, 1), (330, 1)])
def test_block_inverse_matches_numpy(ctx, n, nblk):
    """Whole blocks in LDS up to 3N = 96, the blocked inversion (FP64 matrix-core updates) above: 32-column
    panels up to 300 (a partial last panel at 150 and 300), 16-column panels at 330."""
    print(f"[inverse] n={n} nblk={nblk}", flush=True)
    rng = np.random.default_rng(n)
    B = (rng.normal(size=(nblk, n, n)) + 1j * rng.normal(size=(nblk, n, n))) / np.sqrt(n)
    a, w = 3.0 + 0.5j, 0.2
    inv = api.invert_blocks(B, a, w)
    ref = np.linalg.inv(a * np.eye(n) - w * B)
    assert _rel(inv, ref) <= 1e-11


@pytest.mark.parametrize("n", [150, 288, 330])
def test_blocked_inverse_pivoting(ctx, n, monkeypatch):
    """Blocks whose pivots lie far from the diagonal (a scaled random permutation plus a small random part,
    so every step swaps rows, also across panels): the blocked inversion against numpy and against the
    unblocked sweep, which picks the same pivots."""
    rng = np.random.default_rng(n + 1)
    nblk = 3
    B = np.zeros((nblk, n, n), complex)
    for b in range(nblk):
        perm = rng.permutation(n)
        B[b, np.arange(n), perm] = rng.uniform(1, 2, n) * np.exp(2j * np.pi * rng.uniform(size=n))
    B += 1e-3 * (rng.normal(size=B.shape) + 1j * rng.normal(size=B.shape))
    ref = np.linalg.inv(B)
    inv = api.invert_blocks(B, 0.0, -1.0)
    monkeypatch.setenv("VS_GJ_BLOCKED", "0")
    inv0 = api.invert_blocks(B, 0.0, -1.0)
    e_ref, e_unb = _rel(inv, ref), _rel(inv, inv0)
    print(f"[inverse pivoting] n={n}: vs numpy {e_ref:.2e}, vs unblocked {e_unb:.2e}", flush=True)
    assert e_ref <= 1e-12 and e_unb <= 1e-13


def test_block_inverse_vector_weights(ctx):
    rng = np.random.default_rng(9)
    n = 30
    B = rng.normal(size=(4, n, n)) + 1j * rng.normal(size=(4, n, n))
    w = rng.uniform(0.1, 0.5, size=n) + 0.01j
    inv = api.invert_blocks(B, 2.0, w)
    ref = np.linalg.inv(2.0 * np.eye(n) - w[:, None] * B)
    assert _rel(inv, ref) <= 1e-11


@pytest.mark.parametrize("levels", [1, 2])
def test_apply_matches_oracle(ctx, levels):
    L, M, N = 10, 8, 4
    k = 2 * math.pi
    d = 0.07
    r = G.centred_grid(L, M, N, d)
    rng = np.random.default_rng(11)
    nD = L * M * N
    vox = np.sort(rng.choice(nD, size=200, replace=False))
    idx3 = np.concatenate([vox, nD + vox, 2 * nD + vox])
    Toep = O.toeplitz_dda(r, k)
    alpha = O.alpha_ldr(1.31, k, d, [0, 0, 1], [1, 0, 0])
    x = rng.normal(size=idx3.size) + 1j * rng.normal(size=idx3.size)
    if levels == 2:
        _, cL = Ch.circ_1_level(Toep, L, M, N)
        _, blk = Ch.circ_2_level(cL, L, M, N)
        inv = Ch.invert_blocks_dda(blk, alpha, d ** 3)
        y = api.mvp_circ_2_level(inv, x, L, M, N, idx3)
        yref = Ch.mvp_circ_2_level(inv, x, L, M, N, idx3)
    else:
        blk, _ = Ch.circ_1_level(Toep, L, M, N, 0, "on")
        inv = Ch.invert_blocks_dda(blk, alpha, d ** 3)
        y = api.chan_mvp_idx(inv, x, L, M, N, idx3)
        yref = Ch.chan_mvp_idx(inv, x, L, M, N, idx3)
    assert _rel(y, yref) <= 1e-12
    # the device-built preconditioner of the same operator gives the same apply
    op = api.Operator.dda(r, k, ctx=ctx)
    op.set_domain(idx3, 1 / alpha, 1.0, d ** 3)
    pc = api.Preconditioner.from_operator(op, levels, 1 / alpha, wscalar=d ** 3)
    assert _rel(pc.apply(x), yref) <= 1e-11


@pytest.mark.parametrize("shape", [(10, 8, 4), (9, 7, 3), (6, 5, 2), (2, 3, 2), (4, 3, 30), (6, 5, 36)])
@pytest.mark.parametrize("weights", ["scalar", "vector"])
def test_mirrored_two_level_apply(ctx, shape, weights):
    """The device 2-level preconditioner keeps only the blocks kx <= L/2, ky <= M/2 and applies the others
    through the sign symmetry; its apply equals the oracle's apply with every block inverted, for even and
    odd grids, scalar and per-row weights (packed and full storage), one and five right-hand sides; the
    3N = 90 and 108 blocks take more than 64 KB of LDS per workgroup for four right-hand sides at once."""
    print(f"[mirror] shape={shape} weights={weights}", flush=True)
    L, M, N = shape
    k, d = 2 * math.pi, 0.07
    r = G.centred_grid(L, M, N, d)
    rng = np.random.default_rng(L * 100 + M)
    nD = L * M * N
    vox = np.sort(rng.choice(nD, size=max(1, nD * 2 // 3), replace=False))
    idx3 = np.concatenate([vox, nD + vox, 2 * nD + vox])
    Toep = O.toeplitz_dda(r, k)
    alpha = O.alpha_ldr(1.31 + 0.01j, k, d, [0, 0, 1], [1, 0, 0])
    n = 3 * N
    _, cL = Ch.circ_1_level(Toep, L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    if weights == "scalar":
        w = None
        inv = Ch.invert_blocks_dda(blk, alpha, d ** 3)
    else:
        w = d ** 3 * (1 + 0.1 * rng.uniform(size=n)) + 0.001j
        inv = np.linalg.inv(np.eye(n) / alpha - w[:, None] * blk)
    op = api.Operator.dda(r, k, ctx=ctx)
    op.set_domain(idx3, 1 / alpha, 1.0, d ** 3)
    pc = api.Preconditioner.from_operator(op, 2, 1 / alpha, w=w, wscalar=None if w is not None else d ** 3)
    X = rng.normal(size=(5, idx3.size)) + 1j * rng.normal(size=(5, idx3.size))
    Yref = np.stack([Ch.mvp_circ_2_level(inv, x, L, M, N, idx3) for x in X])
    assert _rel(pc.apply(X[0]), Yref[0]) <= 1e-11
    Xd = api.DeviceArray.from_numpy(X, np.complex128, ctx)
    Yd = api.DeviceArray(X.shape, np.complex128, ctx)
    pc.apply_dev(Xd.ptr, Yd.ptr, nrhs=5)
    assert _rel(Yd.numpy(), Yref) <= 1e-11


@pytest.mark.parametrize("which", ["plate32x32x8", "koch128x128x16"])
def test_fp32_apply_matches_fp64(ctx, which):
    """FP32 2-level apply of one right-hand side against the FP64 apply (relative difference <= 2e-5,
    FP32 rounding), and against the FP32 apply of a two-right-hand-side batch (two right-hand sides per
    block pass) to the same level, on the 32x32x8 plate and the full-size Koch prism."""
    from voxscatter_amd.capi import VS_FP32, VS_FP64
    from voxscatter_amd.dda import DDASystem
    print(f"[fp32 apply] {which}", flush=True)
    s = G.hex_plate() if which.startswith("plate") else G.koch_prism(128, 128, 16, levels=3, kR=20.0,
                                                                        m=1.31 + 0.01j)
    sysd = DDASystem.build(s, precisions=VS_FP64 | VS_FP32, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=VS_FP64 | VS_FP32)
    n = sysd.op.n
    rng = np.random.default_rng(5)
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    y64 = api.DeviceArray((1, n), np.complex128, ctx)
    x64 = api.DeviceArray.from_numpy(X[:1], np.complex128, ctx)
    pc.apply_dev(x64.ptr, y64.ptr, VS_FP64, 1)
    x32 = api.DeviceArray.from_numpy(X, np.complex64, ctx)
    y32 = api.DeviceArray((2, n), np.complex64, ctx)
    pc.apply_dev(x32.ptr, y32.ptr, VS_FP32, 1)
    one = y32.numpy()[0].astype(np.complex128)
    ref = y64.numpy()[0]
    assert _rel(one, ref) <= 2e-5
    pc.apply_dev(x32.ptr, y32.ptr, VS_FP32, 2)
    two = y32.numpy()[0].astype(np.complex128)
    assert _rel(two, ref) <= 2e-5 and _rel(one, two) <= 2e-5


@pytest.mark.parametrize("shape", [(10, 8, 4), (9, 7, 5), (6, 5, 30), (7, 6, 36), (16, 12, 32)])
def test_fp32_matrix_core_apply_matches_oracle(ctx, shape, monkeypatch):
    """The FP32 2-level apply with the block product on the matrix cores (16x16x4 FP32 MFMA, complex as
    four real products) against the oracle's FP64 apply with every block inverted, for 1, 3, 5, 17 and 33
    right-hand sides (one to nine 16-column tiles, one and two column groups) on even and odd grids and
    block sizes 3N = 12 ... 108 that are not multiples of 16; and against the vector-unit FP32 product.
    Both variants of the product: one block per workgroup, and tiles of 8 consecutive blocks."""
    from voxscatter_amd.capi import VS_FP32, VS_FP64
    print(f"[mfma apply] shape={shape}", flush=True)
    L, M, N = shape
    k, d = 2 * math.pi, 0.07
    r = G.centred_grid(L, M, N, d)
    rng = np.random.default_rng(L * 7 + N)
    nD = L * M * N
    vox = np.sort(rng.choice(nD, size=max(1, nD * 2 // 3), replace=False))
    idx3 = np.concatenate([vox, nD + vox, 2 * nD + vox])
    Toep = O.toeplitz_dda(r, k)
    alpha = O.alpha_ldr(1.31 + 0.01j, k, d, [0, 0, 1], [1, 0, 0])
    _, cL = Ch.circ_1_level(Toep, L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    inv = Ch.invert_blocks_dda(blk, alpha, d ** 3)
    op = api.Operator.dda(r, k, ctx=ctx)
    op.set_domain(idx3, 1 / alpha, 1.0, d ** 3)
    # the matrix-core copy of the blocks is built for small blocks only when that product is asked for
    monkeypatch.setenv("VS_BMM_MFMA_MIN", "1")
    pc = api.Preconditioner.from_operator(op, 2, 1 / alpha, wscalar=d ** 3, precisions=VS_FP32 | VS_FP64)
    # tiles: 8 blocks x one wavefront (cfg 0) and 4 blocks x two wavefronts (cfg 1, the default for block
    # sizes that are multiples of 96), both forced on every shape
    for nrhs, tile, cfg in ((1, "1", "0"), (3, "1", "1"), (5, "0", "0"), (5, "1", "0"), (5, "1", "1"),
                            (17, "1", "0"), (17, "1", "1"), (33, "0", "0"), (33, "1", "0"), (33, "1", "1")):
        X = rng.normal(size=(nrhs, idx3.size)) + 1j * rng.normal(size=(nrhs, idx3.size))
        Yref = np.stack([Ch.mvp_circ_2_level(inv, x, L, M, N, idx3) for x in X])
        Xd = api.DeviceArray.from_numpy(X, np.complex64, ctx)
        Yd = api.DeviceArray(X.shape, np.complex64, ctx)
        monkeypatch.setenv("VS_BMM_MFMA_MIN", "1")
        monkeypatch.setenv("VS_BMM_TILE", tile)
        monkeypatch.setenv("VS_BMT_CFG", cfg)
        pc.apply_dev(Xd.ptr, Yd.ptr, VS_FP32, nrhs)
        ym = Yd.numpy().astype(np.complex128)
        monkeypatch.setenv("VS_BMM_MFMA_MIN", "100000")
        pc.apply_dev(Xd.ptr, Yd.ptr, VS_FP32, nrhs)
        yv = Yd.numpy().astype(np.complex128)
        em, ev = _rel(ym, Yref), _rel(ym, yv)
        print(f"[mfma apply] nrhs={nrhs} tile={tile} cfg={cfg}: vs oracle {em:.2e}, vs vector-unit FP32 "
              f"{ev:.2e}", flush=True)
        assert em <= 2e-5 and ev <= 2e-5
