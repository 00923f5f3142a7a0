"""Voxel grids, shape masks and the seeded synthetic scatterers used as solver inputs.

This is host-side problem set-up (the reference spends negligible time here too). It follows the
reference's grid conventions so that a grid built here is the grid the MATLAB drivers would build:

* ``generatedomain_new``  -- src/src_geometry/src_domain/generatedomain_new.m:81-104
* ``grid3d``              -- src/src_geometry/src_domain/grid3d.m:35-61
* ``geometry``            -- src/src_geometry/geometry.m:1-71
* ``koch_polygon``        -- src/src_geometry/koch_snowflake.m:7-21 (iteration count is a parameter)
* ``koch_snowflake``      -- src/src_geometry/koch_snowflake.m:25-61,101

Arrays are MATLAB-ordered: ``r`` has shape (L, M, N, 3) and linear voxel indices are column-major
(x fastest).  Index arrays returned here are 0-based.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

EPS = np.finfo(np.float64).eps


def matlab_round(x):
    """MATLAB ``round``: half away from zero (numpy rounds half to even)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def matlab_colon(a: float, d: float, b: float) -> np.ndarray:
    """Elements of the MATLAB range ``a:d:b``.

    MATLAB builds the first half of the range forwards from ``a`` and the second half backwards from
    the last element, so that symmetric ranges stay symmetric to the last bit.
    """
    tol = 2.0 * EPS * max(abs(a), abs(b))
    n = int(math.floor((b - a) / d + tol / abs(d)))
    if n < 0:
        return np.zeros(0)
    k = np.arange(n + 1, dtype=np.float64)
    last = a + n * d
    out = np.where(k <= n // 2, a + k * d, last - (n - k) * d)
    return out


def grid3d(x, y, z) -> np.ndarray:
    """4-D (L, M, N, 3) coordinate array (grid3d.m:35-61)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    r = np.zeros((x.size, y.size, z.size, 3))
    r[..., 0] = x[:, None, None]
    r[..., 1] = y[None, :, None]
    r[..., 2] = z[None, None, :]
    return r


def generatedomain_new(res: float, x, y=None, z=None) -> np.ndarray:
    """Grid of voxel centres whose edges match the domain (generatedomain_new.m:81-104)."""
    def lims(v):
        v = np.atleast_1d(np.asarray(v, dtype=np.float64))
        return (-v[0], v[0]) if v.size == 1 else (float(v.min()), float(v.max()))

    y = x if y is None else y
    z = x if z is None else z
    axes = []
    for v in (x, y, z):
        lo, hi = lims(v)
        n = max(1, int(matlab_round((hi - lo) / res)))
        c = (hi + lo) / 2.0
        axes.append(matlab_colon(c - (n / 2 - 0.5) * res, res, c + (n / 2 - 0.5) * res))
    return grid3d(*axes)


def centred_grid(L: int, M: int, N: int, d: float) -> np.ndarray:
    """L x M x N grid of spacing ``d`` centred on the origin (the generatedomain_new convention)."""
    axes = [matlab_colon(-(n / 2 - 0.5) * d, d, (n / 2 - 0.5) * d) for n in (L, M, N)]
    return grid3d(*axes)


def hex_mask(r: np.ndarray, a: float) -> np.ndarray:
    """Hexagon of circumradius ``a`` by the four sloping half-planes of geometry.m:56-61.

    The flat top and bottom faces (|y| <= sqrt(3)/2 a) are implied by the grid extent in the reference;
    here they are applied explicitly so that the mask does not depend on the box size.
    """
    x = r[..., 0]
    y = r[..., 1]
    s3 = math.sqrt(3.0)
    l1 = y - s3 * (a - x)
    l2 = y - s3 * (x + a)
    l3 = y - s3 * (-a - x)
    l4 = y - s3 * (x - a)
    b = s3 / 2 * a
    return (l1 < 0) & (l2 < 0) & (l3 > 0) & (l4 > 0) & (np.abs(y) <= b)


def find_F(mask: np.ndarray) -> np.ndarray:
    """0-based column-major linear indices of true voxels (MATLAB ``find``)."""
    return np.flatnonzero(np.asarray(mask).reshape(-1, order="F")).astype(np.int64)


def geometry(geom: str, refInd: complex, sizeParam: float, nPerLam: float):
    """Shapes of geometry.m:1-71: 'sphere', 'spheroid', 'hex', 'cube'.

    Returns ``(r, idx, res, P, lambda_ext, lambda_int)`` with 0-based ``idx``.
    """
    if geom == "spheroid":
        a, c = 1.0, 0.5
        dom = (2 * a, 2 * a, 2 * c)
        P = math.pi * a * c
    elif geom == "sphere":
        a = c = 1.0
        dom = (2 * a, 2 * a, 2 * a)
        P = math.pi * a * a
    elif geom == "hex":
        a = 1.0
        b = math.sqrt(3) / 2 * a
        aspect = 1 / 10
        dom = (2 * a, 2 * b, a * aspect)
        P = dom[1] * dom[2]
    elif geom == "cube":
        a = 1.0
        dom = (2 * a, 2 * a, 2 * a)
        P = 4 * a * a
    else:
        raise ValueError(f"unknown geometry {geom!r}")
    lambda_ext = 2 * math.pi * a / sizeParam
    lambda_int = lambda_ext / np.real(refInd)
    h_pref = dom[0]
    res_temp = lambda_int / nPerLam
    n = math.ceil(h_pref / res_temp)
    res = h_pref / n
    r = generatedomain_new(res, dom[0] / 2, dom[1] / 2, dom[2] / 2)
    if geom == "hex":
        x = r[..., 0]
        y = r[..., 1]
        s3 = math.sqrt(3.0)
        mask = ((y - s3 * (a - x)) < 0) & ((y - s3 * (x + a)) < 0) & \
               ((y - s3 * (-a - x)) > 0) & ((y - s3 * (x - a)) > 0)
    elif geom in ("spheroid", "sphere"):
        mask = (r[..., 0] ** 2 + r[..., 1] ** 2) / a ** 2 + r[..., 2] ** 2 / c ** 2 <= 1
    else:
        mask = np.ones(r.shape[:3], dtype=bool)
    return r, find_F(mask), res, P, lambda_ext, lambda_int


def koch_polygon(levels: int = 5) -> np.ndarray:
    """Closed vertex list of the Koch snowflake after ``levels`` refinements (koch_snowflake.m:7-21).

    The first row is the origin placeholder produced by the reference's preallocation (``newP(1,:)``
    stays zero); it coincides with the first triangle vertex, so the polygon is unchanged by it.
    """
    P = np.array([[0.0, 0.0], [1.0, 0.0], [math.cos(-math.pi / 3), math.sin(-math.pi / 3)], [0.0, 0.0]])
    for _ in range(levels):
        m = P.shape[0]
        newP = np.zeros((m * 4 + 1, 2))
        for i in range(m - 1):
            newP[4 * (i + 1) + 0] = P[i]
            newP[4 * (i + 1) + 1] = (2 * P[i] + P[i + 1]) / 3
            link = P[i + 1] - P[i]
            ang = math.atan2(link[1], link[0])
            leng = math.sqrt(float(np.sum(link ** 2)))
            newP[4 * (i + 1) + 2] = newP[4 * (i + 1) + 1] + (leng / 3) * np.array(
                [math.cos(ang + math.pi / 3), math.sin(ang + math.pi / 3)])
            newP[4 * (i + 1) + 3] = (P[i] + 2 * P[i + 1]) / 3
        newP[4 * m] = P[m - 1]
        P = newP
    # centre on the origin and scale to unit half-height (koch_snowflake.m:25-39)
    xmin, ymin = P.min(axis=0)
    xmax, ymax = P.max(axis=0)
    dom_y = ymax - ymin
    P = P - np.array([(xmax + xmin) / 2, (ymax + ymin) / 2])
    return P / (dom_y / 2)


def inpolygon(xq, yq, xv, yv) -> np.ndarray:
    """MATLAB ``inpolygon``: true for points strictly inside or on an edge (crossing-number rule)."""
    xq = np.asarray(xq, dtype=np.float64).ravel()
    yq = np.asarray(yq, dtype=np.float64).ravel()
    xv = np.asarray(xv, dtype=np.float64).ravel()
    yv = np.asarray(yv, dtype=np.float64).ravel()
    if xv[0] != xv[-1] or yv[0] != yv[-1]:
        xv = np.append(xv, xv[0])
        yv = np.append(yv, yv[0])
    inside = np.zeros(xq.shape, dtype=bool)
    on = np.zeros(xq.shape, dtype=bool)
    scale = max(np.ptp(xv), np.ptp(yv), 1.0)
    tol = 64 * EPS * scale
    for i in range(xv.size - 1):
        x1, y1, x2, y2 = xv[i], yv[i], xv[i + 1], yv[i + 1]
        if x1 == x2 and y1 == y2:
            continue
        cond = (y1 > yq) != (y2 > yq)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = x1 + (yq - y1) * (x2 - x1) / (y2 - y1)
        inside ^= cond & (xq < xint)
        cross = (x2 - x1) * (yq - y1) - (y2 - y1) * (xq - x1)
        seg = math.hypot(x2 - x1, y2 - y1)
        within = (np.minimum(x1, x2) - tol <= xq) & (xq <= np.maximum(x1, x2) + tol) & \
                 (np.minimum(y1, y2) - tol <= yq) & (yq <= np.maximum(y1, y2) + tol)
        on |= within & (np.abs(cross) <= tol * seg)
    return inside | on


def koch_snowflake(lambda_int: float, nPerLam: float, aspectRatio: float, levels: int = 5):
    """Koch-snowflake prism of koch_snowflake.m:25-61,101 (a script there; a function here).

    Returns ``(r, idx, P)`` with 0-based ``idx`` and polygon ``P`` (unit half-height).
    """
    P = koch_polygon(levels)
    xmin, ymin = P.min(axis=0)
    xmax, ymax = P.max(axis=0)
    dom_x, dom_y = xmax - xmin, ymax - ymin
    dom_z = dom_y * aspectRatio
    h_pref = dom_y
    n = math.ceil(h_pref / (lambda_int / nPerLam))
    res = h_pref / n
    r = generatedomain_new(res, dom_x / 2, dom_y / 2, dom_z / 2)
    inside = inpolygon(r[..., 0].reshape(-1, order="F"), r[..., 1].reshape(-1, order="F"), P[:, 0], P[:, 1])
    return r, np.flatnonzero(inside).astype(np.int64), P


# --------------------------------------------------------------------------------------------------
# Seeded synthetic scatterers (the solve-loop workloads).  Units: vacuum wavelength 1, k = 2*pi.
# --------------------------------------------------------------------------------------------------

@dataclass
class Scatterer:
    """A voxelised particle plus its plane-wave excitation."""
    name: str
    r: np.ndarray                    # (L, M, N, 3) voxel centres
    idx: np.ndarray                  # 0-based linear indices of occupied voxels
    refInd: complex                  # refractive index (scalar)
    k0: float                        # vacuum wavenumber
    dInc: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    Eo: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0]))
    eps_r: np.ndarray | None = None  # optional per-voxel relative permittivity (L, M, N)

    @property
    def shape(self):
        return self.r.shape[:3]

    @property
    def d(self) -> float:
        return float(self.r[1, 0, 0, 0] - self.r[0, 0, 0, 0])


def hex_plate(L=32, M=32, N=8, m=1.31 + 0.0j, a_vox=16.0) -> Scatterer:
    """Hexagonal plate on an L x M x N grid, d = lambda/(10|m|), circumradius ``a_vox`` voxels.

    Incidence +z, polarisation +x.  Default: the 32 x 32 x 8 ice plate (aspect 4:1).
    """
    k0 = 2 * math.pi
    d = 1.0 / (10 * abs(m))
    r = centred_grid(L, M, N, d)
    mask = hex_mask(r, a_vox * d)
    return Scatterer("hex_plate", r, find_F(mask), complex(m), k0,
                     np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]))


def koch_prism(L=128, M=128, N=16, levels=3, kR=20.0, m=1.31 + 0.01j) -> Scatterer:
    """Koch-snowflake prism: polygon half-height R with k R = ``kR``, d = 2R/M, thickness N d.

    Incidence +x, polarisation +z (main_koch.m:57-58).
    """
    k0 = 2 * math.pi
    R = kR / k0
    d = 2 * R / M
    P = koch_polygon(levels) * R
    r = centred_grid(L, M, N, d)
    inside = inpolygon(r[..., 0].reshape(-1, order="F"), r[..., 1].reshape(-1, order="F"), P[:, 0], P[:, 1])
    return Scatterer("koch_prism", r, np.flatnonzero(inside).astype(np.int64), complex(m), k0,
                     np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]))


def sphere(n=16, m=1.5 + 0.0j, d=None) -> Scatterer:
    """Sphere inscribed in an n^3 grid (mask of geometry.m:64-68), incidence +z, polarisation +x."""
    k0 = 2 * math.pi
    d = 1.0 / (10 * abs(m)) if d is None else d
    r = centred_grid(n, n, n, d)
    a = n * d / 2
    mask = (r[..., 0] ** 2 + r[..., 1] ** 2 + r[..., 2] ** 2) <= a * a
    return Scatterer("sphere", r, find_F(mask), complex(m), k0)


def perturbed_hex_plate(L=96, M=96, N=24, m=1.31 + 0.01j, rel=0.05, seed=0) -> Scatterer:
    """Hex plate whose per-voxel permittivity is m^2 (1 + rel U(-1,1)), seeded (PCG64)."""
    s = hex_plate(L, M, N, m, a_vox=L / 2)
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(-1.0, 1.0, size=(L, M, N))
    s.eps_r = (complex(m) ** 2) * (1.0 + rel * u)
    s.name = "perturbed_hex_plate"
    return s


# --------------------------------------------------------------------------------------------------
# batches of particles and incident fields (multi-RHS workloads)
# --------------------------------------------------------------------------------------------------

def fibonacci_sphere(n: int) -> np.ndarray:
    """n nearly uniform unit vectors (golden-angle spiral)."""
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = math.pi * (3 - math.sqrt(5)) * i
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=-1)


def polarisations(khat) -> tuple[np.ndarray, np.ndarray]:
    """Two real unit vectors orthogonal to khat and to each other."""
    khat = np.asarray(khat, float)
    a = np.array([0.0, 0.0, 1.0]) if abs(khat[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    e1 = np.cross(khat, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(khat, e1)
    return e1, e2


def plane_wave_batch(scat: Scatterer, ndir: int = 32):
    """Right-hand sides [Ex; Ey; Ez](idx) for ndir Fibonacci directions x 2 polarisations: (2 ndir, 3 nvox).

    PlaneWavePointWise.m:7-19 (E0 exp(i k.r) at the voxel centres) evaluated at the occupied voxels only,
    one phase per direction shared by its two polarisations."""
    rv = np.asarray(scat.r, dtype=np.float64).reshape(-1, 3, order="F")[np.asarray(scat.idx)]
    dirs = fibonacci_sphere(ndir)
    phase = np.exp(1j * (rv @ (scat.k0 * dirs).T))          # (nvox, ndir)
    nv = rv.shape[0]
    out = np.empty((2 * ndir, 3 * nv), dtype=np.complex128)
    meta = []
    for d, khat in enumerate(dirs):
        for p, e in enumerate(polarisations(khat)):
            row = out[2 * d + p]
            for c in range(3):
                row[c * nv:(c + 1) * nv] = e[c] * phase[:, d]
            meta.append((khat, e))
    return out, meta


def plane_wave_batch_dev(scat: Scatterer, ndir: int = 32, ctx=None):
    """``plane_wave_batch`` evaluated on the device: a (2 ndir, 3 nvox) complex128 DeviceArray and the
    (khat, e) of each row (vs_plane_wave_batch_dev)."""
    from . import api
    from .capi import check, dptr
    ctx = ctx or api.get_context()
    rv = np.ascontiguousarray(np.asarray(scat.r, dtype=np.float64).reshape(-1, 3, order="F")[np.asarray(scat.idx)])
    dirs = np.ascontiguousarray(fibonacci_sphere(ndir))
    meta, pol = [], []
    for khat in dirs:
        for e in polarisations(khat):
            meta.append((khat, e))
            pol.append(e)
    pol = np.ascontiguousarray(np.asarray(pol, dtype=np.float64))
    B = api.DeviceArray((2 * ndir, 3 * rv.shape[0]), np.complex128, ctx)
    check(ctx.lib.vs_plane_wave_batch_dev(ctx.h, rv.shape[0], dptr(rv), ndir, dptr(dirs), 2, dptr(pol),
                                          float(scat.k0), api.C.c_void_p(B.ptr)))
    return B, meta


def _rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_particle(rng, n: int, shape: str, m: complex) -> Scatterer:
    """Hex plate (aspect 4:1), hexagonal column (length 2 x diameter) or ellipsoid (1 : 0.7 : 0.5), randomly
    oriented, inscribed in an n^3 grid of spacing lambda/(10|m|)."""
    k0 = 2 * math.pi
    d = 1.0 / (10 * abs(m))
    r = centred_grid(n, n, n, d)
    R = _rotation(rng)
    q = r.reshape(-1, 3) @ R          # body-frame coordinates
    rb = 0.5 * n * d * 0.98           # bounding-sphere radius
    x, y, z = q[:, 0], q[:, 1], q[:, 2]
    s3 = math.sqrt(3.0)
    if shape == "ellipsoid":
        a = rb
        mask = (x / a) ** 2 + (y / (0.7 * a)) ** 2 + (z / (0.5 * a)) ** 2 <= 1
    else:
        if shape == "plate":     # diameter 2a, thickness a/2 -> bounding radius sqrt(a^2 + (a/4)^2)
            a = rb / math.sqrt(1 + 1 / 16)
            h = a / 4
        else:                    # column: diameter 2a, length 4a
            a = rb / math.sqrt(1 + 4)
            h = 2 * a
        hexm = (np.abs(y) <= s3 / 2 * a) & (np.abs(s3 * x) + np.abs(y) <= s3 * a)
        mask = hexm & (np.abs(z) <= h)
    mask = mask.reshape(n, n, n)
    idx = find_F(mask)
    return Scatterer(f"{shape}{n}", r, idx, complex(m), k0)


def particle_batch(count: int = 96, seed: int = 0, sizes=(24, 32, 48, 64, 96)):
    """Seeded particles: size uniform over ``sizes``, shape over plate/column/ellipsoid,
    Re m ~ U[1.31, 1.6], Im m = 0.005, random orientation."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        shape = str(rng.choice(["plate", "column", "ellipsoid"]))
        m = complex(rng.uniform(1.31, 1.6), 0.005)
        out.append(random_particle(rng, n, shape, m))
    return out
