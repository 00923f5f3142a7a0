"""Benchmark of the solve loop on one MI355X.

Workload: level-3 Koch-snowflake prism on a 128 x 128 x 16 grid (k R = 20, k d = 0.3125,
m = 1.31 + 0.01i, incidence +x, polarisation +z), DDA with the LDR polarisability.  One step is one
complete solve, right-hand side resident in HBM to solution resident in HBM with the FP64 relative
residual norm(b - A x)/norm(b) <= 1e-6 verified: BiCGStab with the 2-level Chan circulant
preconditioner in complex FP32 inside FP64 iterative refinement.  The plain FP64 solve (BiCGStab +
FP64 2-level preconditioner) of the same system is timed beside it.

Prints one JSON line (see ``main``).  ``--gpus N`` with N > 1 exits: the build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
TOL = 1e-6


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=2,
                    help="BiCGStab iterations timed on one core for the single-core CPU figure")
    ap.add_argument("--inner-tol", type=float, default=TOL / 3,
                    help="FP32 inner-solve tolerance (default tol/3: one inner solve, the FP64 residual confirms it)")
    ap.add_argument("--inner", default="bicgstab", choices=["bicgstab", "gmres"])
    ap.add_argument("--sweep", action="store_true", help="also time other inner solvers/tolerances (stderr)")
    ap.add_argument("--workload", default="koch", choices=["koch", "particles", "galerkin"],
                    help="koch: the 128x128x16 prism (default); particles: seeded particles of grids 24^3..96^3, "
                         "2 x ndir plane waves each, batched FP32 + FP64-refinement solves; galerkin: the "
                         "high-contrast 64^3 Galerkin cube, FP64 GMRES(100) against the precision form")
    ap.add_argument("--seed", type=int, default=0, help="galerkin workload: seed of the permittivity field")
    ap.add_argument("--particles", type=int, default=96, help="particles workload: number of particles")
    ap.add_argument("--ndir", type=int, default=32, help="particles workload: incident directions per particle")
    ap.add_argument("--matvec-only", action="store_true",
                    help="only run 20 FP32 and 20 FP64 matvecs, 20 preconditioner applies of each precision and "
                         "5 device-to-device copies of 256 MiB (a fixed launch sequence for hardware-counter runs)")
    return ap.parse_args()


COPY_BYTES = 256 << 20
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "matvec_traffic_koch128x128x16.txt")


def measured_traffic(variant: str):
    """HBM bytes per FP32 matvec of the given variant from the committed hardware-counter summary
    (FETCH_SIZE + WRITE_SIZE of the matvec kernels, scaled by the same counters on a device copy of known
    size), or None when the summary holds no figure for that variant."""
    key = f"fp32_matvec_hbm_bytes[{variant}]:"
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(key):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def matvec_only(op, pc, ctx, api, VS_FP32, VS_FP64):
    n = op.n
    rng = np.random.default_rng(0)
    xr = rng.normal(size=n) + 1j * rng.normal(size=n)
    for prec, dt in ((VS_FP32, np.complex64), (VS_FP64, np.complex128)):
        xv = api.DeviceArray.from_numpy(xr, dt, ctx)
        yv = api.DeviceArray(n, dt, ctx)
        for _ in range(20):
            op.mvp_dev(xv.ptr, yv.ptr, False, prec, 1)
        for _ in range(20):
            pc.apply_dev(xv.ptr, yv.ptr, prec, 1)
        ctx.synchronize()
        print(f"[bench] matvec-only: 20 matvecs + 20 pc applies, precision {prec}", file=sys.stderr, flush=True)
    a = api.DeviceArray(COPY_BYTES // 8, np.complex64, ctx)
    c = api.DeviceArray(COPY_BYTES // 8, np.complex64, ctx)
    for _ in range(5):
        api.check(ctx.lib.vs_dev_copy(ctx.h, api.C.c_void_p(c.ptr), api.C.c_void_p(a.ptr), COPY_BYTES, 2))
    print(json.dumps({"matvec_only": True, "matvecs_per_precision": 20, "pc_applies_per_precision": 20,
                      "copy_bytes": COPY_BYTES, "copies": 5}), flush=True)


def particles(args, api, geometry, ctx):
    """Orientation-averaging workload: ``args.particles`` seeded particles (seed 0; grids 24^3, 32^3, 48^3,
    64^3, 96^3; plates, columns, ellipsoids; Re m in [1.31, 1.6], Im m = 0.005; random orientation), each
    solved for 2 x ndir plane waves in one batched FP32 + FP64-refinement solve to FP64 rel-res 1e-6.
    value = right-hand sides per second of solve time (set-up of each particle's operator and inverse
    blocks reported apart).  One progress line per particle on stderr."""
    from voxscatter_amd.particles import solve_particles
    parts = geometry.particle_batch(count=args.particles, seed=0)
    # warm-up: one small particle compiles nothing new but settles the allocator and the autotuner
    small = min(parts, key=lambda p: p.idx.size)
    solve_particles([small], ndir=args.ndir, tol=TOL, ctx=ctx)

    def progress(i, r):
        print(f"[particles] {i + 1}/{len(parts)} {r.name} unknowns={r.unknowns} outer={r.outer} inner={r.inner} "
              f"max_relres={r.relres.max():.2e} setup={r.setup_s:.3f}s solve={r.solve_s:.3f}s",
              file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    res = solve_particles(parts, ndir=args.ndir, tol=TOL, ctx=ctx, on_done=progress)
    wall = time.perf_counter() - t0
    solve = sum(r.solve_s for r in res)
    setup = sum(r.setup_s for r in res)
    nrhs = sum(r.nrhs for r in res)
    by_grid = {}
    for r in res:
        g = by_grid.setdefault(str(r.grid[0]), {"particles": 0, "solve_s": 0.0, "rhs": 0})
        g["particles"] += 1
        g["solve_s"] = round(g["solve_s"] + r.solve_s, 4)
        g["rhs"] += r.nrhs
    out = {
        "metric": "orientation-averaged particle solves: right-hand sides per second @rel-res 1e-6 "
                  "(FP32+refine)",
        "value": round(nrhs / solve, 2),
        "unit": "rhs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(solve * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner / c128 refinement",
        "data": "synthetic",
        "config": {"workload": f"{len(parts)} seeded particles (grids 24^3..96^3), {2 * args.ndir} plane waves "
                               "each, DDA-LDR, BiCGStab + 2-level Chan in FP32 inside FP64 refinement",
                   "particles": len(parts), "rhs": nrhs, "unknowns_total": int(sum(r.unknowns for r in res))},
        "solve_s": round(solve, 3),
        "setup_s": round(setup, 3),
        "wall_s": round(wall, 3),
        "all_converged": bool(all(np.all(r.flags == 0) for r in res)),
        "max_relres": float(max(r.relres.max() for r in res)),
        "by_grid": by_grid,
        "roofline": particles_roofline(api, parts, ctx, args.ndir),
        "cpu_baseline": None if args.no_cpu_baseline else particles_cpu_baseline(parts, args.ndir),
    }
    print(json.dumps(out), flush=True)


MFMA_FP32_PEAK_TFLOPS = 157.3


def particles_roofline(api, parts, ctx, ndir):
    """The dominant kernel of the particle workload: the FP32 2-level block product on the matrix cores
    for all 2 ndir right-hand sides of the largest particle (its 4 mirror points per stored block), timed
    alone with HIP events; algorithmic flops = L M n^2 x 8 per right-hand side (n = 3N: every grid
    point's block applied once, complex multiply-add = 8 real flops)."""
    from voxscatter_amd.capi import VS_FP32, VS_FP64
    from voxscatter_amd.dda import DDASystem
    p = max(parts, key=lambda q: q.r.shape[0] * 10 ** 7 + q.idx.size)
    L, M, N = p.shape
    sysd = DDASystem.build(p, precisions=VS_FP64 | VS_FP32, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=VS_FP32)
    nrhs = 2 * ndir
    ms = api.bench_pc_blocks(pc, VS_FP32, nrhs, 5)
    n = 3 * N
    flops = 8.0 * L * M * n * n * nrhs
    tf = flops / (ms * 1e-3) / 1e12
    pc.close()
    sysd.op.close()
    print(f"[bench] particles roofline: {p.name} {nrhs}-RHS block product {ms:.3f} ms, {tf:.1f} TFLOP/s",
          file=sys.stderr, flush=True)
    return {"bound": "mfma", "achieved": round(tf, 2), "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / MFMA_FP32_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": f"FP32 2-level block product on the matrix cores (k_blockmv2t; "
                      f"{'4 stored blocks x 2 wavefronts' if n % 96 == 0 else '8 stored blocks x 1 wavefront'} "
                      f"per workgroup), {nrhs} right-hand sides of "
                      f"particle {p.name} ({L}x{M}x{N}, blocks {n}x{n})",
            "flops_per_launch": flops, "ms_per_launch": round(ms, 4)}


def particles_cpu_baseline(parts, ndir):
    """Oracle FP64 BiCGStab + 2-level preconditioner on the host cores, on a bounded sample of the workload:
    the smallest particle, its first 2 plane waves, each with its own incidence's polarisability."""
    from oracle import _fft
    from oracle import circulant as Ch
    from oracle import krylov as OK
    from oracle import mvp as OM
    from oracle import operator as O
    from voxscatter_amd.geometry import plane_wave_batch
    from voxscatter_amd.particles import incidence_diagonals
    p = min(parts, key=lambda q: q.idx.size)
    L, M, N = p.shape
    nD = L * M * N
    d = p.r[1, 0, 0, 0] - p.r[0, 0, 0, 0]
    dV = d ** 3
    idx3 = np.concatenate([p.idx, nD + p.idx, 2 * nD + p.idx])
    B, meta = plane_wave_batch(p, ndir=ndir)
    diag = incidence_diagonals(p, meta)
    Tref = O.toeplitz_dda(p.r, p.k0)
    fref = O.fft_operator(O.circulant_nop_const(Tref))
    _, cL = Ch.circ_1_level(Tref, L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    ones = np.ones((L, M, N))
    chi = np.zeros(nD)
    chi[p.idx] = 1
    chi = chi.reshape((L, M, N), order="F")
    t0 = time.perf_counter()
    its = []
    for j in range(2):
        inv = np.linalg.inv(diag[j] * np.eye(3 * N) - dV * blk)

        def afun(v, j=j):
            return OM.MVP(v, fref, ones, dV * chi, diag[j], "notransp", idx3)

        def mfun(v, inv=inv):
            return Ch.mvp_circ_2_level(inv, v, L, M, N, idx3)

        _, flag, relres, it, _ = OK.bicgstab(afun, B[j], TOL, 2000, mfun)
        its.append(it)
    dt = time.perf_counter() - t0
    return {"value": round(2 / dt, 3), "unit": "rhs/s", "cores": _fft.WORKERS, "kind": "port",
            "sample": (f"2 plane waves of particle {p.name} ({L}x{M}x{N}, {p.idx.size} voxels), FP64 BiCGStab + "
                       f"2-level solves by the NumPy/SciPy restatement of the MATLAB path, preconditioner built "
                       f"per incidence as the reference does ({its} iterations, {dt:.1f} s); the smallest "
                       f"particle of the set, so larger particles would be slower per right-hand side")}


def galerkin_eps(L, M, N, seed, sigma=4.0):
    """m = 1.5 + 2u + 0.01i, u = uniform noise smoothed by a periodic Gaussian of sigma voxels, min-max scaled."""
    import math
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=(L, M, N))
    k = [np.fft.fftfreq(n) for n in (L, M, N)]
    KX, KY, KZ = np.meshgrid(*k, indexing="ij")
    g = np.exp(-2 * (math.pi * sigma) ** 2 * (KX ** 2 + KY ** 2 + KZ ** 2))
    sm = np.real(np.fft.ifftn(np.fft.fftn(u) * g))
    sm = (sm - sm.min()) / (sm.max() - sm.min())
    return (1.5 + 2.0 * sm + 0.01j) ** 2


def galerkin(args, api, ctx):
    """High-contrast Galerkin cube: 64^3, m in [1.5, 3.5] + 0.01i (seeded smooth random field, sigma 4
    voxels), voxel lambda0 / 70, form 2 with the averaged-contrast 2-level preconditioner.  One step =
    one solve to relative residual 1e-6 from a device-resident right-hand side: FP64 GMRES(100), and the
    precision form (FP32 matvec and blocks, inner GMRES(100) to 1e-3, FP64 refinement).  value = the
    precision form's time; the internal fields of the two solutions are compared."""
    import math
    from voxscatter_amd.capi import VS_FP32, VS_FP64
    from voxscatter_amd.galerkin import GalerkinSystem
    L = M = N = 64
    eps = galerkin_eps(L, M, N, args.seed)
    sysg = GalerkinSystem.build(eps, 1 / 70, 2 * math.pi, precisions=VS_FP64 | VS_FP32, ctx=ctx)
    pc = sysg.preconditioner(2, precisions=VS_FP64 | VS_FP32)
    op = sysg.op
    b = api.DeviceArray.from_numpy(sysg.b, np.complex128, ctx)
    x64 = api.DeviceArray(op.n, np.complex128, ctx)
    xr = api.DeviceArray(op.n, np.complex128, ctx)

    def run64():
        return api.gmres_dev(op, b.ptr, x64.ptr, 100, TOL, 20, pc)

    def runr():
        return api.refine_dev(op, b.ptr, xr.ptr, pc, tol=TOL, max_outer=10, inner="gmres", inner_tol=1e-3,
                              restart=100)

    def timed(f):
        for _ in range(args.warmup):
            f()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = f()
        ctx.synchronize()
        return (time.perf_counter() - t0) / args.steps, res

    t64, (f64, rr64, it64, _) = timed(run64)
    tr, (fr, rrr, outer, inner) = timed(runr)
    X64, XR = x64.numpy(), xr.numpy()
    nb = np.linalg.norm(sysg.b)
    r64 = float(np.linalg.norm(op.mvp(X64) - sysg.b) / nb)
    rr = float(np.linalg.norm(op.mvp(XR) - sysg.b) / nb)
    E64, Er = sysg.total_field(X64), sysg.total_field(XR)
    err = float(np.linalg.norm((Er - E64).ravel()) / np.linalg.norm(E64.ravel()))
    print(f"[bench] galerkin: FP64 GMRES(100) {t64 * 1e3:.1f} ms iter={it64} flag={f64}; refined {tr * 1e3:.1f} ms "
          f"outer={outer} inner={inner} flag={fr}; field diff {err:.2e}", file=sys.stderr, flush=True)
    out = {
        "metric": "Galerkin time-to-solution @rel-res 1e-6 (FP32+refine vs FP64 GMRES(100)), internal-field error",
        "value": round(tr * 1e3, 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tr * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner / c128 refinement",
        "data": "synthetic",
        "config": {"workload": f"Galerkin JVIE 64^3 cube, m in [1.5,3.5]+0.01i (seed {args.seed}, sigma 4 voxels), "
                               "voxel lambda/70, 2-level averaged-contrast preconditioner, GMRES(100)",
                   "unknowns": int(op.n)},
        "fp64_gmres100_ms": round(t64 * 1e3, 2),
        "fp64_gmres100_iters": list(it64),
        "fp64_flag": f64,
        "refine_flag": fr,
        "refine_outer": outer,
        "refine_inner_iterations": inner,
        "relres_fp64": r64,
        "relres_refined": rr,
        "internal_field_rel_diff": err,
        "speedup_vs_fp64": round(t64 / tr, 3),
        "roofline": galerkin_roofline(op, pc, tr, inner, L, M, N),
        "cpu_baseline": None if args.no_cpu_baseline else galerkin_cpu_baseline(sysg, op, it64, L, M, N),
    }
    print(json.dumps(out), flush=True)


def galerkin_roofline(op, pc, t_solve, inner_iterations, L, M, N):
    """One inner FP32 GMRES(100) iteration of the precision form as the unit: algorithmic bytes = the FP32
    matvec 8 (6 P + 12 V) + the 2-level apply 8 (stored blocks + 18 V) + classical Gram-Schmidt applied
    twice over the Arnoldi basis (per pass the j basis vectors read by the inner products and again by the
    update, j = 50.5 on average over a cycle of 100: 4 x 50.5 n x 8 bytes), against the measured time of
    the whole refined solve divided by its inner iterations (outer FP64 steps included)."""
    V, P = L * M * N, 8 * L * M * N
    n = op.n
    nb = 3 * N
    stored = (L // 2 + 1) * (M // 2 + 1) * nb * nb   # vector weights: full (unpacked) mirrored blocks
    bytes_it = 8 * (6 * P + 12 * V) + 8 * (stored + 18 * V) + 4 * 50.5 * n * 8
    t_it = t_solve / max(1, inner_iterations)
    gbs = bytes_it / t_it / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "one inner FP32 GMRES(100) iteration (FFT matvec + 2-level apply + two classical "
                      "Gram-Schmidt passes over the basis), refined-solve time / inner iterations",
            "bytes_per_launch": int(bytes_it), "ms_per_launch": round(t_it * 1e3, 5)}


def galerkin_cpu_baseline(sysg, op, gpu_iters, L, M, N, sample_iters=6):
    """Oracle FP64 GMRES(100) (MATLAB gmres semantics, left 2-level preconditioner) on the host cores, a
    bounded sample of the Galerkin solve: ``sample_iters`` Arnoldi iterations scaled to the GPU FP64
    iteration count.  The operator's circulant FFT and Toeplitz tensor are taken from the device build
    (input data; the oracle's Python assembly of a 64^3 Galerkin tensor would take hours), the
    preconditioner blocks are built by the oracle."""
    from oracle import _fft
    from oracle import circulant as Ch
    from oracle import krylov as OK
    from oracle import mvp as OM
    from voxscatter_amd.galerkin import averaged_contrast
    fN = op.fN()
    T = op.toeplitz()
    eps = sysg.eps
    Mc = (eps - 1.0) / eps
    ones = np.ones((L, M, N))
    _, cL = Ch.circ_1_level(T, L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    inv = Ch.invert_blocks_galerkin(blk, sysg.Gram, averaged_contrast(eps))

    def afun(v):
        return OM.MVP(v, fN, ones, Mc, sysg.Gram, "notransp", sysg.idx3)

    def mfun(v):
        return Ch.mvp_circ_2_level(inv, v, L, M, N, sysg.idx3)

    t0 = time.perf_counter()
    OK.gmres(afun, sysg.b, sample_iters, 1e-30, 1, mfun)   # tolerance unreachable: exactly sample_iters steps
    dt = time.perf_counter() - t0
    total = int(gpu_iters[0] - 1) * 100 + int(gpu_iters[1]) if isinstance(gpu_iters, (list, tuple)) else gpu_iters
    per = dt / sample_iters
    return {"value": round(per * total * 1e3, 1), "unit": "ms", "cores": _fft.WORKERS, "kind": "port",
            "sample": (f"{sample_iters} GMRES(100) iterations of the NumPy/SciPy FP64 restatement on the same "
                       f"64^3 system ({dt:.1f} s, {per * 1e3:.0f} ms/iteration), scaled to the {total} "
                       f"iterations of the GPU FP64 GMRES(100) solve; scipy.fft workers={_fft.WORKERS}")}


def cpu_baseline(scat, sysd, gpu_iters: float, single_iters: int):
    """Oracle (NumPy/SciPy FP64 restatement of the MATLAB path) on the host cores: one complete
    preconditioned BiCGStab solve to relative residual 1e-6 on every core of this process (scipy.fft
    workers), timed from the right-hand side to the solution (operator and preconditioner built
    beforehand, timed apart), plus ``single_iters`` iterations on one core, scaled to the same
    iteration count."""
    from oracle import circulant as Ch
    from oracle import krylov as OK
    from oracle import mvp as OM
    from oracle import operator as O
    from oracle import _fft

    L, M, N = scat.shape
    t0 = time.perf_counter()
    fref, Tref, alpha = O.getOPERATOR_DDA(scat.r, scat.k0, scat.refInd, sysd.kvec, scat.Eo)
    ones = np.ones((L, M, N))
    chi = np.zeros(L * M * N)
    chi[scat.idx] = 1
    chi = chi.reshape((L, M, N), order="F")
    _, cL = Ch.circ_1_level(Tref, L, M, N)
    _, blk = Ch.circ_2_level(cL, L, M, N)
    inv = Ch.invert_blocks_dda(blk, alpha, sysd.dV)
    t_build = time.perf_counter() - t0

    def afun(v):
        return OM.MVP(v, fref, ones, sysd.dV * chi, 1 / alpha, "notransp", sysd.idx3)

    def mfun(v):
        return Ch.mvp_circ_2_level(inv, v, L, M, N, sysd.idx3)

    workers = _fft.WORKERS
    t0 = time.perf_counter()
    _, flag, relres, it, _ = OK.bicgstab(afun, sysd.b, TOL, 2000, mfun)
    t_solve = time.perf_counter() - t0
    _fft.set_workers(1)
    try:
        t0 = time.perf_counter()
        OK.bicgstab(afun, sysd.b, 1e-30, single_iters, mfun)   # tolerance unreachable: exactly single_iters
        t1 = (time.perf_counter() - t0) / single_iters
    finally:
        _fft.set_workers(workers)
    return {
        "value": round(t_solve * 1e3, 1),
        "unit": "ms",
        "cores": workers,
        "kind": "port",
        "sample": (f"one complete FP64 BiCGStab + 2-level solve of the same Koch system by the NumPy/SciPy "
                   f"restatement of the MATLAB path (flag {flag}, {it:g} iterations against the GPU's "
                   f"{gpu_iters:g}, relres {relres:.2e}), scipy.fft workers={workers}; operator and "
                   f"preconditioner build {t_build:.1f} s apart"),
        "single_core_ms_per_iteration": round(t1 * 1e3, 1),
        "single_core_ms_extrapolated": round(t1 * float(it) * 1e3, 1),
        "single_core_sample": f"{single_iters} iterations on one core, scaled to {it:g} iterations",
        "iterations": float(it),
    }


def solution_errors(sysd, scat, x, x_ref, ctx):
    """Errors of a solution against a reference solution of the same system: the polarisation, the far
    field over a sphere quadrature (max over directions relative to the largest amplitude) and the
    extinction, absorption and scattering cross sections."""
    from voxscatter_amd import fields
    rv = np.asarray(scat.r).reshape(-1, 3, order="F")[np.asarray(scat.idx)]
    dirs, _ = fields.sphere_quadrature(24, 48)
    F = fields.farfield(rv, x, sysd.dV, scat.k0, dirs, ctx)
    Fr = fields.farfield(rv, x_ref, sysd.dV, scat.k0, dirs, ctx)
    cs = fields.cross_sections(x, sysd.b, 1 / sysd.alpha, sysd.dV, scat.k0, 1.0, ctx)
    csr = fields.cross_sections(x_ref, sysd.b, 1 / sysd.alpha, sysd.dV, scat.k0, 1.0, ctx)
    return {
        "solution": float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)),
        "farfield": float(np.max(np.abs(F - Fr)) / np.max(np.abs(Fr))),
        "C_ext": float(abs(cs[0] - csr[0]) / abs(csr[0])),
        "C_abs": float(abs(cs[1] - csr[1]) / abs(csr[1])),
        "C_sca": float(abs(cs[2] - csr[2]) / abs(csr[2])),
    }


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent particles/right-hand sides would shard across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    from voxscatter_amd import api, geometry
    from voxscatter_amd.capi import VS_FP32, VS_FP64
    from voxscatter_amd.dda import DDASystem

    ctx = api.get_context()
    if args.workload == "particles":
        particles(args, api, geometry, ctx)
        return
    if args.workload == "galerkin":
        galerkin(args, api, ctx)
        return
    scat = geometry.koch_prism(128, 128, 16, levels=3, kR=20.0, m=1.31 + 0.01j)
    L, M, N = scat.shape
    sysd = DDASystem.build(scat, precisions=VS_FP64 | VS_FP32, ctx=ctx)
    pc = sysd.preconditioner(2, precisions=VS_FP64 | VS_FP32)
    op = sysd.op
    n = op.n
    if args.matvec_only:
        matvec_only(op, pc, ctx, api, VS_FP32, VS_FP64)
        return
    b = api.DeviceArray.from_numpy(sysd.b, np.complex128, ctx)
    x = api.DeviceArray(n, np.complex128, ctx)

    def refined(inner=args.inner, inner_tol=args.inner_tol):
        return api.refine_dev(op, b.ptr, x.ptr, pc, tol=TOL, max_outer=10, inner=inner,
                              inner_tol=inner_tol, inner_maxit=2000, restart=50)

    if args.sweep:
        for inner in ("bicgstab", "gmres"):
            for itol in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6, TOL / 3):
                refined(inner, itol)
                ctx.synchronize()
                t0 = time.perf_counter()
                f_, rr_, o_, it_ = refined(inner, itol)
                ctx.synchronize()
                print(f"[sweep] inner={inner} inner_tol={itol:g}: {1e3 * (time.perf_counter() - t0):.2f} ms "
                      f"flag={f_} relres={rr_:.2e} outer={o_} inner_steps={it_}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        refined()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        flag, relres, outer, inner_total = refined()
        print(f"[bench] step {i}: flag={flag} relres={relres:.2e} outer={outer} inner={inner_total}",
              file=sys.stderr, flush=True)
    ctx.synchronize()
    t_ref = (time.perf_counter() - t0) / args.steps
    x_ref = x.numpy()
    assert flag == 0 and relres <= TOL, (flag, relres)

    # plain FP64 solve of the same system (timed beside: one warm run, then as many timed runs as steps)
    x64 = api.DeviceArray(n, np.complex128, ctx)
    api.bicgstab_dev(op, b.ptr, x64.ptr, TOL, 2000, pc)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f64, rr64, it64, nres64 = api.bicgstab_dev(op, b.ptr, x64.ptr, TOL, 2000, pc)
    ctx.synchronize()
    t_64 = (time.perf_counter() - t0) / args.steps
    print(f"[bench] fp64: flag={f64} iter={it64} {t_64 * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # plain FP64 GMRES(50) of the same system, to a true relative residual <= 1e-6: MATLAB's gmres stops
    # on the left-preconditioned residual, so the tolerance is tightened (untimed search) until the true
    # residual meets 1e-6, and that solve is timed
    xg = api.DeviceArray(n, np.complex128, ctx)
    nbv = np.linalg.norm(sysd.b)
    tol_g = TOL
    for _ in range(6):
        fg, rrg, itg, nresg = api.gmres_dev(op, b.ptr, xg.ptr, 50, tol_g, 40, pc)
        true_g = float(np.linalg.norm(op.mvp(xg.numpy()) - sysd.b) / nbv)
        if true_g <= TOL:
            break
        tol_g *= 0.5 * TOL / true_g
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fg, rrg, itg, nresg = api.gmres_dev(op, b.ptr, xg.ptr, 50, tol_g, 40, pc)
    ctx.synchronize()
    t_g64 = (time.perf_counter() - t0) / args.steps
    true_g = float(np.linalg.norm(op.mvp(xg.numpy()) - sysd.b) / nbv)
    print(f"[bench] fp64 gmres(50): flag={fg} iter={itg} tol={tol_g:.2e} true relres={true_g:.2e} "
          f"{t_g64 * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # the precision form with GMRES(50) as the inner solver (the FP64 outer residual is the true one)
    xgr = api.DeviceArray(n, np.complex128, ctx)
    api.refine_dev(op, b.ptr, xgr.ptr, pc, tol=TOL, max_outer=10, inner="gmres", inner_tol=args.inner_tol,
                   inner_maxit=2000, restart=50)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fgr, rrgr, ogr, itgr = api.refine_dev(op, b.ptr, xgr.ptr, pc, tol=TOL, max_outer=10, inner="gmres",
                                              inner_tol=args.inner_tol, inner_maxit=2000, restart=50)
    ctx.synchronize()
    t_gr = (time.perf_counter() - t0) / args.steps
    print(f"[bench] refined gmres(50): flag={fgr} relres={rrgr:.2e} outer={ogr} inner={itgr} {t_gr * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    x64n = x64.numpy()
    # true FP64 residual of both solutions, and their difference
    r_ref = np.linalg.norm(op.mvp(x_ref) - sysd.b) / np.linalg.norm(sysd.b)
    r_64 = np.linalg.norm(op.mvp(x64n) - sysd.b) / np.linalg.norm(sysd.b)
    sol_err = float(np.linalg.norm(x_ref - x64n) / np.linalg.norm(x64n))
    # errors of the fields and cross sections against a tightly converged FP64 solution (rel-res 1e-10)
    xt = api.DeviceArray(n, np.complex128, ctx)
    ft, rrt, itt, _ = api.bicgstab_dev(op, b.ptr, xt.ptr, 1e-10, 2000, pc)
    x_tight = xt.numpy()
    err_ref = solution_errors(sysd, scat, x_ref, x_tight, ctx)
    err_64 = solution_errors(sysd, scat, x64n, x_tight, ctx)
    print(f"[bench] errors vs FP64 rel-res 1e-10 solution: refined {err_ref}; FP64 at 1e-6 {err_64}",
          file=sys.stderr, flush=True)

    # matvec and preconditioner-apply timing (HIP events on the library stream)
    P = 8 * L * M * N
    V = L * M * N
    mv = {}
    rng = np.random.default_rng(0)
    xr = rng.normal(size=n) + 1j * rng.normal(size=n)
    # algorithmic bytes per matvec: s_c (6 P + 12 V) -- the padded 3-component spectrum written once and
    # read once (6 P), J read and y written (6 V), and the tensor read once; the tensor is exploited through
    # its octant parity, so its term is the octant, 6 (L+1)(M+1)(N+1) ~ 6 V, not 6 P.
    # Preconditioner apply: the stored blocks (mirror- and transpose-symmetric: (L/2+1)(M/2+1) packed
    # triangles of n(n+1)/2 entries, n = 3N) plus 18 V of vector traffic in the FFT stages.
    nb = 3 * N
    blk_stored = (L // 2 + 1) * (M // 2 + 1) * nb * (nb + 1) // 2
    for prec, dt, sc in ((VS_FP32, np.complex64, 8), (VS_FP64, np.complex128, 16)):
        xv = api.DeviceArray.from_numpy(xr, dt, ctx)
        yv = api.DeviceArray(n, dt, ctx)
        ms = api.bench_mvp(op, xv.ptr, yv.ptr, False, prec, 1, 50)
        ms_pc = api.bench_pc(pc, xv.ptr, yv.ptr, prec, 1, 50)
        B_mv = sc * (6 * P + 12 * V)
        B_pc = sc * (blk_stored + 18 * V)
        mv[prec] = dict(ms=ms, gbs=B_mv / (ms * 1e-3) / 1e9, bytes=B_mv, pc_ms=ms_pc,
                        pc_gbs=B_pc / (ms_pc * 1e-3) / 1e9)

    # batched form beside (no headline): 16 right-hand sides of the same prism (8 Fibonacci directions x 2
    # polarisations) refined together, ms per right-hand side
    from voxscatter_amd import geometry as geo
    rhs16, _ = geo.plane_wave_batch(scat, ndir=8)
    B16 = api.DeviceArray.from_numpy(np.ascontiguousarray(rhs16), np.complex128, ctx)
    X16 = api.DeviceArray(B16.shape, np.complex128, ctx)
    api.refine_batch_dev(op, B16, X16, pc, tol=TOL, inner_tol=args.inner_tol)
    ctx.synchronize()
    t0 = time.perf_counter()
    fb, rrb, ob, itb = api.refine_batch_dev(op, B16, X16, pc, tol=TOL, inner_tol=args.inner_tol)
    ctx.synchronize()
    t_b16 = time.perf_counter() - t0
    print(f"[bench] batch of 16: {t_b16 * 1e3:.1f} ms ({t_b16 * 1e3 / 16:.2f} ms per right-hand side)",
          file=sys.stderr, flush=True)

    m32 = mv[VS_FP32]
    traffic = measured_traffic(op.matvec_variant(VS_FP32))
    roof = {"bound": "hbm", "achieved": round(m32["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(m32["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"FP32 FFT matvec, one vs_mvp_dev call (variant '{op.matvec_variant(VS_FP32)}': "
                      "x FFT + gather, fused y-z slab kernel, inverse x FFT + diagonal + scatter)",
            "bytes_per_launch": m32["bytes"], "ms_per_launch": round(m32["ms"], 5)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scat, sysd, gpu_iters=float(it64), single_iters=args.cpu_iters)

    out = {
        "metric": "time-to-solution @rel-res 1e-6 (FP32+refine vs FP64) + matvec HBM GB/s",
        "value": round(t_ref * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_ref * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner / c128 refinement",
        "data": "synthetic",
        "config": {"workload": "koch_prism level 3, 128x128x16, kR=20, m=1.31+0.01i, DDA-LDR, "
                               "BiCGStab + 2-level Chan, rel-res 1e-6",
                   "unknowns": int(n), "grid": [L, M, N]},
        "fp64_ms": round(t_64 * 1e3, 3),
        "speedup_vs_fp64": round(t_64 / t_ref, 3),
        "fp64_bicgstab_iters": it64,
        "fp64_gmres50_ms": round(t_g64 * 1e3, 3),
        "fp64_gmres50_iters": list(itg),
        "fp64_gmres50_tol": tol_g,
        "fp64_gmres50_true_relres": true_g,
        "refined_gmres50_ms": round(t_gr * 1e3, 3),
        "refined_gmres50_relres": float(rrgr),
        "refined_gmres50_outer_inner": [ogr, itgr],
        "speedup_gmres50_vs_fp64": round(t_g64 / t_gr, 3),
        "inner_solver": args.inner,
        "inner_tol": args.inner_tol,
        "refine_outer": outer,
        "refine_inner_halfsteps": inner_total,
        "relres_refined_fp64": float(r_ref),
        "relres_fp64": float(r_64),
        "solution_rel_diff_vs_fp64": sol_err,
        "errors_vs_converged_fp64": {"reference": f"FP64 BiCGStab to rel-res 1e-10 ({itt:g} iterations)",
                                     "refined": err_ref, "fp64_at_1e-6": err_64},
        "matvec_fp32_ms": round(m32["ms"], 5),
        "matvec_fp32_gbs": round(m32["gbs"], 1),
        "matvec_fp64_ms": round(mv[VS_FP64]["ms"], 5),
        "matvec_fp64_gbs": round(mv[VS_FP64]["gbs"], 1),
        "matvec_variant_fp32": op.matvec_variant(VS_FP32),
        "matvec_variant_fp64": op.matvec_variant(VS_FP64),
        "pc_apply_fp32_ms": round(m32["pc_ms"], 5),
        "pc_apply_fp32_gbs": round(m32["pc_gbs"], 1),
        "pc_apply_fp64_ms": round(mv[VS_FP64]["pc_ms"], 5),
        "pc_apply_fp64_gbs": round(mv[VS_FP64]["pc_gbs"], 1),
        "batch16_refined_ms_per_rhs": round(t_b16 * 1e3 / 16, 3),
        "batch16_converged": int(np.sum(np.asarray(fb) == 0)),
        "batch16_max_relres": float(np.max(rrb)),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
