"""Operator "function handles" beyond the FFT operator (iterapp.m:41-69: A and M1 may be matrices or
handles), against the oracle's Krylov restatements on the same inputs:
* a dense matrix operand with a dense-matrix preconditioner handle: BiCGStab and GMRES with identical
  length(resvec) / iteration counts and solutions within 1e-10; the 'transp' matvec is A^H x;
* the material preconditioner of JVIE_Solver_const.m:232-234, mv_AN_const(J, fN, 1./Mr, 1./Mr - 1, Gram,
  ...), as the matvec of a second operator on the same fN, preconditioning the form-1 Galerkin system
  (Mr = eps, Mc = eps - 1): GMRES iteration count and solution against the oracle."""
import math

import numpy as np
import pytest

from oracle import krylov as OK
from oracle import mvp as OM
from voxscatter_amd import api

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_dense_operator_and_preconditioner_handles(ctx):
    rng = np.random.default_rng(21)
    n = 120
    A = np.eye(n) * (2 + 0.5j) + 0.6 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    Minv = np.linalg.inv(np.diag(np.diag(A)) + np.triu(A, 1) * 0.5)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    Ad = api.Operator.dense(A, ctx=ctx)
    Md = api.Preconditioner.of_operator(api.Operator.dense(Minv, ctx=ctx))
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert _rel(Ad.mvp(x), A @ x) <= 1e-14
    assert _rel(Ad.mvp(x, transp=True), A.conj().T @ x) <= 1e-14
    xg, fg, rg, ig, resg = api.bicgstab(Ad, b, 1e-10, 200, Md)
    xr, fr, rr, ir, resr = OK.bicgstab(lambda v: A @ v, b, 1e-10, 200, lambda v: Minv @ v)
    print(f"[handles] dense bicgstab gpu iter={ig} cpu iter={ir}", flush=True)
    assert fg == fr == 0 and ig == ir and len(resg) == len(resr)
    assert _rel(xg, xr) <= 1e-10 and _rel(xg, np.linalg.solve(A, b)) <= 1e-8
    xg, fg, rg, ig, resg = api.gmres(Ad, b, 20, 1e-10, 10, Md)
    xr, fr, rr, ir, resr = OK.gmres(lambda v: A @ v, b, 20, 1e-10, 10, lambda v: Minv @ v)
    print(f"[handles] dense gmres gpu iter={ig} cpu iter={ir}", flush=True)
    assert fg == fr == 0 and tuple(ig) == tuple(ir)
    assert _rel(xg, xr) <= 1e-10


def test_material_preconditioner_handle(ctx):
    L, M, N = 6, 5, 4
    ko = 2 * math.pi
    dx = 1 / 35
    rng = np.random.default_rng(4)
    eps = (1.8 + 0.6 * rng.uniform(size=(L, M, N))) * (1 + 0.01j)
    nD = L * M * N
    idx3 = np.arange(3 * nD)
    Gram = dx ** 3
    op = api.Operator.galerkin((L, M, N), dx, ko, ctx=ctx)
    fN = op.fN()
    Mr, Mc = eps, eps - 1.0
    op.set_domain(idx3, Gram, Mr.reshape(-1, order="F"), Mc.reshape(-1, order="F"))
    pop = api.Operator.from_fN(fN, ctx=ctx)
    pop.set_domain(idx3, Gram, (1 / Mr).reshape(-1, order="F"), (1 / Mr - 1.0).reshape(-1, order="F"))
    pc = api.Preconditioner.of_operator(pop)
    b = rng.normal(size=idx3.size) + 1j * rng.normal(size=idx3.size)

    def afun(v):
        return OM.MVP(v, fN, Mr, Mc, Gram, "notransp", idx3)

    def mfun(v):
        return OM.MVP(v, fN, 1 / Mr, 1 / Mr - 1.0, Gram, "notransp", idx3)

    assert _rel(op.mvp(b), afun(b)) <= 1e-12
    assert _rel(pc.apply(b), mfun(b)) <= 1e-12
    xg, fg, rg, ig, resg = api.gmres(op, b, 30, 1e-8, 10, pc)
    xr, fr, rr, ir, resr = OK.gmres(afun, b, 30, 1e-8, 10, mfun)
    print(f"[handles] material-preconditioned GMRES gpu iter={ig} cpu iter={ir}", flush=True)
    assert fg == fr == 0 and tuple(ig) == tuple(ir)
    assert _rel(xg, xr) <= 1e-10
    x0, f0, r0, i0, _ = api.gmres(op, b, 30, 1e-8, 10)
    print(f"[handles] unpreconditioned GMRES iter={i0}", flush=True)
