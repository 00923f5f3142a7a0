"""Known-answer tests of the oracle's Krylov solvers (MATLAB bicgstab/gmres semantics, pgmres.m):
exact-arithmetic termination counts, numpy.linalg.solve, documented flags and resvec structure."""
import numpy as np
import pytest

from oracle import krylov as K


def _rand_system(n, seed, shift=4.0):
    rng = np.random.default_rng(seed)
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) + shift * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    return A, b


def test_bicgstab_identity_converges_at_first_half_step():
    b = np.arange(1, 6) + 1j
    x, flag, relres, it, resvec = K.bicgstab(lambda v: v, b, 1e-10, 10)
    assert flag == 0 and it == 0.5 and len(resvec) == 2
    assert np.allclose(x, b) and resvec[0] == pytest.approx(np.linalg.norm(b))


def test_bicgstab_exact_preconditioner():
    A, b = _rand_system(30, 0)
    Ainv = np.linalg.inv(A)
    x, flag, relres, it, resvec = K.bicgstab(lambda v: A @ v, b, 1e-10, 10, mfun=lambda v: Ainv @ v)
    assert flag == 0 and it == 0.5
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9)


def test_bicgstab_matches_direct_solve():
    A, b = _rand_system(60, 1)
    x, flag, relres, it, resvec = K.bicgstab(lambda v: A @ v, b, 1e-12, 200)
    assert flag == 0 and relres <= 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-10 * np.linalg.norm(x)
    # resvec: x0 plus one entry per half step, the last one the true residual
    assert len(resvec) == int(2 * it) + 1
    assert resvec[-1] == pytest.approx(np.linalg.norm(b - A @ x), rel=1e-6)


def test_bicgstab_two_distinct_eigenvalues_terminates():
    d = np.array([1.0] * 10 + [3.0] * 10, dtype=complex)
    b = np.linspace(1, 2, 20) + 0.5j
    x, flag, relres, it, resvec = K.bicgstab(lambda v: d * v, b, 1e-12, 50)
    assert flag == 0 and it <= 2
    assert np.allclose(x, b / d)


def test_bicgstab_maxit_flag_and_zero_rhs():
    A, b = _rand_system(40, 2, shift=0.3)
    x, flag, relres, it, resvec = K.bicgstab(lambda v: A @ v, b, 1e-14, 2)
    assert flag == 1 and len(resvec) == 5
    # the returned iterate is the minimal-residual one
    assert relres == pytest.approx(np.linalg.norm(b - A @ x) / np.linalg.norm(b), rel=1e-10)
    x, flag, relres, it, resvec = K.bicgstab(lambda v: A @ v, np.zeros(40), 1e-6, 5)
    assert flag == 0 and np.all(x == 0) and relres == 0 and list(resvec) == [0.0]


def test_gmres_low_rank_update_terminates():
    n, r = 50, 3
    rng = np.random.default_rng(7)
    U = rng.normal(size=(n, r)) + 1j * rng.normal(size=(n, r))
    Vv = rng.normal(size=(n, r)) + 1j * rng.normal(size=(n, r))
    A = np.eye(n) + U @ Vv.conj().T / n
    b = rng.normal(size=n) + 0j
    x, flag, relres, (outer, inner), resvec = K.gmres(lambda v: A @ v, b, n, 1e-12, 1)
    assert flag == 0 and inner <= r + 1
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_gmres_left_preconditioned_restarted():
    A, b = _rand_system(60, 3, shift=2.0)
    D = np.diag(A).copy()
    x, flag, relres, it, resvec = K.gmres(lambda v: A @ v, b, 10, 1e-11, 20, mfun=lambda v: v / D)
    assert flag == 0
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-9 * np.linalg.norm(x)
    # relres is the preconditioned relative residual
    assert relres == pytest.approx(np.linalg.norm((b - A @ x) / D) / np.linalg.norm(b / D), rel=1e-3, abs=1e-14)


@pytest.mark.parametrize("side", ["L1", "R1", None])
def test_pgmres_matches_direct_solve(side):
    A, b = _rand_system(60, 4, shift=2.0)
    D = np.diag(A).copy()
    pre = (lambda v: v / D)
    kw = {} if side is None else {side: pre}
    x, flag, relres, (p, outit), resvec = K.pgmres(lambda v: A @ v, b, 15, 1e-11, 30, **kw)
    assert flag == 0 and relres <= 1e-11
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-9 * np.linalg.norm(x)
    assert resvec[0] == pytest.approx(1.0)
    assert np.all(np.diff(resvec[:p + 1]) <= 1e-12)  # GMRES residuals are monotone within a cycle


def _outlier_system(n=120, seed=0):
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))[0]
    ev = np.concatenate([[1e-3, 2e-3, 5e-3, 1e-2], 1 + 0.5 * rng.uniform(size=n - 4)])
    A = Q @ np.diag(ev) @ Q.conj().T + 0.01 * rng.normal(size=(n, n))
    b = rng.normal(size=n) + 0j
    return A, b


def test_pgmresdr_solves_and_first_cycle_is_pgmres():
    A, b = _outlier_system()
    af = lambda v: A @ v  # noqa: E731
    xs = np.linalg.solve(A, b)
    x, flag, relres, it, rv = K.pgmresDR(af, b, 20, 1e-10, 30, 6)
    assert flag == 0 and np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
    # without deflated cycles (maxit = 0) the method is the first GMRES(m) cycle of pgmres
    x1, f1, r1, i1, rv1 = K.pgmresDR(af, b, 20, 1e-10, 0, 6)
    xp, fp, rp, ip, rvp = K.pgmres(af, b, 20, 1e-10, 30)
    assert np.allclose(rv1, rvp[:len(rv1)], rtol=1e-12)
    # deflating the four small eigenvalues beats plain restarting
    assert len(rv) < len(rvp)


def test_pgmresdr_preconditioned_sides():
    A, b = _outlier_system(60, 1)
    Minv = np.linalg.inv(np.diag(np.diag(A)))
    af = lambda v: A @ v  # noqa: E731
    mf = lambda v: Minv @ v  # noqa: E731
    xs = np.linalg.solve(A, b)
    for kw in ({"L1": mf}, {"R1": mf}):
        x, flag, relres, it, rv = K.pgmresDR(af, b, 15, 1e-10, 20, 4, **kw)
        assert flag == 0 and np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_pgmresdr_recycling():
    """Recycling the deflation space between solves (GCRO-DR): a space containing the solution solves
    the system by the projection alone; a kept space from one right-hand side cuts the iterations of
    the next; left and right preconditioning give the dense solution."""
    A, b = _outlier_system()
    af = lambda v: A @ v  # noqa: E731
    n = b.size
    rng = np.random.default_rng(4)
    xs = np.linalg.solve(A, b)
    U0 = np.column_stack([xs] + [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(5)])
    x, flag, relres, it, rv, U = K.pgmresDR(af, b, 20, 1e-10, 30, 6, U0=U0, return_U=True)
    assert flag == 0 and tuple(it) == (0, 0) and len(rv) == 2 and rv[1] < 1e-12
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12
    # a second right-hand side with the first solve's space
    b2 = rng.normal(size=n) + 1j * rng.normal(size=n)
    x1, f1, r1, i1, rv1, Ukeep = K.pgmresDR(af, b, 20, 1e-10, 30, 6, return_U=True)
    assert f1 == 0 and Ukeep.shape == (n, 6)
    x2, f2, r2, i2, rv2 = K.pgmresDR(af, b2, 20, 1e-10, 30, 6)
    x3, f3, r3, i3, rv3 = K.pgmresDR(af, b2, 20, 1e-10, 30, 6, U0=Ukeep)
    xs2 = np.linalg.solve(A, b2)
    assert f2 == f3 == 0 and len(rv3) < len(rv2)
    assert np.linalg.norm(x3 - xs2) / np.linalg.norm(xs2) < 1e-9
    Minv = np.linalg.inv(np.diag(np.diag(A)))
    mf = lambda v: Minv @ v  # noqa: E731
    for kw in ({"L1": mf}, {"R1": mf}):
        _, _, _, _, _, Uk = K.pgmresDR(af, b, 15, 1e-10, 20, 4, return_U=True, **kw)
        x4, f4, _, _, _ = K.pgmresDR(af, b2, 15, 1e-10, 20, 4, U0=Uk, **kw)
        assert f4 == 0 and np.linalg.norm(x4 - xs2) / np.linalg.norm(xs2) < 1e-8
    with pytest.raises(ValueError):
        K.pgmresDR(af, b, 20, 1e-10, 30, 6, U0=U0[:, :3])
