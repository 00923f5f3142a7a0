"""Krylov solvers, FP64 (TEST INFRASTRUCTURE: see oracle/__init__.py).

* ``bicgstab``  MATLAB built-in ``bicgstab`` as called at main_DDA.m:82,120 (not vendored in the
                reference).  Restated from MATLAB's documented behaviour: x0 = 0 by default; the
                preconditioner handle returns M\\x; convergence norm(b-A*x) <= tol*norm(b) is tested
                at every half step, confirmed against the explicitly recomputed residual; resvec holds
                the residual norm at x0 and at every half step, so length(resvec) is the "iteration
                count" main_DDA.m:84 prints; flags 0 (converged), 1 (maxit), 2 (ill-conditioned
                preconditioner), 3 (stagnation), 4 (breakdown); on failure the iterate with the
                smallest residual is returned.
* ``gmres``     MATLAB built-in ``gmres(A,b,restart,tol,maxit,M)`` (main_DDA.m:89,128): restarted
                GMRES, LEFT preconditioned, convergence on the preconditioned residual
                norm(M\\(b-Ax)) <= tol*norm(M\\b).  MATLAB orthogonalises with Householder
                reflections; this restatement (and the GPU) uses classical Gram-Schmidt applied twice,
                which spans the same Krylov space and gives the same iterates in exact arithmetic.
* ``pgmres``    the reference's own GMRES(m), src/src_numeric/src_iterative/pgmres.m:70-180 and
                pgmres_iter.m:22-150, followed step by step (one-pass CGS, backslash least squares).

Operators are callables ``afun(x) -> A x`` and ``mfun(x) -> M^{-1} x`` (iterapp.m:41-69).
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps


def _norm(x):
    return float(np.linalg.norm(x))


def bicgstab(afun, b, tol=1e-6, maxit=None, mfun=None, x0=None):
    """Returns ``(x, flag, relres, iter, resvec)`` with MATLAB semantics (see module docstring)."""
    b = np.asarray(b, dtype=np.complex128)
    n = b.shape[0]
    if maxit is None:
        maxit = min(n, 20)
    n2b = _norm(b)
    if n2b == 0:
        return np.zeros(n, dtype=np.complex128), 0, 0.0, 0.0, np.array([0.0])
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    flag = 1
    xmin = x.copy()
    imin = 0.0
    tolb = tol * n2b
    r = b - afun(x)
    normr = _norm(r)
    normr_act = normr
    if normr <= tolb:
        return x, 0, normr / n2b, 0.0, np.array([normr])
    rt = r.copy()
    resvec = np.zeros(2 * maxit + 1)
    resvec[0] = normr
    stag = 0
    moresteps = 0
    maxmsteps = min(n // 50, 10, n - maxit)
    maxstagsteps = 3
    normrmin = normr
    rho = 1.0 + 0j
    omega = 1.0 + 0j
    alpha = 0.0 + 0j
    p = v = None
    nres = None
    ii = 0
    for ii in range(1, maxit + 1):
        rho1 = rho
        rho = np.vdot(rt, r)
        if rho == 0.0 or np.isinf(rho):
            flag = 4
            nres = 2 * ii - 1
            break
        if ii == 1:
            p = r.copy()
        else:
            beta = (rho / rho1) * (alpha / omega)
            if beta == 0 or not np.isfinite(beta):
                flag = 4
                nres = 2 * ii - 1
                break
            p = r + beta * (p - omega * v)
        ph = mfun(p) if mfun is not None else p
        if not np.all(np.isfinite(ph)):
            flag = 2
            nres = 2 * ii - 1
            break
        v = afun(ph)
        rtv = np.vdot(rt, v)
        if rtv == 0 or np.isinf(rtv):
            flag = 4
            nres = 2 * ii - 1
            break
        alpha = rho / rtv
        if np.isinf(alpha):
            flag = 4
            nres = 2 * ii - 1
            break
        if abs(alpha) * _norm(ph) < EPS * _norm(x):
            stag += 1
        else:
            stag = 0
        xhalf = x + alpha * ph
        s = r - alpha * v
        normr = _norm(s)
        normr_act = normr
        resvec[2 * ii - 1] = normr
        if normr <= tolb or stag >= maxstagsteps or moresteps:
            s = b - afun(xhalf)
            normr_act = _norm(s)
            resvec[2 * ii - 1] = normr_act
            if normr_act <= tolb:
                x = xhalf
                flag = 0
                iter_ = ii - 0.5
                nres = 2 * ii
                break
            else:
                if stag >= maxstagsteps and moresteps == 0:
                    stag = 0
                moresteps += 1
                if moresteps >= maxmsteps:
                    flag = 3
                    x = xhalf
                    nres = 2 * ii
                    break
        if normr_act < normrmin:
            normrmin = normr_act
            xmin = xhalf.copy()
            imin = ii - 0.5
        sh = mfun(s) if mfun is not None else s
        if not np.all(np.isfinite(sh)):
            flag = 2
            nres = 2 * ii
            break
        t = afun(sh)
        tt = np.vdot(t, t)
        if tt == 0 or np.isinf(tt):
            flag = 4
            nres = 2 * ii
            break
        omega = np.vdot(t, s) / tt
        if np.isinf(omega):
            flag = 4
            nres = 2 * ii
            break
        if abs(omega) * _norm(sh) < EPS * _norm(xhalf):
            stag += 1
        else:
            stag = 0
        x = xhalf + omega * sh
        r = s - omega * t
        normr = _norm(r)
        normr_act = normr
        resvec[2 * ii] = normr
        if normr <= tolb or stag >= maxstagsteps or moresteps:
            r = b - afun(x)
            normr_act = _norm(r)
            resvec[2 * ii] = normr_act
            if normr_act <= tolb:
                flag = 0
                iter_ = float(ii)
                nres = 2 * ii + 1
                break
            else:
                if stag >= maxstagsteps and moresteps == 0:
                    stag = 0
                moresteps += 1
                if moresteps >= maxmsteps:
                    flag = 3
                    nres = 2 * ii + 1
                    break
        if normr_act < normrmin:
            normrmin = normr_act
            xmin = x.copy()
            imin = float(ii)
        if stag >= maxstagsteps:
            flag = 3
            nres = 2 * ii + 1
            break
    if nres is None:
        nres = 2 * maxit + 1
    if flag == 0:
        relres = normr_act / n2b
    else:
        r_comp = b - afun(xmin)
        if _norm(r_comp) <= normr_act:
            x = xmin
            iter_ = imin
            relres = _norm(r_comp) / n2b
        else:
            iter_ = float(ii)
            relres = normr_act / n2b
    return x, flag, relres, iter_, resvec[:nres].copy()


def gmres(afun, b, restart=None, tol=1e-6, maxit=None, mfun=None, x0=None):
    """Left-preconditioned restarted GMRES; returns ``(x, flag, relres, iter, resvec)``.

    ``iter`` is ``(outer, inner)`` as in MATLAB; ``resvec`` holds preconditioned residual norms
    (initial one included).  relres is the preconditioned relative residual of the returned x.
    """
    b = np.asarray(b, dtype=np.complex128)
    n = b.shape[0]
    if restart is None or restart > n:
        restart = n
    if maxit is None:
        maxit = min(math.ceil(n / restart), 10)
    M = (lambda v: v) if mfun is None else mfun
    n2b = _norm(b)
    if n2b == 0:
        return np.zeros(n, dtype=np.complex128), 0, 0.0, (0, 0), np.array([0.0])
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    n2minv_b = _norm(M(b))
    tolb = tol * n2minv_b
    r = M(b - afun(x))
    normr = _norm(r)
    resvec = [normr]
    if normr <= tolb:
        return x, 0, normr / n2minv_b, (0, 0), np.array(resvec)
    flag = 1
    outer = inner = 0
    for outer in range(1, maxit + 1):
        V = np.zeros((n, restart + 1), dtype=np.complex128)
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        cs = np.zeros(restart, dtype=np.complex128)
        sn = np.zeros(restart, dtype=np.complex128)
        g = np.zeros(restart + 1, dtype=np.complex128)
        beta = normr
        V[:, 0] = r / beta
        g[0] = beta
        k_done = 0
        for k in range(restart):
            w = M(afun(V[:, k]))
            h = V[:, :k + 1].conj().T @ w
            w = w - V[:, :k + 1] @ h
            h2 = V[:, :k + 1].conj().T @ w
            w = w - V[:, :k + 1] @ h2
            h = h + h2
            hn = _norm(w)
            H[:k + 1, k] = h
            H[k + 1, k] = hn
            if hn != 0:
                V[:, k + 1] = w / hn
            for i in range(k):
                t1 = np.conj(cs[i]) * H[i, k] + np.conj(sn[i]) * H[i + 1, k]
                t2 = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k], H[i + 1, k] = t1, t2
            a, bb = H[k, k], H[k + 1, k]
            den = math.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            if den == 0:
                cs[k], sn[k] = 1.0, 0.0
            else:
                cs[k] = a / den
                sn[k] = bb / den
            H[k, k] = np.conj(cs[k]) * a + np.conj(sn[k]) * bb
            H[k + 1, k] = 0
            g[k + 1] = -sn[k] * g[k]
            g[k] = np.conj(cs[k]) * g[k]
            inner = k + 1
            k_done = k + 1
            resvec.append(abs(g[k + 1]))
            if abs(g[k + 1]) <= tolb or hn == 0:
                break
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), g[:k_done])
        x = x + V[:, :k_done] @ y
        r = M(b - afun(x))
        normr = _norm(r)
        resvec[-1] = normr
        if normr <= tolb:
            flag = 0
            break
    return x, flag, normr / n2minv_b, (outer, inner), np.array(resvec)


def pgmres(afun, b, restart=50, tol=1e-3, maxit=10, L1=None, L2=None, R1=None, R2=None, x0=None):
    """``[x,flag,relres,iter,resvec] = pgmres(A,b,restart,tol,maxit,L1,L2,R1,R2,x0)`` (pgmres.m:70-180)."""
    b = np.asarray(b, dtype=np.complex128)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    r = b - afun(x) if np.count_nonzero(x) else b.copy()
    bp = b.copy()
    if L1 is not None:
        r = L1(r)
        bp = L1(bp)
    if L2 is not None:
        r = L2(r)
        bp = L2(bp)
    bnorm = _norm(bp)
    resvec = [_norm(r) / bnorm]
    outit = 0
    p = 0
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        x, r, p, res_inner = pgmres_iter(afun, x, r, restart, L1, L2, R1, R2, tol * bnorm)
        resvec.extend(list(np.asarray(res_inner) / bnorm))
    relres = _norm(b - afun(x)) / _norm(b)
    flag = 0 if relres <= tol else 1
    return x, flag, relres, (p, outit), np.array(resvec)


def pgmres_iter(afun, x, r, m, L1, L2, R1, R2, tol):
    """One GMRES(m) cycle of pgmres_iter.m:22-150."""
    n = r.shape[0]
    V = np.zeros((n, m + 1), dtype=np.complex128)
    V[:, 0] = r / _norm(r)
    H = np.zeros((m + 1, m), dtype=np.complex128)
    resvec = []
    y = res = None
    k = 0
    rn = _norm(r)
    for k in range(1, m + 1):
        w = V[:, k - 1]
        if R1 is not None:
            w = R1(w)
        if R2 is not None:
            w = R2(w)
        w = afun(w)
        if L1 is not None:
            w = L1(w)
        if L2 is not None:
            w = L2(w)
        H[:k, k - 1] = V[:, :k].conj().T @ w
        w = w - V[:, :k] @ H[:k, k - 1]
        H[k, k - 1] = _norm(w)
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, dtype=np.complex128)
        rhs[0] = rn
        Hc = H[:k + 1, :k]
        y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]
        res = rhs - Hc @ y
        resvec.append(_norm(res))
        if resvec[-1] < tol:
            break
    xupd = V[:, :k] @ y
    if R1 is not None:
        xupd = R1(xupd)
    if R2 is not None:
        xupd = R2(xupd)
    x = x + xupd
    r = V[:, :k + 1] @ res
    return x, r, k, resvec


# ------------------------------------------------------------------------------------------------
# GMRES with deflated restarting (pgmresDR.m, pgmresDR_iter.m, getHarmVecs1.m, getHarmVecs2.m)
# ------------------------------------------------------------------------------------------------
def _apply_side(afun, v, L1, R1):
    if R1 is not None:
        v = R1(v)
    v = afun(v)
    if L1 is not None:
        v = L1(v)
    return v


def pgmresDR_iter(afun, r, m, L1, R1, C, tol):
    """One Arnoldi cycle of pgmresDR_iter.m:23-141 (classical Gram-Schmidt after the deflation against C);
    returns V (n x k+1), H (k+1 x k), B (size(C,2) x k), k, resvec (projected residual norms)."""
    n = r.shape[0]
    V = np.zeros((n, m + 1), dtype=np.complex128)
    rn = _norm(r)
    V[:, 0] = r / rn
    kc = 0 if C is None else C.shape[1]
    B = np.zeros((kc, m), dtype=np.complex128)
    H = np.zeros((m + 1, m), dtype=np.complex128)
    resvec = np.zeros(m)
    for k in range(1, m + 1):
        w = _apply_side(afun, V[:, k - 1], L1, R1)
        if kc:
            B[:, k - 1] = C.conj().T @ w
            w = w - C @ B[:, k - 1]
        H[:k, k - 1] = V[:, :k].conj().T @ w
        w = w - V[:, :k] @ H[:k, k - 1]
        H[k, k - 1] = _norm(w)
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, dtype=np.complex128)
        rhs[0] = rn
        y = np.linalg.lstsq(H[:k + 1, :k], rhs, rcond=None)[0]
        resvec[k - 1] = _norm(rhs - H[:k + 1, :k] @ y)
        if resvec[k - 1] < tol:
            return V[:, :k + 1], H[:k + 1, :k], B[:, :k], k, resvec[:k]
    return V, H, B, m, resvec


def getHarmVecs1(m, k, H):
    """getHarmVecs1.m:14-29: harmonic Ritz vectors of H (m+1 x m) with the k smallest |theta|."""
    F = np.linalg.solve(H[:m, :].conj().T, np.eye(m))
    F[:, :m - 1] = 0
    Hr = H[:m, :] + H[m, m - 1] ** 2 * F
    vals, vecs = np.linalg.eig(Hr)
    order = np.argsort(np.abs(vals), kind="stable")
    return vecs[:, order[:k]]


def getHarmVecs2(m, k, H2, V, U, C):
    """getHarmVecs2.m:15-36: generalized harmonic Ritz vectors, the k of largest |lambda| of eig(A, B)."""
    import scipy.linalg
    B = H2.conj().T @ H2
    A = np.zeros((m + 1, m), dtype=np.complex128)
    A[:k, :k] = C.conj().T @ U
    A[k:m + 1, :k] = V.conj().T @ U
    A[k:m, k:m] = np.eye(m - k)
    A = H2.conj().T @ A
    vals, vecs = scipy.linalg.eig(A, B)
    order = np.argsort(np.abs(vals), kind="stable")
    return vecs[:, order[m - k:]]


def pgmresDR(afun, b, restart=50, tol=1e-3, maxit=10, ritz=None, L1=None, R1=None, x0=None, U0=None,
             return_U=False):
    """``[x,flag,relres,iter,resvec] = pgmresDR(A,b,restart,tol,maxit,ritz,L1,[],R1,[],x0)``
    (pgmresDR.m:73-280; one left factor L1 or right factor R1).

    Recycling across solves (no counterpart in pgmresDR.m; GCRO-DR of Parks et al., SIAM J. Sci. Comput.
    28 (2006), Alg. 2 lines 2-6): with ``U0`` (n x ritz, a deflation space kept from an earlier solve with
    the same operator and preconditioner) the first plain cycle is replaced by [C, R] = qr(A U0),
    U = U0 R^-1, x = U C^H r, r = r - C C^H r (one more resvec entry), and the deflated cycles follow.
    ``return_U``: also return the deflation space at the end (None when no deflated cycle ran)."""
    b = np.asarray(b, dtype=np.complex128)
    ritz = min(10, int(math.floor(restart / 2 + 0.5))) if ritz is None else min(ritz, restart - 1)
    x0 = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    r = b - afun(x0)
    bp = b.copy()
    if L1 is not None:
        r = L1(r)
        bp = L1(bp)
    bnorm = _norm(bp)
    resvec = [_norm(r) / bnorm]
    p = 0
    if U0 is not None:
        U0 = np.asarray(U0, dtype=np.complex128)
        if U0.shape != (b.size, ritz):
            raise ValueError("the recycled space must hold ritz vectors of the system size")
        W = np.stack([_apply_side(afun, U0[:, i], L1, R1) for i in range(ritz)], axis=1)
        C, R = np.linalg.qr(W)
        U = np.linalg.solve(R.T, U0.T).T
        cr = C.conj().T @ r
        x = U @ cr
        r = r - C @ cr
        resvec.append(_norm(r) / bnorm)
    else:
        V, H, _, p, rin = pgmresDR_iter(afun, r, restart, L1, R1, None, tol * bnorm)
        resvec += list(rin / bnorm)
        y = np.linalg.lstsq(H, V.conj().T @ r, rcond=None)[0]
        x = V[:, :p] @ y
        r = r - V @ (H @ y)
        U = C = None
        if resvec[-1] > tol:
            k = p - 1 if ritz > p - 1 else ritz
            P = getHarmVecs1(p, k, H)
            U = V[:, :p] @ P
            C, R = np.linalg.qr(H @ P)
            C = V @ C
            U = np.linalg.solve(R.T, U.T).T
    outit = 0
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        V, H, B, p, rin = pgmresDR_iter(afun, r, restart, L1, R1, C, tol * bnorm)
        resvec += list(rin / bnorm)
        d = np.array([_norm(U[:, i]) for i in range(ritz)])
        D = np.diag(1.0 / d)
        U = U @ D
        H2 = np.block([[D, B], [np.zeros((H.shape[0], D.shape[1])), H]])
        CV = np.hstack([C, V])
        y = np.linalg.lstsq(H2, CV.conj().T @ r, rcond=None)[0]
        x = x + np.hstack([U, V[:, :p]]) @ y
        r = r - CV @ (H2 @ y)
        if resvec[-1] > tol:
            P = getHarmVecs2(p + ritz, ritz, H2, V, U, C)
            U = np.hstack([U, V[:, :p]]) @ P
            Q, R = np.linalg.qr(H2 @ P)
            C = CV @ Q
            U = np.linalg.solve(R.T, U.T).T
    if R1 is not None:
        x = R1(x)
    x = x0 + x
    relres = _norm(b - afun(x)) / _norm(b)
    flag = 0 if relres <= tol else 1
    if return_U:
        return x, flag, relres, (p, outit), np.array(resvec), (U if outit > 0 or U0 is not None else None)
    return x, flag, relres, (p, outit), np.array(resvec)
