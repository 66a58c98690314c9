"""Oracle Krylov solvers in the form the reference runs them.

The reference calls ``jax.scipy.sparse.linalg.cg / bicgstab / gmres`` (third-party
JAX, version unpinned by requirements.txt:1-2) from
src/solvers/jax_linear_solver.py:282,305,329 with tol=1e-10, atol=0,
maxiter=100, restart=30 (config/settings.py:104-106) and an optional Jacobi
``M`` (jax_linear_solver.py:12-36).  JAX is absent here, so this file restates
the published algorithms of that module:

* cg:  preconditioned CG; stop when rs <= max(tol^2 |b|^2, atol^2) or k = maxiter,
  rs = r.z without M, r.r with M.
* bicgstab: preconditioned BiCGSTAB with r_hat = r0, rho = alpha = omega = 1,
  early exit when s.s < atol2, breakdown codes -10 (rho = 0) / -11
  (alpha or omega = 0) stop the loop.
* gmres ('batched'): left-preconditioned restarted GMRES; each cycle runs all
  ``restart`` Arnoldi steps unless breakdown; classical Gram-Schmidt (one
  pass, see ``_cgs``); least squares by normal equations + Cholesky; outer loop while k < maxiter and |M(b - A x)| > ptol.

Each function takes ``matvec`` (x -> A x) and ``dinv`` (Jacobi inverse diagonal,
a callable preconditioner, or None) and returns (x, info) where info carries
iteration counts.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

EPS = np.finfo(np.float64).eps


def _M(dinv):
    """Preconditioner: None (identity), a Jacobi inverse diagonal, or any
    callable x -> M x (oracle.precond)."""
    if dinv is None:
        return lambda v: v
    if callable(dinv):
        return dinv
    return lambda v: dinv * v


def cg(matvec, b, x0, tol=1e-10, atol=0.0, maxiter=100, dinv=None):
    M = _M(dinv)
    bs = float(b @ b)
    atol2 = max(tol * tol * bs, atol * atol)
    x = x0.copy()
    r = b - matvec(x)
    z = M(r)
    p = z.copy()
    gamma = float(r @ z)
    k = 0
    while True:
        rs = gamma if dinv is None else float(r @ r)
        if not (rs > atol2 and k < maxiter):
            break
        Ap = matvec(p)
        alpha = gamma / float(p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = M(r)
        gamma_ = float(r @ z)
        beta = gamma_ / gamma
        p = z + beta * p
        gamma = gamma_
        k += 1
    return x, {"iters": k}


def bicgstab(matvec, b, x0, tol=1e-10, atol=0.0, maxiter=100, dinv=None):
    M = _M(dinv)
    bs = float(b @ b)
    atol2 = max(tol * tol * bs, atol * atol)
    x = x0.copy()
    r = b - matvec(x)
    rhat = r.copy()
    alpha = omega = rho = 1.0
    p = r.copy()
    q = r.copy()
    k = 0
    steps = 0
    while True:
        rs = float(r @ r)
        if not (rs > atol2 and k < maxiter and k >= 0):
            break
        rho_ = float(rhat @ r)
        beta = rho_ / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = M(p)
        q = matvec(phat)
        alpha_ = rho_ / float(rhat @ q)
        s = r - alpha_ * q
        exit_early = float(s @ s) < atol2
        if exit_early:
            x = x + alpha_ * phat
            r = s
            omega_ = omega  # unused after an early exit
        else:
            shat = M(s)
            t = matvec(shat)
            omega_ = float(t @ s) / float(t @ t)
            x = x + (alpha_ * phat + omega_ * shat)
            r = s - omega_ * t
        kk = k + 1
        if (not exit_early and omega_ == 0) or alpha_ == 0:
            kk = -11
        if rho_ == 0:
            kk = -10
        k = kk
        alpha, omega, rho = alpha_, omega_, rho_
        steps += 1
    return x, {"iters": steps, "k": k}


def _safe_normalize(x, thresh=None):
    norm = float(np.sqrt(x @ x))
    if thresh is None:
        thresh = EPS
    if norm > thresh:
        return x / norm, norm
    return np.zeros_like(x), 0.0


def _cgs(V, ncols, w, wnorm):
    """Classical Gram-Schmidt of w against V[:, :ncols].

    JAX's ``_iterative_classical_gram_schmidt(max_iterations=2)`` runs one
    pass, then loops while ``k < max_iterations - 1`` with k = 1 already, so
    exactly one pass runs.  (Unpinned below roundoff: the reference's 3D GMRES
    trajectory agrees with one pass and with two passes alike.)
    """
    del wnorm
    h = V[:, :ncols].T @ w
    q = w - V[:, :ncols] @ h
    return q, h


def gmres(matvec, b, x0, tol=1e-10, atol=0.0, restart=30, maxiter=100, dinv=None):
    M = _M(dinv)
    n = b.size
    restart = min(restart, n)
    b_norm = float(np.sqrt(b @ b))
    atol_eff = max(tol * b_norm, atol)
    Mb = M(b)
    Mb_norm = float(np.sqrt(Mb @ Mb))
    ptol = Mb_norm * min(1.0, atol_eff / b_norm)
    x = x0.copy()
    res = M(b - matvec(x))
    unit, rnorm = _safe_normalize(res)
    k = 0
    arnoldi = 0
    while k < maxiter and rnorm > ptol:
        V = np.zeros((n, restart + 1))
        V[:, 0] = unit
        H = np.eye(restart, restart + 1)
        j = 0
        breakdown = False
        while j < restart and not breakdown:
            v = M(matvec(V[:, j]))
            _, v0 = _safe_normalize(v)
            v, h = _cgs(V, j + 1, v, v0)
            unit_v, v1 = _safe_normalize(v, thresh=EPS * v0)
            V[:, j + 1] = unit_v
            H[j, :] = 0.0
            H[j, : j + 1] = h
            H[j, j + 1] = v1
            breakdown = v1 == 0.0
            j += 1
            arnoldi += 1
        beta = np.zeros(restart + 1)
        beta[0] = rnorm
        a = H.T
        a2 = a.T @ a
        b2 = a.T @ beta
        c = sla.cho_factor(a2, lower=True)
        y = sla.cho_solve(c, b2)
        x = x + V[:, :-1] @ y
        res = M(b - matvec(x))
        unit, rnorm = _safe_normalize(res)
        k += 1
    info = -1 if np.isnan(np.sqrt(x @ x)) else 0
    return x, {"iters": k, "arnoldi": arnoldi, "info": info}


def solve(method, matvec, b, x0, dinv=None, tol=1e-10, atol=0.0, maxiter=100, restart=30):
    if method == "cg":
        return cg(matvec, b, x0, tol, atol, maxiter, dinv)
    if method == "bicgstab":
        return bicgstab(matvec, b, x0, tol, atol, maxiter, dinv)
    if method == "gmres":
        return gmres(matvec, b, x0, tol, atol, restart, maxiter, dinv)
    raise ValueError(f"Unsupported solver type: {method}")
