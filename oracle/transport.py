"""Oracle solute / nitrogen transport operators and their time loops (FP64).

TEST INFRASTRUCTURE ONLY (see the package docstring).

Follows, from the reference's source text:

* flux and dispersion recovery: src/models/transport_utilities.py:119-192
  (node value = unweighted mean of the incident elements' gradients, each
  element gradient in the b/c form with |area|; van Genuchten K and theta at
  the node; tau = theta^(7/3) / theta_s^2);
* element operators: src/numerics/assembly_solute.py:10-85 and
  src/numerics/assembly_nitrogen.py:33-307, evaluated as the literal 3-point
  quadrature loop (per point: N(q), J, inv(J), det(J) -- signed), so this file
  does not share the quadrature tensors the HIP kernel uses;
* global sums: assembly_solute.py:87-186, assembly_nitrogen.py:310-607 (COO
  with duplicates summed), lumping nitrogen_solver.py:36-66;
* edge terms: src/models/boundary.py:108-173 (Neumann, free drainage of
  water), :217-277 (Cauchy), :336-421 (free drainage of nitrogen), each on
  consecutive pairs of the boundary node list in list order, with
  length x_J - x_I and the 2-point Gauss rule (+-1/sqrt(3), weights 1);
* time loops: src/solvers/nitrogen_solver.py:100-413 and
  src/solvers/coupled_solver_v2.py:80-331.

Parity unpinned against the reference (it commits no numeric output of these
paths).  Pinned by the tests in tests/test_oracle_transport.py: exact
gradients of linear fields, K 1 = 0 for the transport stiffness, mass rows
against exact integrals for constant theta, dispersion/advection against an
independent high-order rule, edge-term totals against closed forms, and a
constant concentration staying constant without sources.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import krylov
from .assembly import TRI_Q, TRI_W, _ref_deriv, _shape
from .soil import van_genuchten

KSI_1D = np.array([-1.0 / np.sqrt(3.0), 1.0 / np.sqrt(3.0)])
W_1D = np.array([1.0, 1.0])


# ----------------------------------------------------------------------------
# flux / dispersion recovery (transport_utilities.py:119-192)
# ----------------------------------------------------------------------------
def element_gradients(points, tris, h):
    X = points[tris]
    x0, y0 = X[:, 0, 0], X[:, 0, 1]
    x1, y1 = X[:, 1, 0], X[:, 1, 1]
    x2, y2 = X[:, 2, 0], X[:, 2, 1]
    b1, b2, b3 = y1 - y2, y2 - y0, y0 - y1
    c1, c2, c3 = x2 - x1, x0 - x2, x1 - x0
    aire = np.abs(c3 * b2 - c2 * b3) / 2
    he = h[tris]
    dhx = (b1 * he[:, 0] + b2 * he[:, 1] + b3 * he[:, 2]) / (2 * aire)
    dhz = (c1 * he[:, 0] + c2 * he[:, 1] + c3 * he[:, 2]) / (2 * aire)
    return dhx, dhz


def fluxes_dispersion(points, tris, h, phi, DL, DT, Dm, ks_elem=None):
    """(qx, qz, |q|, Dxx, Dxz, Dzz, theta) at every node.  ks_elem: optional
    per-element Ks (an extension): K = K(h) * mean Ks of the node's elements."""
    n = points.shape[0]
    dhx, dhz = element_gradients(points, tris, h)
    sx = np.zeros(n)
    sz = np.zeros(n)
    sk = np.zeros(n)
    cnt = np.zeros(n)
    for e in range(tris.shape[0]):          # element order, as the masked sum
        for a in range(3):
            i = tris[e, a]
            sx[i] += dhx[e]
            sz[i] += dhz[e]
            if ks_elem is not None:
                sk[i] += ks_elem[e]
            cnt[i] += 1.0
    gx, gz = sx / cnt, sz / cnt
    _, K, theta = van_genuchten(h, phi)
    if ks_elem is not None:
        K = K * (sk / cnt)
    qx = -K * gx
    qz = -K * (gz + 1.0)
    aq = np.sqrt(qx ** 2 + qz ** 2)
    tau = theta ** (7.0 / 3.0) / phi[1] ** 2
    Dxx = (DL * qx ** 2 + DT * qz ** 2) / aq + theta * Dm * tau
    Dzz = (DL * qz ** 2 + DT * qx ** 2) / aq + theta * Dm * tau
    Dxz = (DL - DT) * qx * qz / aq
    return qx, qz, aq, Dxx, Dxz, Dzz, theta


# ----------------------------------------------------------------------------
# element operators (assembly_solute.py:10-85, assembly_nitrogen.py:33-307)
# ----------------------------------------------------------------------------
def element_operators(points, tris, theta, qx, qz, Dxx, Dxz, Dzz, mu=0.0, rho_kd=0.0, mu_src=0.0, c_src=None):
    """Per-element stiffness (with +mu * int theta N N^T), mass
    (-int (theta + rho_kd) N N^T) and source (-mu_src int theta c_src N)."""
    dN = _ref_deriv(2)
    X = points[tris]
    J = np.einsum("ak,ekd->ead", dN, X)
    Jinv = np.linalg.inv(J)
    detJ = np.linalg.det(J)
    G = np.einsum("ead,dk->eak", Jinv, dN)             # (E, 2, 3)
    Gx, Gz = G[:, 0, :], G[:, 1, :]
    E = tris.shape[0]
    stiff = np.zeros((E, 3, 3))
    mass = np.zeros((E, 3, 3))
    src = np.zeros((E, 3))
    loc = {k: v[tris] for k, v in dict(th=theta, qx=qx, qz=qz, xx=Dxx, xz=Dxz, zz=Dzz).items()}
    cs = c_src[tris] if c_src is not None else None
    for q, w in zip(TRI_Q, TRI_W):
        N = _shape(q)
        f = {k: v @ N for k, v in loc.items()}
        disp = -(f["xx"][:, None, None] * np.einsum("ea,eb->eab", Gx, Gx)
                 + f["xz"][:, None, None] * np.einsum("ea,eb->eab", Gx, Gz)
                 + f["xz"][:, None, None] * np.einsum("ea,eb->eab", Gz, Gx)
                 + f["zz"][:, None, None] * np.einsum("ea,eb->eab", Gz, Gz))
        adv = -(f["qx"][:, None, None] * np.einsum("a,eb->eab", N, Gx)
                + f["qz"][:, None, None] * np.einsum("a,eb->eab", N, Gz))
        NN = np.outer(N, N)[None]
        stiff += w * (disp + adv) * detJ[:, None, None]
        if mu != 0.0:
            stiff += w * (mu * f["th"])[:, None, None] * NN * detJ[:, None, None]
        mass += w * (-f["th"] - rho_kd)[:, None, None] * NN * detJ[:, None, None]
        if cs is not None:
            src += -w * mu_src * (f["th"] * (cs @ N))[:, None] * N[None, :] * detJ[:, None]
    return stiff, mass, src


def coo_matrix(tris, n, Ae):
    rows = np.repeat(tris, 3, axis=1).ravel()
    cols = np.tile(tris, (1, 3)).ravel()
    return sp.coo_matrix((Ae.ravel(), (rows, cols)), shape=(n, n)).tocsr()


def element_vector(tris, n, be):
    b = np.zeros(n)
    for e in range(tris.shape[0]):
        for a in range(3):
            b[tris[e, a]] += be[e, a]
    return b


def lump(M):
    """nitrogen_solver.py:36-66: row sums on the diagonal."""
    return sp.diags(np.asarray(M.sum(axis=1)).ravel()).tocsr()


# ----------------------------------------------------------------------------
# edge terms (boundary.py:108-421)
# ----------------------------------------------------------------------------
def _edge_shape():
    return np.stack([(1 - KSI_1D) / 2, (1 + KSI_1D) / 2])      # N[:, j]


def neumann(b, flux, nodes, points):
    b = b.copy()
    N = _edge_shape()
    for k in range(len(nodes) - 1):
        I, J = nodes[k], nodes[k + 1]
        jac = (points[J, 0] - points[I, 0]) / 2
        loc = np.zeros(2)
        for j in range(2):
            loc = loc - jac * W_1D[j] * flux * N[:, j]
        b[I] += loc[0]
        b[J] += loc[1]
    return b


def freedrain_water(b, K, nodes, points):
    b = b.copy()
    N = _edge_shape()
    for k in range(len(nodes) - 1):
        I, J = nodes[k], nodes[k + 1]
        jac = (points[J, 0] - points[I, 0]) / 2
        loc = np.zeros(2)
        for j in range(2):
            loc = loc + (-jac * W_1D[j] * (N[:, j] @ K[[I, J]]) * N[:, j])
        b[I] += loc[0]
        b[J] += loc[1]
    return b


def _edge_matrix(n, nodes, entries):
    rows, cols, vals = [], [], []
    for (I, J), m in zip(zip(nodes[:-1], nodes[1:]), entries):
        rows += [I, I, J, J]
        cols += [I, J, I, J]
        vals += list(m)
    return sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()


def cauchy(A, b, flux, c_in, nodes, points):
    n = A.shape[0]
    N = _edge_shape()
    b = b.copy()
    ents = []
    for k in range(len(nodes) - 1):
        I, J = nodes[k], nodes[k + 1]
        jac = (points[J, 0] - points[I, 0]) / 2
        m = np.zeros(4)
        for j in range(2):
            wgt = jac * W_1D[j] * flux
            Nv = N[:, j]
            m += wgt * np.array([Nv[0] * Nv[0], Nv[0] * Nv[1], Nv[1] * Nv[0], Nv[1] * Nv[1]])
        ents.append(m)
        b[I] += jac * c_in * flux * (W_1D[0] * N[0, 0] + W_1D[1] * N[0, 1])
        b[J] += jac * c_in * flux * (W_1D[0] * N[1, 0] + W_1D[1] * N[1, 1])
    if len(nodes) > 1:
        A = A + _edge_matrix(n, nodes, ents)
    return A, b


def freedrain_nitrogen(A, K, theta, nodes, points):
    n = A.shape[0]
    N = _edge_shape()
    ents = []
    for k in range(len(nodes) - 1):
        I, J = nodes[k], nodes[k + 1]
        jac = (points[J, 0] - points[I, 0]) / 2
        m = np.zeros(4)
        for j in range(2):
            v = -(N[:, j] @ K[[I, J]]) / (N[:, j] @ theta[[I, J]])
            wgt = jac * W_1D[j] * v
            Nv = N[:, j]
            m += wgt * np.array([Nv[0] * Nv[0], Nv[0] * Nv[1], Nv[1] * Nv[0], Nv[1] * Nv[1]])
        ents.append(m)
    if len(nodes) > 1:
        A = A + _edge_matrix(n, nodes, ents)
    return A


# ----------------------------------------------------------------------------
# species systems
# ----------------------------------------------------------------------------
NITROGEN_DEFAULT = dict(DL=0.04, DT=0.0004, Dm=0.288e-4, rho=1.6e-6, Kd=0.34e-6, mu1=0.12, mu2=0.048, mu3=0.012)


def nitrogen_system(points, tris, species, fields, prm, c_n, c_prev, dt, K, top, bottom, flux, c_in):
    """System (A, b) of one species (1 NH4, 2 NO2, 3 NO3), nitrogen_solver.py:148-238."""
    n = points.shape[0]
    qx, qz, _, Dxx, Dxz, Dzz, theta = fields
    mu = (prm["mu1"], prm["mu2"], prm["mu3"])[species - 1]
    rho_kd = prm["rho"] * prm["Kd"] if species == 1 else 0.0
    mu_src = 0.0 if species == 1 else (prm["mu1"], prm["mu2"])[species - 2]
    st, ms, src = element_operators(points, tris, theta, qx, qz, Dxx, Dxz, Dzz, mu=mu, rho_kd=rho_kd,
                                    mu_src=mu_src, c_src=None if species == 1 else c_prev)
    S = coo_matrix(tris, n, st)
    ML = lump(coo_matrix(tris, n, ms))
    b = (1.0 / dt) * (ML @ c_n)
    if species != 1:
        b = b + element_vector(tris, n, src)
    if len(top):
        S, b = cauchy(S, b, flux, c_in, top, points)
    if len(bottom):
        S = freedrain_nitrogen(S, K, theta, bottom, points)
    A = ((1.0 / dt) * ML + S).tocsr()
    return A, b


def solute_system(points, tris, fields, c_n, dt, top, flux, c_in):
    """coupled_solver_v2.py:154-200: consistent mass, Cauchy inlet."""
    n = points.shape[0]
    qx, qz, _, Dxx, Dxz, Dzz, theta = fields
    st, ms, _ = element_operators(points, tris, theta, qx, qz, Dxx, Dxz, Dzz)
    S = coo_matrix(tris, n, st)
    M = coo_matrix(tris, n, ms)
    b = (1.0 / dt) * (M @ c_n)
    if len(top):
        S, b = cauchy(S, b, flux, c_in, top, points)
    A = ((1.0 / dt) * M + S).tocsr()
    return A, b


# ----------------------------------------------------------------------------
# coupled Richards step with edge terms (nitrogen_solver.py:100-145,
# coupled_solver_v2.py:82-135)
# ----------------------------------------------------------------------------
def richards_edge_step(pattern, points, phi, h_n, dt, solver, precond, top, bottom, flux, freedrain, tol, maxit,
                       ktol=1e-10, maxiter=100, restart=30):
    from .assembly import assemble, jacobi_inverse
    _, _, theta_n = van_genuchten(h_n, phi)
    h_m = h_n
    err = np.inf
    it = 0
    kry = []
    while err >= tol and it < maxit:
        C, K, th = van_genuchten(h_m, phi)
        vals, b = assemble(pattern, points, th, theta_n, h_m, K, C, dt)
        b = neumann(b, flux, top, points)
        if freedrain and len(bottom):
            b = freedrain_water(b, K, bottom, points)
        A = pattern.to_scipy(vals)
        dinv = jacobi_inverse(pattern, vals) if precond == "jacobi" else None
        h, info = krylov.solve(solver, A.dot, b, h_m, dinv, ktol, 0.0, maxiter, restart)
        kry.append(info["iters"])
        err = np.linalg.norm(h - h_m)
        h_m = h
        it += 1
    return h_m, err, it


def _solve(A, b, x0, solver, precond, ktol=1e-10, maxiter=100, restart=30):
    dinv = None
    if precond == "jacobi":
        d = A.diagonal()
        dinv = 1.0 / np.where(d != 0, d, 1.0)
    x, _ = krylov.solve(solver, A.dot, b, x0, dinv, ktol, 0.0, maxiter, restart)
    return x


def nitrogen_run(points, tris, phi, nsteps, dt0, solver="gmres", precond="none", prm=None, c_inlet=(50.0, 0.0, 20.0),
                 c_init=(0.0, 0.0, 0.0), h_init=-1.0, flux=-0.05, adaptive=True):
    """NitrogenSolver.solve (nitrogen_solver.py:305-398) for nsteps steps."""
    from .assembly import Pattern
    prm = dict(NITROGEN_DEFAULT if prm is None else prm)
    n = points.shape[0]
    y = points[:, 1]
    top = np.where(np.abs(y - 1.0) < 1e-10)[0]
    bottom = np.where(np.abs(y - 0.0) < 1e-10)[0]
    pattern = Pattern(tris, n)
    h_n = np.full(n, h_init)
    c = [np.full(n, v) for v in c_init]
    out = {k: [] for k in ("pressure_head", "theta", "NH4", "NO2", "NO3", "times", "iterations", "errors", "dt_values")}
    t, dt = 0.0, dt0
    for _ in range(nsteps):
        h, err, it = richards_edge_step(pattern, points, phi, h_n, dt, solver, precond, top, bottom, flux, True,
                                        1e-6, 100)
        fields = fluxes_dispersion(points, tris, h, phi, prm["DL"], prm["DT"], prm["Dm"])
        _, K, _ = van_genuchten(h, phi)
        new = []
        for s in (1, 2, 3):
            A, b = nitrogen_system(points, tris, s, fields, prm, c[s - 1], new[-1] if new else None, dt, K, top,
                                   bottom, flux, c_inlet[s - 1])
            new.append(_solve(A, b, c[s - 1], solver, precond))
        if adaptive:
            dt_new = min(1.2 * dt, 0.01) if it <= 5 else (dt if it <= 15 else max(0.5 * dt, 1e-6))
        else:
            dt_new = dt
        for k, v in zip(("pressure_head", "theta", "NH4", "NO2", "NO3", "times", "iterations", "errors", "dt_values"),
                        (h, fields[6], new[0], new[1], new[2], t, it, err, dt)):
            out[k].append(v)
        h_n, c = h, new
        t += dt
        dt = dt_new
    return {k: np.array(v) for k, v in out.items()}


def solute_run(points, tris, phi, nsteps, dt0, top, solver="gmres", precond="none", DL=0.5, DT=0.1, Dm=0.0,
               c_init=0.1, c_in=1.0, h_init=-1.3, flux=-0.1, cflux=-0.1, adaptive=True, m_it=5, M_it=7, amp=2.0,
               red=0.5):
    """CoupledSolver.solve (coupled_solver_v2.py:221-331) for nsteps steps."""
    from .assembly import Pattern
    n = points.shape[0]
    pattern = Pattern(tris, n)
    h_n = np.full(n, h_init)
    c = np.full(n, c_init)
    out = {k: [] for k in ("pressure_head", "theta", "solute", "times", "iterations", "errors", "dt_values")}
    t, dt = 0.0, dt0
    for _ in range(nsteps):
        _, _, theta_n = van_genuchten(h_n, phi)
        h, err, it = richards_edge_step(pattern, points, phi, h_n, dt, solver, precond, top, [], flux, False,
                                        1e-4, 100)
        fields = fluxes_dispersion(points, tris, h, phi, DL, DT, Dm)
        A, b = solute_system(points, tris, fields, c, dt, top, cflux, c_in)
        c_new = _solve(A, b, c, solver, precond)
        if adaptive:
            dr = amp * dt if it <= m_it else (dt if it <= M_it else red * dt)
            dt_new = max(dr, 1e-6)       # coupled_solver_v2.py:219-220: the min(., 0.5) is overwritten
        else:
            dt_new = dt
        for k, v in zip(("pressure_head", "theta", "solute", "times", "iterations", "errors", "dt_values"),
                        (h, fields[6], c_new, t, it, err, dt_new)):
            out[k].append(v)
        h_n, c = h, c_new
        t += dt
        dt = dt_new
    return {k: np.array(v) for k, v in out.items()}
