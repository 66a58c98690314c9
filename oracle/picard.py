"""Oracle Picard and time loops for the mixed-form Richards equation.

2D: src/solvers/richards_solver_v2.py:91-168 (Picard while err >= tol and
it < 100, err = |h - h_m| / |h|), :272-352 (time loop, adaptive dt, time
recorded before ``t += dt_new``, theta stored at the end of the step).
3D: src/solvers/richards_solver_3D_v2.py:177-302 (err = |h - h_m| absolute,
tol 1e-4, <= 30 iterations, theta stored at the start of the step).
Adaptivity: config/settings.py:85-97 (dt*2 if it <= 5, dt if it <= 7, else
dt/2).  A fixed-step mode (``nsteps``) runs exactly that many steps of ``dt0``.
"""
from __future__ import annotations

import time as _time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse.linalg as spla

from . import assembly, precond, krylov, soil


@dataclass
class Problem:
    points: np.ndarray
    elems: np.ndarray
    soil_model: str                 # "gardner" | "vg"
    phi: tuple
    bc_lists: list                  # [(nodes, values), ...] applied in order
    h0: np.ndarray
    dt0: float
    tmax: float = 1.0
    nsteps: int | None = None       # fixed-dt mode when set
    adaptive: bool = True
    picard_tol: float = 1e-6
    picard_maxit: int = 100
    picard_mode: str = "rel"        # "rel" (2D reference) | "abs" (3D reference)
    theta_at_start: bool = False    # 3D stores theta of h_n
    solver: str = "gmres"
    precond: str = "none"
    tol: float = 1e-10
    maxiter: int = 100
    restart: int = 30
    Ks_elem: np.ndarray | None = None
    m_it: int = 5
    M_it: int = 7
    lambda_amp: float = 2.0
    lambda_red: float = 0.5
    extra: dict = field(default_factory=dict)


def adapt_dt(pb: Problem, dt, it):
    if it <= pb.m_it:
        return pb.lambda_amp * dt
    if it <= pb.M_it:
        return dt
    return pb.lambda_red * dt


class Stepper:
    def __init__(self, pb: Problem):
        self.pb = pb
        self.n = pb.points.shape[0]
        self.pattern = assembly.Pattern(pb.elems, self.n)
        phi = np.asarray(pb.phi, dtype=np.float64)
        if pb.Ks_elem is not None:
            phi = phi.copy()
            phi[-1] = 1.0                  # nodal model returns Kr; Ks_e is per element
        self.phi = phi
        self.krylov_iters = []

    def props(self, h):
        return soil.soil(self.pb.soil_model, h, self.phi)

    def picard_iteration(self, h_m, theta_n, dt):
        pb = self.pb
        C, K, th = self.props(h_m)
        vals, b = assembly.assemble(self.pattern, pb.points, th, theta_n, h_m, K, C, dt,
                                    pb.bc_lists, pb.Ks_elem)
        A = self.pattern.to_scipy(vals)
        if pb.solver == "direct":
            # solver_types.py:60-63: jnp.linalg.solve (LU with partial pivoting) on
            # the dense matrix; SciPy's sparse LU solves the same system
            self.krylov_iters.append(0)
            return spla.spsolve(A.tocsc(), b)
        if pb.precond == "jacobi":
            dinv = assembly.jacobi_inverse(self.pattern, vals)
        elif pb.precond in ("block_jacobi", "ssor", "ilu"):
            dinv = precond.make(pb.precond, A)
        else:
            dinv = None
        h, info = krylov.solve(pb.solver, A.dot, b, h_m, dinv, pb.tol, 0.0, pb.maxiter, pb.restart)
        self.krylov_iters.append(info["iters"])
        return h

    def step(self, h_n, dt):
        pb = self.pb
        _, _, theta_n = self.props(h_n)
        h_m = h_n
        err = np.inf
        it = 0
        while err >= pb.picard_tol and it < pb.picard_maxit:
            h = self.picard_iteration(h_m, theta_n, dt)
            d = np.linalg.norm(h - h_m)
            err = d / np.linalg.norm(h) if pb.picard_mode == "rel" else d
            h_m = h
            it += 1
        return h_m, err, it, theta_n


def run(pb: Problem, max_steps: int | None = None):
    st = Stepper(pb)
    h_n = pb.h0.astype(np.float64).copy()
    t = 0.0
    dt = float(pb.dt0)
    out = {k: [] for k in ("pressure_head", "theta", "times", "iterations", "errors", "dt_values")}
    t0 = _time.perf_counter()
    k = 0
    while True:
        if pb.nsteps is not None:
            if k >= pb.nsteps:
                break
        elif not t < pb.tmax:
            break
        if max_steps is not None and k >= max_steps:
            break
        h, err, it, theta_n = st.step(h_n, dt)
        dt_new = adapt_dt(pb, dt, it) if (pb.adaptive and pb.nsteps is None) else dt
        theta = theta_n if pb.theta_at_start else st.props(h)[2]
        out["pressure_head"].append(h)
        out["theta"].append(theta)
        out["times"].append(t)
        out["iterations"].append(it)
        out["errors"].append(float(err))
        out["dt_values"].append(dt_new)
        h_n = h
        t += dt_new
        dt = dt_new
        k += 1
    res = {k2: np.array(v) for k2, v in out.items()}
    res["simulation_time"] = _time.perf_counter() - t0
    res["krylov_iters"] = np.array(st.krylov_iters)
    return res
