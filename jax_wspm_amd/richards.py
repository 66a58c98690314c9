"""Richards solvers with the reference's ``Solver(config).solve() -> dict`` protocol.

``RichardsSolver`` mirrors src/solvers/richards_solver_v2.py:42-352 and
``RichardsSolver3D`` mirrors src/solvers/richards_solver_3D_v2.py:41-302.  The
Picard loop of one step runs in libwspm (``wspm_richards_step``); this file
keeps the host time loop of the reference: dt adaptation by iteration count
(config/settings.py:85-97; richards_solver_v2.py:91-105), the time recorded
before ``t += dt_new`` (richards_solver_v2.py:316,323), and theta stored at the
end (2D v2) or at the start (3D, and the earlier 2D solver that produced the
published CSVs) of each step.

``mass_balance=True`` (Test1-3) adds the records of the mass-balance solver
(src/solvers/.ipynb_checkpoints/richards_mass_balance-checkpoint.py:210-335):
per step the domain integral of theta(h) at the end of the step and of the
exact Tracy theta at the step's recorded time (``wspm_domain_integral``), and
|1 - numerical / exact| * 100.
"""
from __future__ import annotations

import time
from dataclasses import replace
from typing import Any, Dict, Optional, Union

import numpy as np
import torch

from .config import SimulationConfig
from .device import Context
from .problems import Problem, from_config
from .reorder import renumbering


def _renumbered(pb: Problem, rn) -> Problem:
    """The problem on the renumbered mesh (boundary lists in their order)."""
    return replace(pb, points=rn.points(pb.points), elements=rn.elements(pb.elements),
                   bc_lists=[(rn.node_ids(nd), np.asarray(v)) for nd, v in pb.bc_lists],
                   h0=rn.nodal(pb.h0), Ks_elem=rn.elemental(pb.Ks_elem))


def adapt_time_step(dt: float, iters: int, m_it=5, M_it=7, amp=2.0, red=0.5) -> float:
    if iters <= m_it:
        return amp * dt
    if iters <= M_it:
        return dt
    return red * dt


class RichardsSolver:
    """Solve a Richards problem on the GPU.

    ``history``: "all" keeps h and theta of every step on the host (reference
    behaviour), "last" keeps only the final step, an int k keeps every k-th
    step plus the last."""

    def __init__(self, config: Union[SimulationConfig, Problem], mesh_dir: Optional[str] = None,
                 history: Union[str, int] = "all", mass_balance: bool = False, reorder: Optional[str] = None):
        """reorder "rcm": the GPU works on the reverse Cuthill-McKee
        renumbering of the mesh (jax_wspm_amd.reorder); inputs and results
        stay in the problem's own numbering."""
        pb0 = config if isinstance(config, Problem) else from_config(config, mesh_dir)
        self.renum = renumbering(reorder, pb0.elements, len(pb0.points))
        self.problem = pb0 if self.renum is None else _renumbered(pb0, self.renum)
        self.config = config
        self.history = history
        pb = self.problem
        self.ctx = Context(pb.points, pb.elements)
        self.ctx.set_dirichlet(pb.bc_lists)
        self.points = pb0.points
        self.elements = pb0.elements
        self.mesh_info = pb.mesh_info
        self.mass_balance = bool(mass_balance)
        if self.mass_balance and not (pb.name in ("Test1", "Test2", "Test3") and pb.soil_model == "gardner"):
            raise ValueError("mass balance needs a Tracy test case (Test1, Test2, Test3)")

    def _exact_theta(self, t: float) -> torch.Tensor:
        """theta of Tracy's solution at time t (exact_solutions.py, as the
        mass-balance solver evaluates it at :272-276)."""
        from .tracy import exact_solution
        pb = self.problem
        alpha0, ts, tr, ks = (float(v) for v in pb.phi[:4])
        hd = -15.24
        eps0 = np.exp(alpha0 * hd)
        d = alpha0 * (ts - tr) / ks
        _, S = exact_solution(pb.points[:, 0], pb.points[:, 1], t, int(pb.name[-1]), -hd, alpha0, eps0, d)
        return torch.as_tensor((ts - tr) * S + tr, dtype=torch.float64, device=self.ctx.device)

    def _host(self, x: torch.Tensor) -> np.ndarray:
        """A nodal GPU array on the host, in the problem's own numbering."""
        a = x.cpu().numpy()
        return a if self.renum is None else self.renum.nodal_back(a)

    def init_state(self):
        pb = self.problem
        self.ctx.richards_init(pb.soil_model, pb.phi, pb.solver, pb.precond, pb.h0, Ks_elem=pb.Ks_elem,
                               tol=pb.tol, atol=pb.atol, maxiter=pb.maxiter, restart=pb.restart,
                               picard_tol=pb.picard_tol, picard_maxit=pb.picard_maxit,
                               picard_mode=pb.picard_mode, precision=pb.precision, inner_tol=pb.inner_tol,
                               inner=pb.inner, inner_iters=pb.inner_iters)

    def _keep(self, k: int) -> bool:
        if self.history == "all":
            return True
        if isinstance(self.history, int) and self.history > 0:
            return k % self.history == 0
        return False

    def solve(self, max_steps: Optional[int] = None) -> Dict[str, Any]:
        pb = self.problem
        self.init_state()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = 0.0
        dt = float(pb.dt0)
        rec = {k: [] for k in ("pressure_head", "theta", "times", "iterations", "errors", "dt_values",
                               "krylov_iterations")}
        if self.mass_balance:
            rec.update(mass_balance_numerical=[], mass_balance_exact=[])
        k = 0
        last = None
        while True:
            if pb.nsteps is not None:
                if k >= pb.nsteps:
                    break
            elif not t < pb.tmax:
                break
            if max_steps is not None and k >= max_steps:
                break
            st = self.ctx.richards_step(dt)
            dt_new = (dt if pb.nsteps is not None
                      else adapt_time_step(dt, st["picard_iters"], pb.m_it, pb.M_it, pb.lambda_amp, pb.lambda_red))
            rec["times"].append(t)
            rec["iterations"].append(st["picard_iters"])
            rec["errors"].append(st["err"])
            rec["dt_values"].append(dt_new)
            rec["krylov_iterations"].append(st["krylov_iters"])
            if self.mass_balance:
                th_end = self.ctx.richards_state()[1]
                rec["mass_balance_numerical"].append(self.ctx.domain_integral(th_end))
                rec["mass_balance_exact"].append(self.ctx.domain_integral(self._exact_theta(t)))
            if self._keep(k):
                h, th, thn = self.ctx.richards_state()
                rec["pressure_head"].append(self._host(h))
                rec["theta"].append(self._host(thn if pb.theta_at_start else th))
                last = k
            t += dt_new
            dt = dt_new
            k += 1
        if k > 0 and last != k - 1:
            h, th, thn = self.ctx.richards_state()
            rec["pressure_head"].append(self._host(h))
            rec["theta"].append(self._host(thn if pb.theta_at_start else th))
        torch.cuda.synchronize()
        sim = time.perf_counter() - t0
        out: Dict[str, Any] = {key: np.array(v) for key, v in rec.items()}
        out.update({
            "points": self.points,
            ("tetrahedra" if pb.dim == 3 else "triangles"): self.elements,
            "simulation_time": sim,
            "mesh_info": pb.mesh_info,
            "config": self.config,
        })
        if out["theta"].size:
            out["final_theta"] = out["theta"][-1]
        if self.mass_balance:
            from .diagnostics import mass_balance_error
            out["mass_balance_error"] = mass_balance_error(out["mass_balance_numerical"], out["mass_balance_exact"])
        return out


class RichardsSolver3D(RichardsSolver):
    """3D variant (src/solvers/richards_solver_3D_v2.py:41-302): same driver,
    the problem definition carries the 3D Picard rule and theta-at-start."""
