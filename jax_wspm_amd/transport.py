"""Coupled water flow + solute / nitrogen transport on the GPU.

``NitrogenSolver`` mirrors src/solvers/nitrogen_solver.py:68-413 and
``CoupledSolver`` mirrors src/solvers/coupled_solver_v2.py:25-331 (the
reference's ``Solver(config).solve() -> dict`` protocol, same result keys).
Per time step:

1. the Richards Picard loop (``wspm_richards_step``) with the reference's
   edge terms on b: Neumann inflow on the inlet (top) list and, for the
   nitrogen case, free drainage on the outlet (bottom) list
   (nitrogen_solver.py:100-145; coupled_solver_v2.py:82-135); absolute Picard
   error, tolerance 1e-6 (nitrogen) or 1e-4 (coupled), at most 100 iterations;
2. node fluxes and dispersion from the new h (``wspm_transport_fields``,
   src/models/transport_utilities.py:119-192);
3. one linear system per species (``wspm_assemble_transport`` + Krylov, x0 =
   the previous concentration): nitrogen NH4 -> NO2 -> NO3 in sequence with
   lumped mass, Cauchy inflow and free-drainage outflow
   (nitrogen_solver.py:148-238); solute with consistent mass and Cauchy
   inflow (coupled_solver_v2.py:137-200);
4. the reference's dt rule: nitrogen min(1.2 dt, 0.01) / dt / max(dt/2, 1e-6)
   at <= 5 / <= 15 / more Picard iterations (nitrogen_solver.py:400-412);
   coupled amp / keep / reduce at m_it / M_it with only the lower clamp
   1e-6 in effect (coupled_solver_v2.py:202-220, where the min(., 0.5) is
   overwritten).

The recorded dt follows each reference: the dt used (nitrogen) or the next
dt (coupled).
"""
from __future__ import annotations

import time
from typing import Any, Dict, Optional

import numpy as np
import torch

from .config import SimulationConfig
from .device import Context
from .mesh import load_and_validate_mesh
from .reorder import renumbering


def transport_boundary_nodes(points: np.ndarray, test_case: str):
    """(inlet, outlet) node lists of src/models/boundary.py:499-528."""
    x, y = points[:, 0], points[:, 1]
    if test_case == "SoluteTest":
        top = np.nonzero((y == 2.0) & (x <= 0.05) & (x >= -0.05))[0]
        bottom = np.nonzero(y == 0)[0]
    elif test_case == "NitrogenTest":
        top = np.nonzero(np.abs(y - 1.0) < 1e-10)[0]
        bottom = np.nonzero(np.abs(y - 0.0) < 1e-10)[0]
    else:
        raise ValueError(f"Invalid test case: {test_case}")
    return top, bottom


def _mesh(config: SimulationConfig, mesh_dir: Optional[str], points, triangles):
    if points is not None:
        return np.asarray(points, dtype=np.float64), np.asarray(triangles, dtype=np.int64), {}
    if config.test_case == "NitrogenTest" and mesh_dir is None:
        from .mesh import mesh_info, unit_square
        cells = int(config.mesh.mesh_size)
        pts, tri = unit_square(cells, 1.0)      # equals data/mesh/nitrogen/p_,t_{25,50,100}
        return pts, tri, mesh_info(pts, tri, config.test_case, config.mesh.mesh_size)
    pts, tri, info = load_and_validate_mesh(config.mesh.mesh_size, mesh_dir or str(config.mesh.mesh_dir),
                                            config.test_case)
    return np.asarray(pts), np.asarray(tri), info


class _TransportBase:
    picard_tol = 1e-6
    h_init = -1.0
    neumann_flux = -0.05
    freedrain = True

    def __init__(self, config: SimulationConfig, mesh_dir: Optional[str] = None, points=None, triangles=None,
                 history: Any = "all", diagnostics: bool = False, reorder: Optional[str] = None):
        """reorder "rcm": the GPU works on the reverse Cuthill-McKee
        renumbering of the mesh (jax_wspm_amd.reorder); the boundary lists
        keep the reference's order and results are in the mesh's own
        numbering."""
        self.config = config
        self.history = history
        self.diagnostics = bool(diagnostics)
        self.points, self.triangles, self.mesh_info = _mesh(config, mesh_dir, points, triangles)
        self.inlet, self.outlet = transport_boundary_nodes(self.points, config.test_case)
        self.renum = renumbering(reorder, self.triangles, len(self.points))
        rn = self.renum
        if rn is None:
            self.ctx = Context(self.points, self.triangles)
            self.ctx.edge_set("inlet", self.inlet)
            self.ctx.edge_set("outlet", self.outlet)
        else:
            self.ctx = Context(rn.points(self.points), rn.elements(self.triangles))
            self.ctx.edge_set("inlet", rn.node_ids(self.inlet))
            self.ctx.edge_set("outlet", rn.node_ids(self.outlet))
        self.ctx.richards_edges(self.neumann_flux, freedrain=self.freedrain)
        self.phi = np.asarray(config.van_genuchten.to_array(), dtype=np.float64)
        s = config.solver
        self.ctx.richards_init("vg", self.phi, s.solver_type, s.precond_type, np.full(len(self.points), self.h_init),
                               tol=s.tol, maxiter=s.maxiter, restart=s.restart, picard_tol=self.picard_tol,
                               picard_maxit=100, picard_mode="abs")

    def _krylov(self, A, b, x0, dinv):
        """One species / solute system with the configured solver
        (solver_types.solve_system's dispatch)."""
        s = self.config.solver
        if s.solver_type == "direct":
            return self.ctx.direct(A, b)
        if s.precond_type in ("block_jacobi", "ssor", "ilu"):
            x, _ = self.ctx.krylov_pc(s.solver_type, s.precond_type, A, b, x0, tol=s.tol, maxiter=s.maxiter,
                                      restart=s.restart)
            return x
        x, _ = self.ctx.krylov(s.solver_type, A, b, x0, dinv=dinv if s.precond_type == "jacobi" else None,
                               tol=s.tol, maxiter=s.maxiter, restart=s.restart)
        return x

    def _host(self, x) -> np.ndarray:
        """A nodal GPU array on the host, in the mesh's own numbering."""
        a = x.cpu().numpy()
        return a if self.renum is None else self.renum.nodal_back(a)

    def _keep(self, k: int, last: bool) -> bool:
        if self.history == "all":
            return True
        if self.history == "last":
            return last
        return last or k % int(self.history) == 0

    def _diag(self, out, fields, dt):
        """Whole-mesh maxima of the element Peclet / Courant numbers of this
        step's fields (transport_utilities.py:206-242 over every element)."""
        if self.diagnostics:
            pe, cr = self.ctx.peclet_courant(fields[0], fields[1], fields[6], fields[3], fields[5], dt,
                                             all_elements=True)
            out.setdefault("peclet_max", []).append(pe)
            out.setdefault("courant_max", []).append(cr)

    def _done(self, t, k):
        ts = self.config.time
        if getattr(ts, "nsteps", None):
            return k >= ts.nsteps
        return not t < ts.Tmax


class NitrogenSolver(_TransportBase):
    """NH4+ -> NO2- -> NO3- transport coupled to Richards flow."""

    def solve(self, max_steps: Optional[int] = None) -> Dict[str, Any]:
        nc = self.config.nitrogen
        n = len(self.points)
        dev = self.ctx.device
        c = [torch.full((n,), float(v), dtype=torch.float64, device=dev)
             for v in (nc.c_init_NH4, nc.c_init_NO2, nc.c_init_NO3)]
        inlet = (nc.c_inlet_NH4, nc.c_inlet_NO2, nc.c_inlet_NO3)
        mus = (nc.mu1, nc.mu2, nc.mu3)
        keys = ("pressure_head", "theta", "NH4_concentration", "NO2_concentration", "NO3_concentration")
        out = {k: [] for k in keys + ("times", "iterations", "errors", "dt_values", "krylov_iterations")}
        t, dt, k = 0.0, float(self.config.time.dt_init), 0
        t0 = time.perf_counter()
        while not self._done(t, k) and (max_steps is None or k < max_steps):
            st = self.ctx.richards_step(dt)
            h = self.ctx.richards_state()[0]
            fields = self.ctx.transport_fields(self.phi, h, nc.DL, nc.DT, nc.Dm)
            K = self.ctx.soil("vg", self.phi, h)[1]
            new = []
            for s in range(3):
                A, b, dinv = self.ctx.assemble_transport(
                    fields, c[s], dt, lumped=True, mu=mus[s], rho_kd=nc.rho * nc.Kd if s == 0 else 0.0,
                    mu_src=mus[s - 1] if s > 0 else 0.0, c_src=new[-1] if s > 0 else None,
                    cauchy=(self.neumann_flux, inlet[s]), K=K)
                new.append(self._krylov(A, b, c[s], dinv))
            it = st["picard_iters"]
            dt_new = min(1.2 * dt, 0.01) if it <= 5 else (dt if it <= 15 else max(0.5 * dt, 1e-6))
            if self.config.time.nsteps:
                dt_new = dt
            last = self._done(t + dt, k + 1) or (max_steps is not None and k + 1 >= max_steps)
            if self._keep(k, last):
                for key, v in zip(keys, (h, fields[6], new[0], new[1], new[2])):
                    out[key].append(self._host(v))
            out["times"].append(t)
            out["iterations"].append(it)
            out["errors"].append(st["err"])
            out["dt_values"].append(dt)
            out["krylov_iterations"].append(st["krylov_iters"])
            self._diag(out, fields, dt)
            c = new
            t += dt
            dt = dt_new
            k += 1
        res = {key: np.array(v) for key, v in out.items()}
        res.update(points=self.points, triangles=self.triangles, simulation_time=time.perf_counter() - t0,
                   mesh_info=self.mesh_info, config=self.config)
        return res


class CoupledSolver(_TransportBase):
    """Single-solute transport coupled to Richards flow (SoluteTest)."""
    picard_tol = 1e-4
    h_init = -1.3
    neumann_flux = -0.1
    freedrain = False

    def solve(self, max_steps: Optional[int] = None) -> Dict[str, Any]:
        sc = self.config.solute
        ts = self.config.time
        n = len(self.points)
        c = torch.full((n,), float(sc.c_init), dtype=torch.float64, device=self.ctx.device)
        keys = ("pressure_head", "theta", "solute")
        out = {k: [] for k in keys + ("times", "iterations", "errors", "dt_values", "krylov_iterations")}
        t, dt, k = 0.0, float(ts.dt_init), 0
        t0 = time.perf_counter()
        while not self._done(t, k) and (max_steps is None or k < max_steps):
            st = self.ctx.richards_step(dt)
            h = self.ctx.richards_state()[0]
            fields = self.ctx.transport_fields(self.phi, h, sc.DL, sc.DT, sc.Dm)
            A, b, dinv = self.ctx.assemble_transport(fields, c, dt, lumped=False, cauchy=(-0.1, sc.c_inlet))
            c_new = self._krylov(A, b, c, dinv)
            it = st["picard_iters"]
            dr = ts.lambda_amp * dt if it <= ts.m_it else (dt if it <= ts.M_it else ts.lambda_red * dt)
            dt_new = max(dr, 1e-6) if not ts.nsteps else dt
            last = self._done(t + dt, k + 1) or (max_steps is not None and k + 1 >= max_steps)
            if self._keep(k, last):
                for key, v in zip(keys, (h, fields[6], c_new)):
                    out[key].append(self._host(v))
            out["times"].append(t)
            out["iterations"].append(it)
            out["errors"].append(st["err"])
            out["dt_values"].append(dt_new)
            out["krylov_iterations"].append(st["krylov_iters"])
            self._diag(out, fields, dt)
            c = c_new
            t += dt
            dt = dt_new
            k += 1
        res = {key: np.array(v) for key, v in out.items()}
        res.update(points=self.points, triangles=self.triangles, simulation_time=time.perf_counter() - t0,
                   mesh_info=self.mesh_info, config=self.config)
        if len(res["solute"]):
            res.update(final_pressure=res["pressure_head"][-1], final_theta=res["theta"][-1],
                       final_solute=res["solute"][-1])
        return res
