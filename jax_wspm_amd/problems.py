"""Problem definitions: the reference's test cases and the unit-domain
van Genuchten cases used for scale runs.

Test1-3 (2D) follow src/solvers/richards_solver_v2.py:50-89: Gardner model,
h0 = -15.24, top Dirichlet from the Tracy functions (src/models/boundary.py:17-33,
eps0 = exp(alpha0 * h0)), then bottom/left/right Dirichlet with the value
``jnp.full_like(int_nodes, -15.24)`` = -15 (richards_solver_v2.py:141), applied
in that order.  Test3D follows src/solvers/richards_solver_3D_v2.py:49-88 with
the 3D top function (boundary.py:423-442) and bottom + four sides at -15
(boundary.py:676-684).

VGSquare / VGCube: unit square / cube, van Genuchten soil (explicit m),
uniform initial head, Dirichlet head on the top face (y = 1 or z = 1), no-flux
elsewhere (natural boundary), optional per-element log-normal Ks field.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from . import mesh as M
from .config import SimulationConfig

TRACY_BOTTOM = float(int(-15.24))   # jnp.full_like(int64 nodes, -15.24) -> -15


@dataclass
class Problem:
    name: str
    points: np.ndarray
    elements: np.ndarray
    soil_model: str
    phi: Tuple[float, ...]
    bc_lists: List[Tuple[np.ndarray, np.ndarray]]
    h0: np.ndarray
    dt0: float
    tmax: float = 1.0
    nsteps: Optional[int] = None
    picard_tol: float = 1e-6
    picard_maxit: int = 100
    picard_mode: str = "rel"
    theta_at_start: bool = False
    solver: str = "gmres"
    precond: Optional[str] = "none"
    tol: float = 1e-10
    atol: float = 0.0
    maxiter: int = 100
    restart: int = 30
    precision: str = "fp64"
    inner_tol: float = 1e-6
    inner: str = "krylov"
    inner_iters: int = 0
    Ks_elem: Optional[np.ndarray] = None
    m_it: int = 5
    M_it: int = 7
    lambda_amp: float = 2.0
    lambda_red: float = 0.5
    mesh_info: Dict[str, Any] = field(default_factory=dict)

    @property
    def dim(self) -> int:
        return self.points.shape[1]

    @property
    def n(self) -> int:
        return self.points.shape[0]


def tracy_top(test_case: str, pts_top: np.ndarray, alpha0: float, L: float = M.L_TRACY, h0: float = -15.24):
    eps0 = np.exp(alpha0 * h0)
    x = pts_top[:, 0]
    if test_case == "Test1":
        f = (3 / 4) * np.sin(np.pi * x / L) - (1 / 4) * np.sin(3 * np.pi * x / L)
        return (1 / alpha0) * np.log(eps0 + (1 - eps0) * f)
    if test_case == "Test2":
        return (1 / alpha0) * np.log(eps0 + (1 - eps0) * np.sin(np.pi * x / L))
    if test_case == "Test3":
        return (1 / alpha0) * np.log(eps0 + 0.5 * (1 - eps0) * (1 - np.cos(2 * np.pi * x / L)))
    if test_case == "Test3D":
        y = pts_top[:, 1]
        return (1 / alpha0) * np.log(eps0 + (1 - eps0) * np.sin(np.pi * x / L) * np.sin(np.pi * y / L))
    raise ValueError(f"Invalid test case: {test_case}")


def _mesh_2d(mesh_size: str, mesh_dir: Optional[str], test_case: str):
    if mesh_dir:
        p = Path(mesh_dir) / "richards" / f"p{mesh_size}.csv"
        if p.exists():
            return M.load_and_validate_mesh(mesh_size, mesh_dir, test_case)
    pts, tri = M.unit_square(int(mesh_size), M.L_TRACY)
    return pts, tri, M.mesh_info(pts, tri, test_case, mesh_size)


def _mesh_3d(cells: int, mesh_dir: Optional[str]):
    if mesh_dir and cells == 20 and (Path(mesh_dir) / "richards" / "p_3D.zip").exists():
        return M.load_and_validate_mesh("3D", mesh_dir, "Test3D")
    pts, tets = M.unit_cube(cells, M.L_TRACY)
    return pts, tets, M.mesh_info(pts, tets, "Test3D", str(cells))


def from_config(cfg: SimulationConfig, mesh_dir: Optional[str] = None) -> Problem:
    """Build the problem the reference's CLI would run for ``cfg.test_case``."""
    tc = cfg.test_case
    sv = cfg.solver
    common = dict(solver=sv.solver_type, precond=sv.precond_type, tol=sv.tol, maxiter=sv.maxiter,
                  restart=sv.restart, precision=cfg.device.precision, inner_tol=cfg.device.inner_tol,
                  inner=cfg.device.inner, inner_iters=cfg.device.inner_iters,
                  m_it=cfg.time.m_it, M_it=cfg.time.M_it, lambda_amp=cfg.time.lambda_amp,
                  lambda_red=cfg.time.lambda_red, dt0=cfg.time.dt_init, tmax=cfg.time.Tmax,
                  nsteps=cfg.time.nsteps)
    ex = cfg.exponential
    if tc in ("Test1", "Test2", "Test3"):
        pts, tri, info = _mesh_2d(cfg.mesh.mesh_size, mesh_dir, tc)
        top, bottom = M.extract_boundary_nodes(pts, tc)
        hup = tracy_top(tc, pts[top], ex.alpha0)
        tol, maxit, mode = cfg.picard.resolve(2)
        return Problem(name=tc, points=pts, elements=tri, soil_model="gardner", phi=tuple(ex.to_array()),
                       bc_lists=[(top, hup), (bottom, np.full(bottom.shape, TRACY_BOTTOM))],
                       h0=np.full(len(pts), -15.24), picard_tol=tol, picard_maxit=maxit, picard_mode=mode,
                       theta_at_start=cfg.picard.rule == "published", mesh_info=info, **common)
    if tc == "Test3D":
        cells = int(cfg.problem.cells or 20)
        pts, tets, info = _mesh_3d(cells, mesh_dir)
        top, bottom = M.extract_boundary_nodes(pts, "Test3D")
        hup = tracy_top("Test3D", pts[top], ex.alpha0)
        tol, maxit, mode = cfg.picard.resolve(3)
        return Problem(name=tc, points=pts, elements=tets, soil_model="gardner", phi=tuple(ex.to_array()),
                       bc_lists=[(top, hup), (bottom, np.full(bottom.shape, TRACY_BOTTOM))],
                       h0=np.full(len(pts), -15.24), picard_tol=tol, picard_maxit=maxit, picard_mode=mode,
                       theta_at_start=True, mesh_info=info, **common)
    if tc in ("VGSquare", "VGCube"):
        return vg_problem(cfg, **common)
    raise NotImplementedError(f"test case {tc!r} is not on the MI355X path (supported: Test1, Test2, Test3, "
                              "Test3D, VGSquare, VGCube)")


def vg_problem(cfg: SimulationConfig, **common) -> Problem:
    pr = cfg.problem
    vg = cfg.van_genuchten
    dim = 3 if cfg.test_case == "VGCube" else 2
    cells = int(pr.cells or (64 if dim == 3 else 25))
    pts, el = (M.unit_cube if dim == 3 else M.unit_square)(cells, 1.0)
    top = np.nonzero(pts[:, dim - 1] == 1.0)[0]
    ks = None
    if pr.ks_field:
        f = dict(pr.ks_field)
        ks = M.correlated_log_ks(pts, el, cells, mean_ks=vg.Ks, sigma=f.get("sigma", 0.5),
                                 length=f.get("length", 0.05), seed=f.get("seed", 0))
    tol, maxit, mode = (cfg.picard.tol or 1e-6, cfg.picard.maxit or 100, cfg.picard.mode or "rel")
    return Problem(name=cfg.test_case, points=pts, elements=el, soil_model="vg", phi=tuple(vg.to_array()),
                   bc_lists=[(top, np.full(top.shape, pr.h_top))], h0=np.full(len(pts), pr.h_init),
                   picard_tol=tol, picard_maxit=maxit, picard_mode=mode, theta_at_start=False, Ks_elem=ks,
                   mesh_info=M.mesh_info(pts, el, cfg.test_case, str(cells)), **common)


def baseline_vg_config(cells: int, dim: int = 2, *, ks_field=None, nsteps=None, dt=1e-4, tmax=0.25,
                       solver="gmres", precond="jacobi", precision="fp64", h_init=None, h_top=None):
    """van Genuchten alpha=0.1, n=2, m=0.5 (explicit), theta_r=0.06,
    theta_s=0.40, Ks=1e-3 on the unit square / cube."""
    from .config import (DeviceParameters, ProblemParameters, SolverParameters, TimeSteppingParameters,
                         VanGenuchtenParameters)
    cfg = SimulationConfig(
        van_genuchten=VanGenuchtenParameters(Ks=1e-3, thetas=0.40, thetar=0.06, alpha=0.1, n=2.0, m=0.5),
        time=TimeSteppingParameters(dt_init=dt, Tmax=tmax, nsteps=nsteps),
        solver=SolverParameters(solver_type=solver, precond_type=precond),
        test_case="VGCube" if dim == 3 else "VGSquare",
        device=DeviceParameters(precision=precision),
        problem=ProblemParameters(soil_model="vg", cells=cells,
                                  h_init=(-500.0 if dim == 3 else -1000.0) if h_init is None else h_init,
                                  h_top=(-10.0 if dim == 3 else -75.0) if h_top is None else h_top,
                                  ks_field=ks_field),
    )
    return cfg
