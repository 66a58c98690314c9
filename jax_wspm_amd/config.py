"""Simulation parameters, mirroring config/settings.py of the reference.

The dataclass names, fields and defaults follow config/settings.py:9-215 so a
reference configuration dict loads unchanged (``create_config_from_dict``).
Fields the reference hard-codes inside its solvers are exposed here with the
reference's values as defaults (``PicardParameters``), and the options of the
MI355X path sit in ``DeviceParameters`` and ``ProblemParameters``.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass, field
from typing import Any, Dict, Optional

import numpy as np


@dataclass
class VanGenuchtenParameters:
    """config/settings.py:9-25.  ``m`` defaults to 1 - 1/1.56 and, as in the
    reference (settings.py:17), does not follow a changed ``n`` unless set."""
    Ks: float = 0.2496
    thetas: float = 0.43
    thetar: float = 0.078
    alpha: float = 3.6
    n: float = 1.56
    m: float = 1 - 1 / 1.56

    def to_array(self) -> np.ndarray:
        return np.array([self.alpha, self.thetas, self.thetar, self.n, self.m, self.Ks])

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class ExponentialParameters:
    """config/settings.py:39-50 (Gardner model)."""
    Ks: float = 0.1
    thetas: float = 0.45
    thetar: float = 0.15
    alpha0: float = 0.164

    def to_array(self) -> np.ndarray:
        return np.array([self.alpha0, self.thetas, self.thetar, self.Ks])

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class SoluteParameters:
    """config/settings.py:27-37."""
    DL: float = 0.5
    DT: float = 0.1
    Dm: float = 0.0
    c_init: float = 0.1
    c_inlet: float = 1.0

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class NitrogenParameters:
    """config/settings.py:52-83."""
    DL: float = 0.04
    DT: float = 0.0004
    Dm: float = 0.288e-4
    rho: float = 1.6e-6
    Kd: float = 0.34e-6
    mu1: float = 0.12
    mu2: float = 0.048
    mu3: float = 0.012
    c_init_NH4: float = 0.0
    c_init_NO2: float = 0.0
    c_init_NO3: float = 0.0
    c_inlet_NH4: float = 50.0
    c_inlet_NO2: float = 0.0
    c_inlet_NO3: float = 20.0

    def to_array(self) -> np.ndarray:
        return np.array([self.DL, self.DT, self.Dm, self.rho, self.Kd, self.mu1, self.mu2, self.mu3])

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class TimeSteppingParameters:
    """config/settings.py:85-97."""
    dt_init: float = 1e-5
    Tmax: float = 1.0
    m_it: int = 5
    M_it: int = 7
    W_it: int = 10
    lambda_amp: float = 2.0
    lambda_red: float = 0.5
    nsteps: Optional[int] = None     # fixed-dt mode: exactly nsteps steps; None = adaptive to Tmax

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class SolverParameters:
    """config/settings.py:99-114."""
    solver_type: str = "gmres"
    precond_type: str = "ilu"
    tol: float = 1e-10
    maxiter: int = 100
    restart: int = 30
    precond_params: Dict[str, Any] = field(default_factory=lambda: {"alpha": 1.0, "omega": 1.0, "block_size": 2})

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class PicardParameters:
    """Picard stopping rule.  The reference hard-codes it per solver:
    2D ``richards_solver_v2.py:155,162`` relative 1e-6 / 100 iterations,
    3D ``richards_solver_3D_v2.py:186,216`` absolute 1e-4 / 30.  The published
    2D CSVs (resultsRE2D/*.csv) were produced by the earlier 2D solver
    (src/solvers/.ipynb_checkpoints/richards_solver-checkpoint.py:157-164:
    absolute 1e-4 / 30, theta stored at the start of the step): ``rule="published"``.
    ``None`` fields take the rule's values."""
    rule: str = "v2"                 # "v2" | "published"
    tol: Optional[float] = None
    maxit: Optional[int] = None
    mode: Optional[str] = None       # "rel" | "abs"

    def resolve(self, dim: int):
        if dim == 3 or self.rule == "published":
            tol, maxit, mode = 1e-4, 30, "abs"
        else:
            tol, maxit, mode = 1e-6, 100, "rel"
        return (self.tol if self.tol is not None else tol,
                self.maxit if self.maxit is not None else maxit,
                self.mode if self.mode is not None else mode)

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class DeviceParameters:
    precision: str = "fp64"          # "fp64" | "mixed"
    inner_tol: float = 1e-6          # FP32 inner relative tolerance in mixed mode
    inner: str = "krylov"            # mixed mode inner solves: "krylov" | "jacobi"
    inner_iters: int = 0             # "jacobi": Richardson steps per refinement sweep (0: 3)

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class ProblemParameters:
    """Problem definitions beyond the reference's Test1-3/Test3D (the unit
    square / cube van Genuchten cases).  See ``jax_wspm_amd.problems``."""
    soil_model: str = "gardner"      # used by VGSquare/VGCube: "vg"
    cells: Optional[int] = None      # structured mesh cells per side
    h_init: float = -1000.0
    h_top: float = -75.0
    ks_field: Optional[Dict[str, Any]] = None   # {"sigma": 0.5, "length": 0.05, "seed": 0}

    def to_dict(self) -> Dict:
        return asdict(self)


@dataclass
class MeshParameters:
    """config/settings.py:116-167."""
    mesh_size: str = "25"
    mesh_dir: str = "data/mesh"
    test_case: str = "Test1"

    @property
    def mesh_files(self):
        from pathlib import Path
        base = Path(self.mesh_dir)
        if self.test_case == "SoluteTest":
            return base / "solute" / "p_Pinns.csv", base / "solute" / "t_Pinns.csv"
        if self.test_case == "Test3D":
            return base / "richards" / "p_3D.csv", base / "richards" / "t_3D.csv"
        return base / "richards" / f"p{self.mesh_size}.csv", base / "richards" / f"t{self.mesh_size}.csv"

    def to_dict(self) -> Dict:
        p, t = self.mesh_files
        return {"mesh_size": self.mesh_size, "mesh_dir": str(self.mesh_dir), "test_case": self.test_case,
                "points_file": str(p), "triangles_file": str(t)}


@dataclass
class SimulationConfig:
    """config/settings.py:169-215."""
    exponential: ExponentialParameters = field(default_factory=ExponentialParameters)
    van_genuchten: VanGenuchtenParameters = field(default_factory=VanGenuchtenParameters)
    solute: SoluteParameters = field(default_factory=SoluteParameters)
    nitrogen: NitrogenParameters = field(default_factory=NitrogenParameters)
    time: TimeSteppingParameters = field(default_factory=TimeSteppingParameters)
    solver: SolverParameters = field(default_factory=SolverParameters)
    test_case: str = "Test1"
    save_frequency: int = 1
    picard: PicardParameters = field(default_factory=PicardParameters)
    device: DeviceParameters = field(default_factory=DeviceParameters)
    problem: ProblemParameters = field(default_factory=ProblemParameters)

    def __post_init__(self):
        self.mesh = MeshParameters(test_case=self.test_case)

    def to_dict(self) -> Dict:
        return {
            "exponential": self.exponential.to_dict(),
            "van_genuchten": self.van_genuchten.to_dict(),
            "solute": self.solute.to_dict(),
            "nitrogen": self.nitrogen.to_dict(),
            "time": self.time.to_dict(),
            "solver": self.solver.to_dict(),
            "mesh": self.mesh.to_dict(),
            "test_case": self.test_case,
            "save_frequency": self.save_frequency,
            "picard": self.picard.to_dict(),
            "device": self.device.to_dict(),
            "problem": self.problem.to_dict(),
        }


def create_config_from_dict(d: Dict[str, Any]) -> SimulationConfig:
    """config/settings.py:218-228, plus the MI355X-path sections."""
    return SimulationConfig(
        exponential=ExponentialParameters(**d.get("exponential", {})),
        van_genuchten=VanGenuchtenParameters(**d.get("van_genuchten", {})),
        solute=SoluteParameters(**d.get("solute", {})),
        nitrogen=NitrogenParameters(**d.get("nitrogen", {})),
        time=TimeSteppingParameters(**d.get("time", {})),
        solver=SolverParameters(**d.get("solver", {})),
        test_case=d.get("test_case", "Test1"),
        save_frequency=d.get("save_frequency", 1),
        picard=PicardParameters(**d.get("picard", {})),
        device=DeviceParameters(**d.get("device", {})),
        problem=ProblemParameters(**d.get("problem", {})),
    )
