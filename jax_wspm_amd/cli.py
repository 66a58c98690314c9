"""Command line, a drop-in for src/cli.py of the reference.

    python -m jax_wspm_amd.cli --test-case Test1 --solver gmres --preconditioner jacobi \
        --mesh-size 25 --dt 1e-3 --tmax 0.25 --output-dir results

Flags, defaults and outputs follow src/cli.py:51-236: the same flags with the
same defaults, ``simulation_config.json`` (cli.py:123-125), and
``numerical_results_<case>.npz`` with the reference's keys (cli.py:155-196),
including the Tracy error metrics for Test1-3 (src/utils/error_calculation.py).
Extra flags select what the reference hard-codes (Picard rule, Krylov
tolerance/maxiter/restart), the unit-domain van Genuchten cases and the
MI355X options (precision, fixed-step mode).
"""
from __future__ import annotations

import argparse
import json
import resource
import sys
from pathlib import Path

import numpy as np

from .config import (DeviceParameters, ExponentialParameters, PicardParameters, ProblemParameters,
                     SimulationConfig, SolverParameters, TimeSteppingParameters, VanGenuchtenParameters)
from .tracy import final_errors

REFERENCE_CASES = ["Test1", "Test2", "Test3", "Test3D", "SoluteTest", "NitrogenTest"]
EXTRA_CASES = ["VGSquare", "VGCube"]
MESH_SIZES = ["25", "50", "100", "200", "400", "512", "1024", "2048", "4096", "8192", "16384", "65536",
              "267039", "1014454"]


def parse_args(argv=None):
    p = argparse.ArgumentParser(description="Richards Equation and Solute Transport Solver (MI355X)",
                                formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    # reference flags (src/cli.py:51-84)
    p.add_argument("--mesh-size", choices=MESH_SIZES, default="25", help="Mesh resolution")
    p.add_argument("--test-case", choices=REFERENCE_CASES + EXTRA_CASES, default="Test1",
                   help="Test case to simulate")
    p.add_argument("--solver", choices=["gmres", "cg", "bicgstab", "direct"], default="gmres",
                   help="Linear solver type")
    p.add_argument("--preconditioner", choices=["none", "jacobi", "ilu", "block_jacobi", "ssor"],
                   default="none", help="Preconditioner type")
    p.add_argument("--dt", type=float, default=1e-6, help="Initial time step size")
    p.add_argument("--tmax", type=float, default=1.0, help="Maximum simulation time")
    p.add_argument("--output-dir", type=str, default="results", help="Directory for output files")
    p.add_argument("--save-frequency", type=int, default=5, help="Frequency of saving results")
    # options the reference hard-codes, and the MI355X path
    p.add_argument("--mesh-dir", type=str, default=None,
                   help="reference data/mesh directory (default: generate the structured meshes)")
    p.add_argument("--picard-rule", choices=["v2", "published"], default="v2",
                   help="2D Picard rule: v2 = richards_solver_v2 (rel 1e-6, 100); published = the rule "
                        "that produced resultsRE2D/*.csv (abs 1e-4, 30, theta at step start)")
    p.add_argument("--tol", type=float, default=1e-10, help="Krylov relative tolerance")
    p.add_argument("--maxiter", type=int, default=100, help="Krylov max iterations (GMRES: cycles)")
    p.add_argument("--restart", type=int, default=30, help="GMRES restart")
    p.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    p.add_argument("--inner", choices=["krylov", "jacobi"], default="krylov",
                   help="mixed precision: FP32 inner solves of each refinement sweep (jacobi: Richardson steps, "
                        "needs --preconditioner jacobi)")
    p.add_argument("--inner-iters", type=int, default=0, help="--inner jacobi: steps per refinement sweep (0: 3)")
    p.add_argument("--nsteps", type=int, default=None, help="fixed-dt mode: exactly this many steps")
    p.add_argument("--cells", type=int, default=None, help="VGSquare/VGCube cells per side / Test3D cells")
    p.add_argument("--ks-sigma", type=float, default=None, help="VG cases: per-element log-normal Ks std")
    p.add_argument("--ks-length", type=float, default=0.05)
    p.add_argument("--ks-seed", type=int, default=0)
    p.add_argument("--history", default="all", help="'all', 'last' or k (every k-th step) kept on the host")
    p.add_argument("--mass-balance", action="store_true",
                   help="Test1-3: record the domain integrals of theta and of the exact theta every step")
    p.add_argument("--diagnostics", action="store_true",
                   help="SoluteTest/NitrogenTest: record the largest element Peclet and Courant numbers every step")
    p.add_argument("--reorder", choices=["none", "rcm"], default="none",
                   help="renumber the mesh's nodes (reverse Cuthill-McKee) for the GPU; results stay in the "
                        "mesh's own numbering")
    return p.parse_args(argv)


def build_config(args) -> SimulationConfig:
    time = TimeSteppingParameters(dt_init=args.dt, Tmax=args.tmax, nsteps=args.nsteps)
    solver = SolverParameters(solver_type=args.solver, precond_type=args.preconditioner, tol=args.tol,
                              maxiter=args.maxiter, restart=args.restart)
    vg = VanGenuchtenParameters()
    problem = ProblemParameters(cells=args.cells)
    if args.test_case in EXTRA_CASES:
        vg = VanGenuchtenParameters(Ks=1e-3, thetas=0.40, thetar=0.06, alpha=0.1, n=2.0, m=0.5)
        problem = ProblemParameters(soil_model="vg", cells=args.cells,
                                    h_init=-500.0 if args.test_case == "VGCube" else -1000.0,
                                    h_top=-10.0 if args.test_case == "VGCube" else -75.0,
                                    ks_field=None if args.ks_sigma is None else
                                    {"sigma": args.ks_sigma, "length": args.ks_length, "seed": args.ks_seed})
    cfg = SimulationConfig(exponential=ExponentialParameters(), van_genuchten=vg, time=time, solver=solver,
                           test_case=args.test_case, save_frequency=args.save_frequency,
                           picard=PicardParameters(rule=args.picard_rule),
                           device=DeviceParameters(precision=args.precision, inner=args.inner,
                                                   inner_iters=args.inner_iters), problem=problem)
    cfg.mesh.mesh_size = args.mesh_size
    if args.mesh_dir:
        cfg.mesh.mesh_dir = args.mesh_dir
    return cfg


def _history(v):
    return int(v) if str(v).isdigit() else v


def run(args):
    cfg = build_config(args)
    if args.precision == "mixed" and args.solver == "direct":
        raise ValueError("the direct solver is FP64 (--precision fp64)")
    if args.precision == "mixed" and args.preconditioner not in ("none", "jacobi"):
        raise NotImplementedError("mixed precision takes --preconditioner none or jacobi")
    if args.precision == "mixed" and args.inner == "jacobi" and args.preconditioner != "jacobi":
        raise ValueError("--inner jacobi needs --preconditioner jacobi")
    from .richards import RichardsSolver, RichardsSolver3D

    out_dir = Path(args.output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    with open(out_dir / "simulation_config.json", "w") as fh:
        json.dump(cfg.to_dict(), fh, indent=4)
    print("Starting simulation with configuration:")
    print(f"Test case: {args.test_case}")
    print(f"Solver: {args.solver} with {args.preconditioner} preconditioner")
    if args.test_case in ("SoluteTest", "NitrogenTest"):
        from .transport import CoupledSolver, NitrogenSolver
        cls = CoupledSolver if args.test_case == "SoluteTest" else NitrogenSolver
    else:
        cls = RichardsSolver3D if args.test_case == "Test3D" else RichardsSolver
    extra = {}
    if getattr(args, "reorder", "none") != "none":
        extra["reorder"] = args.reorder
    if args.mass_balance:
        if args.test_case not in ("Test1", "Test2", "Test3"):
            raise ValueError("--mass-balance needs Test1, Test2 or Test3")
        extra["mass_balance"] = True
    if args.diagnostics:
        if args.test_case not in ("SoluteTest", "NitrogenTest"):
            raise ValueError("--diagnostics needs SoluteTest or NitrogenTest")
        extra["diagnostics"] = True
    solver = cls(cfg, mesh_dir=args.mesh_dir, history=_history(args.history), **extra)
    results = solver.solve()
    solver.ctx.close()

    if args.test_case in ("Test1", "Test2", "Test3"):
        ex = cfg.exponential
        results.update(final_errors(results["points"], results["pressure_head"][-1], results["theta"][-1],
                                    results["times"][-1], int(args.test_case[-1]), alpha0=ex.alpha0,
                                    thetas=ex.thetas, thetar=ex.thetar, Ks=ex.Ks))
    save = {
        "pressure_head": results["pressure_head"],
        "theta": results["theta"],
        "times": results["times"],
        "points": results["points"],
        "iterations": results["iterations"],
        "errors": results["errors"],
        "dt_values": results["dt_values"],
        "simulation_time": results["simulation_time"],
        "memory_usage": float(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0),
        "elements": results["tetrahedra"] if "tetrahedra" in results else results["triangles"],
        "krylov_iterations": results["krylov_iterations"],
    }
    # src/cli.py:167-179; the reference's "solute": None entries are left out
    # (an object array would need allow_pickle to load)
    if args.test_case == "SoluteTest":
        save["solute"] = results["solute"]
    elif args.test_case == "NitrogenTest":
        for k in ("NH4_concentration", "NO2_concentration", "NO3_concentration"):
            save[k] = results[k]
    if args.test_case in ("Test1", "Test2", "Test3"):
        for k in ("l2_pressure", "linf_pressure", "l2_saturation", "linf_saturation", "l2_relative_pressure",
                  "l2_relative_saturation"):
            save[k] = float(results[k])
    for k in ("mass_balance_numerical", "mass_balance_exact", "mass_balance_error", "peclet_max", "courant_max"):
        if k in results:
            save[k] = np.asarray(results[k])
    res_file = out_dir / f"numerical_results_{args.test_case}.npz"
    np.savez(res_file, **save)
    steps = len(results["times"])
    print("Results saved successfully:")
    print(f"  - Configuration: {out_dir / 'simulation_config.json'}")
    print(f"  - Numerical results: {res_file}")
    print(f"Total Simulation Time: {results['simulation_time']:.2f} seconds")
    if steps:
        print(f"Physical Time Simulated: {results['times'][-1]:.2f} time units")
        print(f"Average Time per Step: {results['simulation_time'] / steps:.4f} seconds")
        print(f"Total Iterations: {int(np.sum(results['iterations']))}")
        print(f"Final Error: {results['errors'][-1]:.2e}")
    return results, res_file


def main(argv=None):
    args = parse_args(argv)
    try:
        run(args)
    except Exception as e:
        print("Error during simulation:", file=sys.stderr)
        print(str(e), file=sys.stderr)
        raise


if __name__ == "__main__":
    main()
