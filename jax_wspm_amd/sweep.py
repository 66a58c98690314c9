"""Benchmark sweeps in the reference's harness format (job.sh).

The reference's job.sh loops over test cases, meshes, solvers and
preconditioners, runs the CLI for each and appends one CSV row read from
``numerical_results_<case>.npz`` (job.sh:17, 34-54).  This module does the
same in-process on the MI355X path:

    python -m jax_wspm_amd.sweep --test-cases Test1 Test2 Test3 --mesh-sizes 25 50 \
        --solvers cg bicgstab gmres --preconditioners none jacobi --dt 1e-3 --tmax 0.25 \
        --picard-rule published --csv results/benchmark_results.csv

The CSV header and number formats are job.sh's, so the files compare with
resultsRE2D/benchmark_results_cpu_*.csv directly.
"""
from __future__ import annotations

import argparse
from pathlib import Path

import numpy as np

HEADER = ("test_case,mesh_size,solver,preconditioner,simulation_time,iterations,l2_pressure,linf_pressure,"
          "l2_saturation,linf_saturation,l2_relative_pressure,l2_relative_saturation,memory_usage")


def csv_row(npz_path, test_case: str, mesh: str, solver: str, precond: str) -> str:
    """One row in job.sh's format (job.sh:41-54)."""
    d = np.load(npz_path)
    return (f"{test_case},{mesh},{solver},{precond},"
            f"{float(d['simulation_time']):.4f},"
            f"{int(np.sum(d['iterations']))},"
            f"{float(d['l2_pressure']):.4e},"
            f"{float(d['linf_pressure']):.4e},"
            f"{float(d['l2_saturation']):.4e},"
            f"{float(d['linf_saturation']):.4e},"
            f"{float(d['l2_relative_pressure']):.4e},"
            f"{float(d['l2_relative_saturation']):.4e},"
            f"{float(d['memory_usage']):.2f}")


def main(argv=None):
    ap = argparse.ArgumentParser(description="Richards benchmark sweep (job.sh format)")
    ap.add_argument("--test-cases", nargs="+", default=["Test1", "Test2", "Test3"])
    ap.add_argument("--mesh-sizes", nargs="+", default=["25"])
    ap.add_argument("--solvers", nargs="+", default=["bicgstab", "gmres"])
    ap.add_argument("--preconditioners", nargs="+", default=["none"])
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--tmax", type=float, default=0.25)
    ap.add_argument("--picard-rule", choices=["v2", "published"], default="v2")
    ap.add_argument("--output-dir", default="results")
    ap.add_argument("--mesh-dir", default=None)
    ap.add_argument("--csv", default=None, help="default: <output-dir>/benchmark_results.csv")
    a = ap.parse_args(argv)
    from .cli import parse_args, run

    out = Path(a.output_dir)
    out.mkdir(parents=True, exist_ok=True)
    csv = Path(a.csv) if a.csv else out / "benchmark_results.csv"
    with open(csv, "w") as fh:
        fh.write(HEADER + "\n")
    for tc in a.test_cases:
        for mesh in a.mesh_sizes:
            for solver in a.solvers:
                for pc in a.preconditioners:
                    print(f"Running: {tc} - mesh:{mesh} - solver:{solver} - precond:{pc}", flush=True)
                    cli_args = ["--test-case", tc, "--mesh-size", str(mesh), "--solver", solver,
                                "--preconditioner", pc, "--output-dir", str(out), "--tmax", str(a.tmax),
                                "--dt", str(a.dt), "--picard-rule", a.picard_rule, "--history", "last"]
                    if a.mesh_dir:
                        cli_args += ["--mesh-dir", a.mesh_dir]
                    _, npz = run(parse_args(cli_args))
                    with open(csv, "a") as fh:
                        fh.write(csv_row(npz, tc, str(mesh), solver, pc) + "\n")
    return csv


if __name__ == "__main__":
    main()
