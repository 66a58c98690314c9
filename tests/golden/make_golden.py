"""Write the small golden fixtures under tests/golden from the reference's data files.

Run in the build container only (it reads the reference checkout, which the GPU
box does not have):

    python tests/golden/make_golden.py /root/reference

Everything written is data the reference itself holds (mesh checksums, rows of
its published result CSVs, slices of its committed 3D result NPZ files, two of
its unstructured solute meshes), so
the tests can pin the oracle without the reference present.
"""
from __future__ import annotations

import csv
import hashlib
import io
import json
import sys
import zipfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _read(path: Path) -> np.ndarray:
    if path.suffix == ".zip":
        with zipfile.ZipFile(path) as z:
            with z.open(z.namelist()[0]) as fh:
                return np.loadtxt(io.TextIOWrapper(fh), delimiter=",", ndmin=2)
    return np.loadtxt(path, delimiter=",", ndmin=2)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(ref: Path):
    mesh_dir = ref / "data" / "mesh"
    meshes = {
        "richards/p25.csv": ("tri", 25, 15.24), "richards/p50.csv": ("tri", 50, 15.24),
        "richards/p100.csv": ("tri", 100, 15.24), "richards/p200.csv": ("tri", 200, 15.24),
        "nitrogen/p_25.csv": ("tri", 25, 1.0), "nitrogen/p_50.csv": ("tri", 50, 1.0),
        "nitrogen/p_100.csv": ("tri", 100, 1.0),
        "richards/p_3D.zip": ("tet", 20, 15.24),
    }
    out = {}
    for pfile, (kind, N, L) in meshes.items():
        tfile = pfile.replace("/p", "/t", 1)
        p = _read(mesh_dir / pfile).astype(np.float64)
        t = _read(mesh_dir / tfile).astype(np.int64) - 1
        out[pfile] = {"kind": kind, "cells": N, "length": L, "points_shape": list(p.shape),
                      "elems_shape": list(t.shape), "points_sha256": sha(p), "elems_sha256": sha(t)}
    (HERE / "mesh_checksums.json").write_text(json.dumps(out, indent=1))

    rows = []
    for tc in ("Test1", "Test2", "Test3"):
        with open(ref / "resultsRE2D" / f"benchmark_results_cpu_{tc}_WA.csv") as fh:
            for r in csv.DictReader(fh):
                if r["mesh_size"] in ("25", "50") and r["preconditioner"] in ("none", "jacobi") \
                        and r["solver"] in ("cg", "bicgstab", "gmres", "direct"):
                    rows.append({k: r[k] for k in ("test_case", "mesh_size", "solver", "preconditioner",
                                                   "iterations", "l2_pressure", "linf_pressure",
                                                   "l2_saturation", "linf_saturation",
                                                   "l2_relative_pressure", "l2_relative_saturation")})
    (HERE / "tracy_metrics.json").write_text(json.dumps(rows, indent=1))

    # the unstructured solute meshes of SoluteTest (data/mesh/solute, read by
    # src/mesh/loader.py:33-36): points, and elements shifted to 0-based as
    # the loader does; 25 is the CLI's default --mesh-size
    pinns = {}
    for tag in ("25", "1024"):
        pinns[f"points_{tag}"] = _read(mesh_dir / "solute" / f"p_Pinns_{tag}.csv").astype(np.float64)
        pinns[f"elements_{tag}"] = (_read(mesh_dir / "solute" / f"t_Pinns_{tag}.csv").astype(np.int64) - 1
                                    ).astype(np.int32)
    np.savez_compressed(HERE / "pinns_meshes.npz", **pinns)

    for solver in ("bicgstab", "gmres"):
        z = np.load(ref / "results3D" / f"numerical_results_Test3D_{solver}_{solver}_none_mesh25.npz",
                    allow_pickle=False)
        np.savez_compressed(HERE / f"test3d_{solver}.npz",
                            pressure_head_steps=np.array([0, 4, 9]),
                            pressure_head=z["pressure_head"][[0, 4, 9]],
                            theta_final=z["theta"][9],
                            times=z["times"], iterations=z["iterations"],
                            errors=z["errors"], dt_values=z["dt_values"])


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
