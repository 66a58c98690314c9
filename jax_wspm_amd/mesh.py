"""Meshes: fixture loader, structured generators, boundary sets, Ks fields.

Loader semantics follow src/mesh/loader.py:11-108 (CSV or zipped CSV, 1-based
connectivity shifted to 0-based, shape and index checks, ``mesh_info``).  The
structured generators produce the reference's fixtures bit-for-bit
(data/mesh/richards/p{N}.csv / t{N}.csv: diagonal-split squares, x fastest;
p_3D / t_3D: Kuhn 6-tet cubes), so the fixtures need not travel with the code.
Boundary sets follow src/models/boundary.py:444-547.
"""
from __future__ import annotations

import io
import zipfile
from pathlib import Path
from typing import Dict, Optional, Tuple

import numpy as np

L_TRACY = 15.24

# corner offsets (di, dj, dk) of the 6 tets of a cube, data/mesh/richards/t_3D.csv rows 1-6
_KUHN = np.array([
    [(0, 0, 1), (0, 0, 0), (0, 1, 1), (1, 1, 1)],
    [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)],
    [(1, 0, 0), (0, 0, 0), (1, 0, 1), (1, 1, 1)],
    [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)],
    [(0, 1, 0), (0, 0, 0), (1, 1, 0), (1, 1, 1)],
    [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1)],
])


def _axis(cells: int, length: float) -> np.ndarray:
    # the fixtures hold coordinates printed to 12 decimals
    return np.round(np.arange(cells + 1, dtype=np.float64) * (length / cells), 12)


def unit_square(cells: int, length: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    n1 = cells + 1
    ax = _axis(cells, length)
    pts = np.empty((n1 * n1, 2))
    pts[:, 0] = np.tile(ax, n1)
    pts[:, 1] = np.repeat(ax, n1)
    c = np.arange(cells)
    base = (c[:, None] * n1 + c[None, :]).ravel()          # row j, column i
    tri = np.empty((2 * cells * cells, 3), dtype=np.int64)
    tri[0::2, 0], tri[0::2, 1], tri[0::2, 2] = base, base + 1, base + n1 + 1
    tri[1::2, 0], tri[1::2, 1], tri[1::2, 2] = base, base + n1 + 1, base + n1
    return pts, tri


def unit_cube(cells: int, length: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    n1 = cells + 1
    ax = _axis(cells, length)
    pts = np.empty((n1 ** 3, 3))
    pts[:, 0] = np.tile(ax, n1 * n1)
    pts[:, 1] = np.tile(np.repeat(ax, n1), n1)
    pts[:, 2] = np.repeat(ax, n1 * n1)
    c = np.arange(cells)
    base = (c[:, None, None] * n1 * n1 + c[None, :, None] * n1 + c[None, None, :]).ravel()
    off = _KUHN[..., 2] * n1 * n1 + _KUHN[..., 1] * n1 + _KUHN[..., 0]   # (6, 4)
    tets = (base[:, None, None] + off[None]).reshape(-1, 4).astype(np.int64)
    return pts, tets


def _read_table(path: Path) -> np.ndarray:
    if path.suffix == ".zip":
        with zipfile.ZipFile(path) as z:
            with z.open(z.namelist()[0]) as fh:
                return np.loadtxt(io.TextIOWrapper(fh), delimiter=",", ndmin=2)
    return np.loadtxt(path, delimiter=",", ndmin=2)


def load_and_validate_mesh(mesh_size: str, mesh_dir: str, test_case: str = "Test1",
                           prefix: Optional[Tuple[str, str]] = None):
    """src/mesh/loader.py:11-108: file choice per test case, -1 shift, checks."""
    base = Path(mesh_dir)
    if test_case == "SoluteTest":
        p_file, t_file, is_3d = base / "solute" / f"p_Pinns_{mesh_size}.csv", base / "solute" / f"t_Pinns_{mesh_size}.csv", False
    elif test_case == "NitrogenTest":
        p_file, t_file, is_3d = base / "nitrogen" / f"p_{mesh_size}.csv", base / "nitrogen" / f"t_{mesh_size}.csv", False
    elif test_case == "Test3D":
        pp, tp = prefix if prefix else ("p_3D", "t_3D")
        p_file, t_file, is_3d = base / "richards" / f"{pp}.zip", base / "richards" / f"{tp}.zip", True
    else:
        p_file, t_file, is_3d = base / "richards" / f"p{mesh_size}.csv", base / "richards" / f"t{mesh_size}.csv", False
    if not p_file.exists():
        raise FileNotFoundError(f"Points file not found: {p_file}")
    if not t_file.exists():
        raise FileNotFoundError(f"Elements file not found: {t_file}")
    points = _read_table(p_file).astype(np.float64)
    elements = _read_table(t_file).astype(np.int64) - 1
    validate(points, elements, 3 if is_3d else 2)
    return points, elements, mesh_info(points, elements, test_case, mesh_size)


def validate(points: np.ndarray, elements: np.ndarray, dim: int):
    if points.ndim != 2 or points.shape[1] != dim:
        raise ValueError(f"Points array must have {dim} columns")
    if elements.ndim != 2 or elements.shape[1] != dim + 1:
        raise ValueError(f"Elements array must have {dim + 1} columns")
    if elements.size and (elements.min() < 0 or elements.max() >= len(points)):
        raise ValueError("Element indices out of bounds")


def mesh_info(points, elements, test_case, mesh_size) -> Dict:
    dim = points.shape[1]
    return {
        "num_points": int(len(points)), "num_elements": int(len(elements)), "dimension": dim,
        "element_type": "tetrahedra" if dim == 3 else "triangles",
        "x_range": (float(points[:, 0].min()), float(points[:, 0].max())),
        "y_range": (float(points[:, 1].min()), float(points[:, 1].max())),
        "z_range": (float(points[:, 2].min()), float(points[:, 2].max())) if dim == 3 else None,
        "mesh_type": "solute" if test_case == "SoluteTest" else "richards",
        "mesh_size": mesh_size if test_case != "SoluteTest" else "Pinns",
    }


def extract_boundary_nodes(points: np.ndarray, test_case: str = "Test1"):
    """src/models/boundary.py:482-547 (2D, exact coordinate equality) and
    :444-480 (3D, |.| < 1e-10).  Returns (top, bottom) index arrays."""
    if test_case == "Test3D":
        tol, L = 1e-10, L_TRACY
        x, y, z = points.T
        top = np.nonzero(np.abs(z - L) < tol)[0]
        bottom = np.nonzero((np.abs(z) < tol) | (np.abs(x) < tol) | (np.abs(x - L) < tol)
                            | (np.abs(y) < tol) | (np.abs(y - L) < tol))[0]
        return top, bottom
    x, y = points[:, 0], points[:, 1]
    if test_case == "Test3":
        bottom = np.nonzero(y == 0)[0]
    else:
        bottom = np.nonzero((y == 0) | (x == 0) | (x == L_TRACY))[0]
    top = np.nonzero(y == L_TRACY)[0]
    return top, bottom


def correlated_log_ks(points: np.ndarray, elements: np.ndarray, cells: int, mean_ks: float = 1e-3,
                      sigma: float = 0.5, length: float = 0.05, seed: int = 0, domain: float = 1.0):
    """Per-element Ks = exp(ln mean_ks + sigma z), z a spatially correlated
    standard normal field: i.i.d. N(0, 1) per structured cell from
    numpy's PCG64(seed), smoothed with a Gaussian kernel exp(-d^2 / 2 l^2)
    truncated at 4 l, renormalised to zero mean and unit variance; every element
    takes the value of the cell that holds its centroid."""
    from scipy.ndimage import gaussian_filter

    dim = points.shape[1]
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((cells,) * dim)
    z = gaussian_filter(w, sigma=length / (domain / cells), mode="wrap", truncate=4.0)
    z = (z - z.mean()) / z.std()
    cent = points[elements].mean(axis=1)
    idx = np.clip(np.floor(cent / (domain / cells)).astype(np.int64), 0, cells - 1)
    ze = z[tuple(idx[:, d] for d in reversed(range(dim)))]
    return np.exp(np.log(mean_ks) + sigma * ze)
