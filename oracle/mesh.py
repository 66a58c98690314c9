"""Oracle meshes: structured generators, fixture loader, boundary node sets.

Follows src/mesh/loader.py:11-108 (1-based CSV -> 0-based), the structured
diagonal-split triangulation of data/mesh/richards/p{N}.csv / t{N}.csv and the
Kuhn 6-tet grid of data/mesh/richards/p_3D.zip / t_3D.zip, and the boundary
extraction of src/models/boundary.py:444-547.
"""
from __future__ import annotations

import io
import zipfile
from pathlib import Path

import numpy as np

L_TRACY = 15.24

# Kuhn subdivision of a unit cube into 6 tets sharing the (0,0,0)-(1,1,1)
# diagonal, in the local corner order of data/mesh/richards/t_3D.csv rows 1-6.
KUHN_TETS = (
    ((0, 0, 1), (0, 0, 0), (0, 1, 1), (1, 1, 1)),
    ((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)),
    ((1, 0, 0), (0, 0, 0), (1, 0, 1), (1, 1, 1)),
    ((0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)),
    ((0, 1, 0), (0, 0, 0), (1, 1, 0), (1, 1, 1)),
    ((0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1)),
)


def grid_coords(n_cells: int, length: float) -> np.ndarray:
    """1-D node coordinates, rounded to 12 decimals as the fixtures were written."""
    return np.round(np.arange(n_cells + 1) * (length / n_cells), 12)


def structured_tri(n_cells: int, length: float = 1.0):
    """Diagonal-split square grid: 2 triangles per cell, x fastest.

    Node (i, j) -> j*(N+1) + i.  Cell (i, j) -> triangles
    [n00, n10, n11] and [n00, n11, n01] (data/mesh/richards/t25.csv rows 1-2).
    """
    N = int(n_cells)
    xs = grid_coords(N, length)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    points = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    n00 = (j * (N + 1) + i).ravel()
    n10 = n00 + 1
    n01 = n00 + N + 1
    n11 = n00 + N + 2
    tri = np.empty((2 * N * N, 3), dtype=np.int64)
    tri[0::2] = np.stack([n00, n10, n11], axis=1)
    tri[1::2] = np.stack([n00, n11, n01], axis=1)
    return points, tri


def structured_tet(n_cells: int, length: float = 1.0):
    """Kuhn 6-tet cube grid, x fastest then y then z."""
    N = int(n_cells)
    xs = grid_coords(N, length)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    base = (k * (N + 1) ** 2 + j * (N + 1) + i).ravel()
    tets = np.empty((6 * N ** 3, 4), dtype=np.int64)
    for t, corners in enumerate(KUHN_TETS):
        for c, (di, dj, dk) in enumerate(corners):
            tets[t::6, c] = base + dk * (N + 1) ** 2 + dj * (N + 1) + di
    return points, tets


def _read_csv(path: Path) -> np.ndarray:
    if path.suffix == ".zip":
        with zipfile.ZipFile(path) as z:
            name = z.namelist()[0]
            with z.open(name) as fh:
                return np.loadtxt(io.TextIOWrapper(fh), delimiter=",", ndmin=2)
    return np.loadtxt(path, delimiter=",", ndmin=2)


def load_mesh(p_file, t_file):
    """Mirror of src/mesh/loader.py:70-71: read, shift connectivity to 0-based."""
    points = _read_csv(Path(p_file)).astype(np.float64)
    elems = _read_csv(Path(t_file)).astype(np.int64) - 1
    return points, elems


def boundary_nodes_2d(points: np.ndarray, test_case: str = "Test1"):
    """src/models/boundary.py:482-547 (exact float equality for Test1-3)."""
    x, y = points[:, 0], points[:, 1]
    if test_case == "Test3":
        bottom = np.where(y == 0)[0]
    else:
        bottom = np.where((y == 0) | (x == 0) | (x == L_TRACY))[0]
    top = np.where(y == L_TRACY)[0]
    return top, bottom


def boundary_nodes_3d(points: np.ndarray, L: float = L_TRACY, tol: float = 1e-10):
    """src/models/boundary.py:444-480."""
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    top = np.where(np.abs(z - L) < tol)[0]
    bottom = np.where((np.abs(z) < tol) | (np.abs(x) < tol) | (np.abs(x - L) < tol)
                      | (np.abs(y) < tol) | (np.abs(y - L) < tol))[0]
    return top, bottom


def csr_pattern(elems: np.ndarray, n: int):
    """Structural CSR pattern: union of all (i, j) pairs within elements, sorted."""
    k = elems.shape[1]
    rows = np.repeat(elems, k, axis=1).ravel()
    cols = np.tile(elems, (1, k)).ravel()
    key = np.unique(rows.astype(np.int64) * n + cols)
    r = key // n
    c = key % n
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr, c.astype(np.int64)
