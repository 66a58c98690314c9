"""Oracle P1 element matrices, CSR assembly and Dirichlet rows (FP64).

Element integrals follow src/numerics/assembly_re_v2.py:13-60 (triangles, 3-point
rule of src/numerics/gauss.py:31-37 with its 15-digit literals) and
src/numerics/assembly_re_3D.py:12-70 (tets, 4-point rule of
src/numerics/gauss_3D.py:29-39 with weights 0.25 each).  The quadrature loop is
evaluated literally (per point: N(q), J, inv(J), det(J)), vectorised over
elements, so this file does not share the closed form the HIP kernel uses.

Global assembly follows assembly_re_v2.py:111-164 (sum of duplicates) on the
structural pattern; Dirichlet rows follow src/models/boundary.py:38-53 (dense,
2D) and :603-686 (sparse, 3D): for each BC node in list order the row is zeroed,
the diagonal set to 1 and b set to the value; later lists overwrite earlier ones.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .mesh import csr_pattern

TRI_Q = np.array([[0.166666666666667, 0.166666666666667],
                  [0.166666666666667, 0.666666666666667],
                  [0.666666666666667, 0.166666666666667]])
TRI_W = np.array([0.166666666666667, 0.166666666666667, 0.166666666666667])
_A3, _B3 = 0.585410196624969, 0.138196601125011
TET_Q = np.array([[_A3, _B3, _B3], [_B3, _A3, _B3], [_B3, _B3, _A3], [_B3, _B3, _B3]])
TET_W = np.array([0.25, 0.25, 0.25, 0.25])


def _ref_deriv(dim):
    d = np.zeros((dim, dim + 1))
    d[:, 0] = -1.0
    d[np.arange(dim), np.arange(1, dim + 1)] = 1.0
    return d


def _shape(q):
    return np.concatenate([[1.0 - q.sum()], q])


def element_matrices(points, elems, theta_m, theta_n, h_m, K, C, dt, Ks_elem=None):
    """Per-element local matrix (E, k, k) and source (E, k), k = dim+1.

    If ``Ks_elem`` is given, ``K`` holds the relative conductivity Kr(h) and the
    conductivity at a quadrature point is Ks_e * (N . Kr)  (the per-element
    extension; equal to the reference when Ks_e is constant).
    """
    dim = points.shape[1]
    Q, W = (TRI_Q, TRI_W) if dim == 2 else (TET_Q, TET_W)
    dN = _ref_deriv(dim)
    X = points[elems]                                   # (E, k, dim)
    J = np.einsum("ak,ekd->ead", dN, X)                 # (E, dim, dim)
    Jinv = np.linalg.inv(J)
    detJ = np.linalg.det(J)
    G = np.einsum("ead,dk->eak", Jinv, dN)              # (E, dim, k)
    GtG = np.einsum("eak,eal->ekl", G, G)
    Kl = K[elems]
    E, k = elems.shape
    stiff = np.zeros((E, k, k))
    mass = np.zeros((E, k, k))
    src = np.zeros((E, k))
    for q, w in zip(Q, W):
        Nq = _shape(q)
        Kq = Kl @ Nq
        if Ks_elem is not None:
            Kq = Ks_elem * Kq
        stiff += (w * Kq)[:, None, None] * GtG * detJ[:, None, None]
        mass += w * np.outer(Nq, Nq)[None] * detJ[:, None, None]
        src -= (w * Kq)[:, None] * G[:, dim - 1, :] * detJ[:, None]
    ML = mass.sum(axis=2)                               # row-sum lumping
    Cl = C[elems]
    A = stiff.copy()
    idx = np.arange(k)
    A[:, idx, idx] += (1.0 / dt) * (ML * Cl)
    rhs = src - (1.0 / dt) * (ML * (theta_m[elems] - theta_n[elems] - Cl * h_m[elems]))
    return A, rhs


class Pattern:
    """Structural CSR pattern of a P1 mesh plus the COO->CSR position map."""

    def __init__(self, elems, n):
        self.n = int(n)
        self.elems = np.asarray(elems, dtype=np.int64)
        self.indptr, self.indices = csr_pattern(self.elems, self.n)
        k = self.elems.shape[1]
        rows = np.repeat(self.elems, k, axis=1).ravel()
        cols = np.tile(self.elems, (1, k)).ravel()
        # position of every (element, a, b) entry inside the CSR value array
        start = self.indptr[rows]
        pos = np.empty_like(rows)
        for r0 in range(0, rows.size, 1 << 22):
            sl = slice(r0, r0 + (1 << 22))
            pos[sl] = _row_search(self.indptr, self.indices, rows[sl], cols[sl])
        self.coo_pos = pos.reshape(-1, k, k)
        self.diag_pos = _row_search(self.indptr, self.indices, np.arange(self.n), np.arange(self.n))
        del start

    @property
    def nnz(self):
        return int(self.indices.size)

    def scatter(self, Ae, be):
        vals = np.zeros(self.nnz)
        np.add.at(vals, self.coo_pos.ravel(), Ae.ravel())
        b = np.zeros(self.n)
        np.add.at(b, self.elems.ravel(), be.ravel())
        return vals, b

    def to_scipy(self, vals):
        return sp.csr_matrix((vals, self.indices, self.indptr), shape=(self.n, self.n))


def _row_search(indptr, indices, rows, cols):
    """CSR position of (row, col): binary search inside each row."""
    lo = indptr[rows].copy()
    hi = indptr[rows + 1].copy()
    while True:
        active = lo < hi
        if not active.any():
            break
        mid = (lo + hi) // 2
        go_right = active & (indices[np.minimum(mid, indices.size - 1)] < cols)
        lo = np.where(go_right, mid + 1, lo)
        hi = np.where(active & ~go_right, mid, hi)
    return lo


def apply_dirichlet(pattern: Pattern, vals, b, bc_lists):
    """Row replacement for each (nodes, values) list in order (later lists win)."""
    vals = vals.copy()
    b = b.copy()
    ip = pattern.indptr
    for nodes, values in bc_lists:
        nodes = np.asarray(nodes, dtype=np.int64)
        values = np.broadcast_to(np.asarray(values, dtype=np.float64), nodes.shape)
        for i, v in zip(nodes, values):
            vals[ip[i]:ip[i + 1]] = 0.0
            vals[pattern.diag_pos[i]] = 1.0
            b[i] = v
    return vals, b


def assemble(pattern: Pattern, points, theta_m, theta_n, h_m, K, C, dt, bc_lists=(), Ks_elem=None):
    Ae, be = element_matrices(points, pattern.elems, theta_m, theta_n, h_m, K, C, dt, Ks_elem)
    vals, b = pattern.scatter(Ae, be)
    if bc_lists:
        vals, b = apply_dirichlet(pattern, vals, b, bc_lists)
    return vals, b


def jacobi_inverse(pattern: Pattern, vals):
    """src/solvers/jax_linear_solver.py:12-36: 1/diag with 0 -> 1."""
    d = vals[pattern.diag_pos]
    return 1.0 / np.where(d != 0, d, 1.0)


def dense_bruteforce(points, elems, n, theta_m, theta_n, h_m, K, C, dt, Ks_elem=None):
    """Dense n x n assembly by explicit loops (tiny meshes only)."""
    Ae, be = element_matrices(points, elems, theta_m, theta_n, h_m, K, C, dt, Ks_elem)
    A = np.zeros((n, n))
    b = np.zeros(n)
    for e in range(elems.shape[0]):
        for a, i in enumerate(elems[e]):
            b[i] += be[e, a]
            for c, j in enumerate(elems[e]):
                A[i, j] += Ae[e, a, c]
    return A, b
