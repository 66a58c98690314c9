"""Oracle preconditioners of src/solvers/jax_linear_solver.py (FP64).

TEST INFRASTRUCTURE ONLY (see the package docstring).

* ``block_jacobi`` (jax_linear_solver.py:39-87): 2x2 diagonal blocks of
  consecutive unknowns (2k, 2k+1), each inverted after adding 1e-12 I; a
  trailing 1x1 block when n is odd.
* ``ssor`` (:90-172, omega = 1): forward sweep y = (D + L)^-1 x.  The
  backward sweep as written reads z at rows that it has not reached yet
  (still zero), so it reduces to z_i = (d_i y_i) / d_i: that is what the
  reference computes, and what this restates.
* ``ilu`` (:175-252, the "ilu0" of the reference): L = strict lower part of A
  scaled by 1/d_col (d = 1 where the diagonal is 0), U = strict upper part;
  y = (I + L)^-1 x, then z = (D + U)^-1 y.

Each returns a callable x -> M x.

The reference's ``create_ssor`` / ``create_ilu0`` store their factor values
with ``L_data.at[index_val[0]]`` where ``index_val[0]`` is the (row, col) pair
of the entry (jax_linear_solver.py:105-112, 187-199): the values land at the
positions ``row`` and ``col`` of the entry-indexed arrays instead of at the
entry, so its sweeps read scrambled coefficients that depend on JAX's BCOO
padding and ordering.  That behaviour is not restated (it cannot be pinned
without JAX); these are the factorisations the code describes.  The published
rows where the preconditioned solves converge (resultsRE2D, meshes 25 and 50)
do not depend on the difference and are reproduced; rows where they do
(Test1 mesh 100 cg+ilu: 183 iterations) are not.

Pinned by tests/test_oracle_precond.py (dense restatements, converged
published rows).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular


def block_jacobi(A, block_size: int = 2):
    A = sp.csr_matrix(A)
    n = A.shape[0]
    invs = []
    for s in range(0, n, block_size):
        e = min(s + block_size, n)
        B = np.zeros((block_size, block_size))
        B[: e - s, : e - s] = A[s:e, s:e].toarray()
        B = B + np.eye(block_size) * 1e-12
        B = B[: e - s, : e - s]
        invs.append(np.linalg.inv(B))

    def apply(x):
        y = np.empty_like(x)
        for k, s in enumerate(range(0, n, block_size)):
            e = min(s + block_size, n)
            y[s:e] = invs[k] @ x[s:e]
        return y
    return apply


def ssor(A, omega: float = 1.0):
    A = sp.csr_matrix(A)
    d = A.diagonal()
    d_omega = d / omega
    lower = (sp.tril(A, -1) + sp.diags(d_omega)).tocsr()

    def apply(x):
        y = spsolve_triangular(lower, x, lower=True)
        return (d * y) / d_omega
    return apply


def ilu(A):
    A = sp.csr_matrix(A)
    d = A.diagonal()
    dg = np.where(d == 0, 1.0, d)
    L = sp.tril(A, -1).tocsr()
    L.data = L.data / dg[L.indices]         # val / diag[col], as :196
    unit_lower = (L + sp.eye(A.shape[0])).tocsr()
    upper = (sp.triu(A, 1) + sp.diags(dg)).tocsr()

    def apply(x):
        y = spsolve_triangular(unit_lower, x, lower=True, unit_diagonal=True)
        return spsolve_triangular(upper, y, lower=False)
    return apply


def make(kind: str, A):
    if kind == "block_jacobi":
        return block_jacobi(A)
    if kind == "ssor":
        return ssor(A)
    if kind == "ilu":
        return ilu(A)
    raise ValueError(f"unknown preconditioner {kind!r}")
