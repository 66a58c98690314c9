"""Entry points with the reference's names, over the device context.

* ``solve_system(matrix, rhs, x0, solver_config)`` -- src/solvers/solver_types.py:7-68
  (dispatch on solver_type / precond_type; ValueError for an unknown solver).
* ``assemble_global_matrices_sparse_re(...)`` -- src/numerics/assembly_re_v2.py:62-166
  and assembly_re_3D.py:72-170, returning the device CSR matrix and b.
* ``van_genuchten_model`` / ``exponential_model`` -- src/models/van_genuchten.py:9-47,
  src/models/exponential.py:10-42, vectorised over device tensors.

``CSRMatrix`` is the device analogue of the reference's BCOO: the context's
fixed pattern plus a value tensor (and the Jacobi inverse diagonal computed by
the assembly).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

from .device import Context


@dataclass
class CSRMatrix:
    ctx: Context
    values: torch.Tensor
    dinv: Optional[torch.Tensor] = None

    @property
    def shape(self):
        return (self.ctx.n, self.ctx.n)

    def __matmul__(self, x):
        return self.ctx.spmv(self.values, x)

    def to_scipy(self):
        import scipy.sparse as sp
        rp, ci = self.ctx.pattern()
        return sp.csr_matrix((self.values.cpu().numpy(), ci, rp), shape=self.shape)


def solve_system(matrix: CSRMatrix, rhs, x0, solver_config):
    """Returns (solution, convergence) like the reference: GMRES gives 0 / -1
    (NaN) in ``convergence``; CG and BiCGSTAB give None."""
    st = solver_config.solver_type
    pt = solver_config.precond_type
    if st not in ("gmres", "bicgstab", "cg", "direct"):
        raise ValueError(f"Unsupported solver type: {st}")
    if st == "direct":   # solver_types.py:60-63 (dense LU there, banded LU here); preconditioner ignored
        return matrix.ctx.direct(matrix.values, rhs), None
    if pt in ("block_jacobi", "ssor", "ilu"):
        x, info = matrix.ctx.krylov_pc(st, pt, matrix.values, rhs, x0, tol=solver_config.tol,
                                       maxiter=solver_config.maxiter, restart=solver_config.restart)
        return x, (info["code"] if st == "gmres" else None)
    if pt not in (None, "none", "jacobi"):
        raise ValueError(f"Unsupported preconditioner: {pt}")
    dinv = matrix.dinv if pt == "jacobi" else None
    if pt == "jacobi" and dinv is None:
        raise ValueError("Jacobi needs the assembled inverse diagonal (matrix.dinv)")
    x, info = matrix.ctx.krylov(st, matrix.values, rhs, x0, dinv=dinv, tol=solver_config.tol,
                                maxiter=solver_config.maxiter, restart=solver_config.restart)
    return x, (info["code"] if st == "gmres" else None)


def assemble_global_matrices_sparse_re(ctx: Context, theta_m, theta_n, pressure_head_m, conductivity_m,
                                       capacity_m, dt, Ks_elem=None):
    """Assembly of the mixed-form Richards system on the context's mesh, with the
    context's Dirichlet rows (``ctx.set_dirichlet``) applied."""
    A, b, dinv = ctx.assemble_richards(pressure_head_m, capacity_m, conductivity_m, theta_m, theta_n, dt,
                                       Ks_elem=Ks_elem)
    return CSRMatrix(ctx, A, dinv), b


def van_genuchten_model(ctx: Context, h, phi):
    """(C, K, theta) of van Genuchten-Mualem, phi = [alpha, theta_s, theta_r, n, m, Ks]."""
    return ctx.soil("vg", phi, h)


def exponential_model(ctx: Context, h, phi):
    """(C, K, theta) of Gardner, phi = [alpha0, theta_s, theta_r, Ks]."""
    return ctx.soil("gardner", phi, h)
