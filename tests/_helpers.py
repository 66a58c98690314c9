"""Shared test helpers: product problem -> oracle problem, comparisons."""
from __future__ import annotations

import numpy as np

from oracle import picard as opicard


def to_oracle(pb) -> opicard.Problem:
    return opicard.Problem(
        points=pb.points, elems=pb.elements, soil_model=pb.soil_model, phi=tuple(pb.phi),
        bc_lists=[(np.asarray(n), np.asarray(v)) for n, v in pb.bc_lists], h0=pb.h0, dt0=pb.dt0,
        tmax=pb.tmax, nsteps=pb.nsteps, adaptive=pb.nsteps is None, picard_tol=pb.picard_tol,
        picard_maxit=pb.picard_maxit, picard_mode=pb.picard_mode, theta_at_start=pb.theta_at_start,
        solver=pb.solver, precond=pb.precond, tol=pb.tol, maxiter=pb.maxiter, restart=pb.restart,
        Ks_elem=pb.Ks_elem, m_it=pb.m_it, M_it=pb.M_it, lambda_amp=pb.lambda_amp, lambda_red=pb.lambda_red)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def row_scaled_err(vals_a, vals_b, indptr):
    """max_i max_k |a_k - b_k| / max_k |b_k| over row i (entries that cancel to
    ~0 in one summation order are compared at the row's scale)."""
    d = np.abs(np.asarray(vals_a) - np.asarray(vals_b))
    mag = np.abs(np.asarray(vals_b))
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    rmax = np.zeros(len(indptr) - 1)
    np.maximum.at(rmax, rows, mag)
    dmax = np.zeros(len(indptr) - 1)
    np.maximum.at(dmax, rows, d)
    return float(np.max(dmax / np.maximum(rmax, 1e-300)))
