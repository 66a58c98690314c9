"""Oracle diagnostics (FP64, NumPy).

TEST INFRASTRUCTURE ONLY (see the package docstring).

* ``integrate_2d_domain`` restates
  src/solvers/.ipynb_checkpoints/richards_mass_balance-checkpoint.py:25-65:
  points (1/6,1/6), (2/3,1/6), (1/6,2/3), weights 1/3 each, times |det J|.
  The weights sum to 1 on a reference triangle of area 1/2, so the result is
  twice the integral (for the ratios the mass balance forms, the factor
  cancels).
* ``mass_balance_error`` restates :300-323: |1 - M_num / M_exact| * 100 of
  the stored domain integrals.
* ``element_size`` restates src/models/transport_utilities.py:196-204.
* ``peclet_courant`` restates transport_utilities.py:206-242, including the
  reference's collection of the vmapped (Pe[E], Cr[E]) pair with
  ``[pc[0] for pc in Pe_Cr]`` / ``[pc[1] for pc in Pe_Cr]``: iterating the
  pair yields the two arrays, so the reference returns
  (max(Pe_0, Cr_0), max(Pe_1, Cr_1)).  ``all_elements=True`` gives the
  maxima over the mesh that the docstring describes.

Pinned by tests/test_oracle_diagnostics.py: closed-form integrals of
constant and linear fields, a hand-computed single-triangle Peclet / Courant
case, and brute-force loops on small meshes.
"""
from __future__ import annotations

import numpy as np

_GP = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
_W = np.array([1 / 3, 1 / 3, 1 / 3])


def integrate_2d_domain(f, points, triangles) -> float:
    f = np.asarray(f, dtype=np.float64)
    P = np.asarray(points, dtype=np.float64)[np.asarray(triangles)]          # [E,3,2]
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]], axis=-1)            # [[x1-x0, x2-x0], [y1-y0, y2-y0]]
    det = np.abs(J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0])
    N = np.stack([1 - _GP[:, 0] - _GP[:, 1], _GP[:, 0], _GP[:, 1]], axis=1)  # [q, a]
    fv = f[np.asarray(triangles)]                                           # [E,3]
    tri = (fv @ N.T) @ _W                                                   # sum_q w_q f(x_q)
    return float(np.sum(tri * det))


def mass_balance_error(m_num, m_exact):
    return np.abs(1.0 - np.asarray(m_num) / np.asarray(m_exact)) * 100.0


def element_size(points, triangles) -> np.ndarray:
    V = np.asarray(points, dtype=np.float64)[np.asarray(triangles)]
    edges = V - np.roll(V, 1, axis=1)
    return np.min(np.sqrt(np.sum(edges ** 2, axis=2)), axis=1)


def element_peclet_courant(points, triangles, qx, qz, theta, dxx, dzz, dt):
    """Per-element (Pe[E], Cr[E]) of transport_utilities.py:220-235."""
    tri = np.asarray(triangles)
    h = element_size(points, tri)[:, None]
    vx = np.asarray(qx)[tri] / np.asarray(theta)[tri]
    vz = np.asarray(qz)[tri] / np.asarray(theta)[tri]
    with np.errstate(divide="ignore", invalid="ignore"):
        pex = (np.abs(vx) * h) / (2 * np.asarray(dxx)[tri])
        pez = (np.abs(vz) * h) / (2 * np.asarray(dzz)[tri])
        crx = (np.abs(vx) * dt) / h
        crz = (np.abs(vz) * dt) / h
    pe = np.maximum(np.max(pex, axis=1), np.max(pez, axis=1))
    cr = np.maximum(np.max(crx, axis=1), np.max(crz, axis=1))
    return pe, cr


def peclet_courant(points, triangles, qx, qz, theta, dxx, dzz, dt, all_elements: bool = False):
    pe, cr = element_peclet_courant(points, triangles, qx, qz, theta, dxx, dzz, dt)
    if all_elements:
        return float(np.max(pe)), float(np.max(cr))
    i1 = min(1, len(pe) - 1)               # JAX clamps out-of-range indices
    return float(np.max([pe[0], cr[0]])), float(np.max([pe[i1], cr[i1]]))
