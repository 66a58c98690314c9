"""Diagnostics on the GPU: domain integrals for mass balance and the element
Peclet / Courant numbers of the transport fields.

Mirrors, with the mesh held by a ``Context`` instead of (points, triangles):

* ``integrate_2d_domain(ctx, f)`` -
  src/solvers/.ipynb_checkpoints/richards_mass_balance-checkpoint.py:25-65
  (3-point rule, weights 1/3, times |det J|: twice the integral, as there);
* ``mass_balance_error(num, exact)`` - the same file, :300-323;
* ``calculate_element_size(ctx)`` - src/models/transport_utilities.py:196-204;
* ``calculate_peclet_courant(ctx, qx, qz, theta, Dxx, Dzz, dt)`` -
  transport_utilities.py:206-242.  By default it returns what the reference
  returns, (max(Pe_0, Cr_0), max(Pe_1, Cr_1)): the reference collects its
  vmapped (Pe[E], Cr[E]) pair by iterating it (:237-240).  Pass
  ``all_elements=True`` for the maxima over the mesh.

Every function takes device tensors (or arrays, copied to the device) and
runs one reduction kernel pass on the context's stream (libwspm).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from .device import Context


def _ctx(ctx_or_points, triangles=None) -> Tuple[Context, bool]:
    if isinstance(ctx_or_points, Context):
        return ctx_or_points, False
    return Context(np.asarray(ctx_or_points), np.asarray(triangles)), True


def integrate_2d_domain(f, ctx_or_points, triangles=None) -> float:
    ctx, own = _ctx(ctx_or_points, triangles)
    try:
        return ctx.domain_integral(f)
    finally:
        if own:
            ctx.close()


def mass_balance_error(m_num, m_exact) -> np.ndarray:
    return np.abs(1.0 - np.asarray(m_num, dtype=np.float64) / np.asarray(m_exact, dtype=np.float64)) * 100.0


def calculate_element_size(ctx_or_points, triangles=None):
    ctx, own = _ctx(ctx_or_points, triangles)
    try:
        return ctx.element_size()
    finally:
        if own:
            ctx.close()


def calculate_peclet_courant(ctx_or_points, triangles, water_flux_x, water_flux_z, theta, Dispersion_xx,
                             Dispersion_zz, dt: float, *, all_elements: bool = False) -> Tuple[float, float]:
    ctx, own = _ctx(ctx_or_points, triangles)
    try:
        return ctx.peclet_courant(water_flux_x, water_flux_z, theta, Dispersion_xx, Dispersion_zz, dt,
                                  all_elements=all_elements)
    finally:
        if own:
            ctx.close()
