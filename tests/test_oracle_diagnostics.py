"""Oracle diagnostics pinned by closed forms and hand-computed cases.

* integrate_2d_domain: for a P1 field the 3-point rule with weights 1/3 times
  |det J| gives twice the integral of the interpolant.  Checked on constant
  and linear fields (exact integrals over the domain) and, for random nodal
  fields on a perturbed mesh, against the P1 integral from shoelace areas.
* element_size: structured N x N meshes have legs 1/N (the smallest edge).
* peclet_courant: a single triangle computed by hand, and a three-triangle
  mesh where only element 2 holds the maxima, which separates the reference's
  (max(Pe_0, Cr_0), max(Pe_1, Cr_1)) result from the mesh maxima.
"""
import numpy as np
import pytest

from oracle import diagnostics as D
from oracle import mesh as omesh


def _perturbed(cells, seed=0):
    pts, tri = omesh.structured_tri(cells, 1.0)
    pts = pts.copy()
    rng = np.random.default_rng(seed)
    interior = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts[interior] += rng.uniform(-0.15, 0.15, size=(interior.sum(), 2)) / cells   # keeps every triangle upright
    return pts, np.asarray(tri)


@pytest.mark.parametrize("cells", [1, 4, 13])
def test_integral_closed_forms(cells):
    pts, tri = _perturbed(cells)
    x, y = pts[:, 0], pts[:, 1]
    assert D.integrate_2d_domain(np.ones(len(pts)), pts, tri) == pytest.approx(2.0, rel=1e-14)
    # int_0^1 int_0^1 (3x + 2y + 1) = 1.5 + 1 + 1
    assert D.integrate_2d_domain(3 * x + 2 * y + 1, pts, tri) == pytest.approx(2 * 3.5, rel=1e-13)
    # scaling the domain by L scales the integral by L^2
    assert D.integrate_2d_domain(np.ones(len(pts)), 15.24 * pts, tri) == pytest.approx(2 * 15.24 ** 2, rel=1e-13)


def test_integral_matches_shoelace_p1():
    pts, tri = _perturbed(9, seed=3)
    f = np.random.default_rng(1).normal(size=len(pts))
    P = pts[tri]
    area = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
                        - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    exact = np.sum(area * f[tri].mean(axis=1))
    assert D.integrate_2d_domain(f, pts, tri) == pytest.approx(2 * exact, rel=1e-12)


def test_mass_balance_error():
    assert np.allclose(D.mass_balance_error([1.0, 2.0], [1.0, 1.6]), [0.0, 25.0])


@pytest.mark.parametrize("cells", [2, 7])
def test_element_size_structured(cells):
    pts, tri = omesh.structured_tri(cells, 1.0)
    # the generator rounds coordinates as the reference CSVs do (about 1e-12)
    assert np.allclose(D.element_size(pts, tri), 1.0 / cells, rtol=1e-11, atol=0)


def test_peclet_courant_hand_computed():
    pts = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 1.0]])      # edges 2, sqrt(5), 1 -> h = 1
    tri = np.array([[0, 1, 2]])
    qx, qz = np.array([1.0, 2.0, 3.0]), np.array([0.5, -4.0, 1.0])
    th = np.array([0.5, 0.5, 0.25])
    dxx, dzz = np.array([1.0, 1.0, 2.0]), np.array([1.0, 0.5, 1.0])
    # v_x = (2, 4, 12), v_z = (1, -8, 4); Pe_x = (1, 2, 3), Pe_z = (0.5, 8, 2);
    # Cr_x = (0.2, 0.4, 1.2), Cr_z = (0.1, 0.8, 0.4)
    pe, cr = D.element_peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1)
    assert pe[0] == pytest.approx(8.0) and cr[0] == pytest.approx(1.2)
    assert D.peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1, all_elements=True) == pytest.approx((8.0, 1.2))
    # one element: index 1 clamps to 0, both results are max(Pe_0, Cr_0)
    assert D.peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1) == pytest.approx((8.0, 8.0))


def test_peclet_courant_reference_collection():
    # three unit right triangles side by side; nodes 0-3 on y=0, 4-7 on y=1
    pts = np.array([[x, y] for y in (0.0, 1.0) for x in (0.0, 1.0, 2.0, 3.0)])
    tri = np.array([[0, 1, 4], [1, 2, 5], [2, 3, 6]])
    n = len(pts)
    qx, qz = np.zeros(n), np.zeros(n)
    qx[[0, 1, 4]] = [0.1, 0.2, 0.1]            # element 0: Pe = 0.1, Cr = 0.02
    qx[5] = 0.4                               # element 1 (nodes 1, 2, 5): Pe = 0.2, Cr = 0.04
    qx[[3, 6]] = [6.0, 2.0]                   # element 2: Pe = 3, Cr = 0.6
    th, dxx, dzz = np.ones(n), np.ones(n), np.ones(n)
    pe, cr = D.element_peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1)
    assert np.allclose(pe, [0.1, 0.2, 3.0]) and np.allclose(cr, [0.02, 0.04, 0.6])
    assert D.peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1) == pytest.approx((0.1, 0.2))
    assert D.peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.1, all_elements=True) == pytest.approx((3.0, 0.6))


def test_peclet_nan_propagates():
    pts = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    tri = np.array([[0, 1, 2]])
    z = np.zeros(3)
    with np.errstate(divide="ignore", invalid="ignore"):
        pe, cr = D.peclet_courant(pts, tri, z, z, z, np.ones(3), np.ones(3), 0.1, all_elements=True)
    assert np.isnan(pe) and np.isnan(cr)
