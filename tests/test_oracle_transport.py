"""Pins of the oracle's transport operators (CPU).  The reference commits no
numeric output of the solute / nitrogen paths, so each check compares the
oracle with mathematics that a plausible slip (wrong sign, index, weight or
transposed operand) would violate."""
import numpy as np
import pytest

from oracle import mesh as omesh
from oracle import transport as T
from oracle.soil import van_genuchten

PHI = np.array([3.6, 0.43, 0.078, 1.56, 1 - 1 / 1.56, 0.2496])


def _mesh(n=6, jitter=0.0, seed=0):
    pts, tri = omesh.structured_tri(n, 1.0)
    if jitter:
        rng = np.random.default_rng(seed)
        inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
        pts = pts.copy()
        pts[inner] += jitter / n * rng.uniform(-1, 1, size=(inner.sum(), 2))
    return pts, np.asarray(tri)


def test_linear_head_gives_exact_fluxes():
    """h = a x + c z + d: every element gradient is (a, c), so q = -K (a, c + 1)
    at every node, and D follows the closed form."""
    pts, tri = _mesh(5, jitter=0.3)
    a, c, d = 0.7, -1.3, -2.0
    h = a * pts[:, 0] + c * pts[:, 1] + d
    DL, DT, Dm = 0.5, 0.1, 0.01
    qx, qz, aq, Dxx, Dxz, Dzz, th = T.fluxes_dispersion(pts, tri, h, PHI, DL, DT, Dm)
    _, K, theta = van_genuchten(h, PHI)
    np.testing.assert_allclose(qx, -K * a, rtol=1e-12)
    np.testing.assert_allclose(qz, -K * (c + 1.0), rtol=1e-12)
    tau = theta ** (7 / 3) / PHI[1] ** 2
    q2 = (K * a) ** 2 + (K * (c + 1)) ** 2
    np.testing.assert_allclose(Dxx, (DL * (K * a) ** 2 + DT * (K * (c + 1)) ** 2) / np.sqrt(q2) + theta * Dm * tau,
                               rtol=1e-12)
    np.testing.assert_allclose(Dzz, (DL * (K * (c + 1)) ** 2 + DT * (K * a) ** 2) / np.sqrt(q2) + theta * Dm * tau,
                               rtol=1e-12)
    np.testing.assert_allclose(Dxz, (DL - DT) * (K * a) * (K * (c + 1)) / np.sqrt(q2), rtol=1e-12)
    np.testing.assert_array_equal(th, theta)


def _fields(pts, seed=1):
    rng = np.random.default_rng(seed)
    n = len(pts)
    return [rng.uniform(0.1, 0.4, n), rng.normal(size=n), rng.normal(size=n), rng.uniform(0.1, 1, n),
            rng.normal(scale=0.1, size=n), rng.uniform(0.1, 1, n)]


def _exact_element(X, th, qx, qz, dxx, dxz, dzz):
    """Closed forms: P1 integrals int N_a = A/3, int N_a N_b = A (1 + d_ab) / 12,
    int N_a N_b N_c = 2A * (multiplicity factorials) / 5! ."""
    x, y = X[:, 0], X[:, 1]
    A2 = (x[1] - x[0]) * (y[2] - y[0]) - (y[1] - y[0]) * (x[2] - x[0])   # 2A (signed)
    bx = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / A2
    bz = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / A2
    A = A2 / 2
    mean = lambda f: f.sum() / 3
    disp = -A * (mean(dxx) * np.outer(bx, bx) + mean(dxz) * (np.outer(bx, bz) + np.outer(bz, bx))
                 + mean(dzz) * np.outer(bz, bz))
    M2 = A / 12 * (np.ones((3, 3)) + np.eye(3))          # int N_a N_b
    adv = -(np.outer(M2 @ qx, bx) + np.outer(M2 @ qz, bz))
    return disp + adv, M2


def test_element_operators_against_closed_forms():
    """Dispersion (linear integrand) and advection (quadratic) are integrated
    exactly by the 3-point rule: compare with the closed forms; the mass with
    constant theta likewise; the mass total equals -int theta exactly."""
    pts, tri = _mesh(4, jitter=0.4, seed=3)
    th, qx, qz, dxx, dxz, dzz = _fields(pts)
    st, ms, _ = T.element_operators(pts, tri, th, qx, qz, dxx, dxz, dzz)
    for e in range(len(tri)):
        idx = tri[e]
        ex, M2 = _exact_element(pts[idx], th[idx], qx[idx], qz[idx], dxx[idx], dxz[idx], dzz[idx])
        np.testing.assert_allclose(st[e], ex, rtol=1e-9, atol=1e-12 * np.abs(ex).max())
        A = M2.sum()                  # sum_ab int N_a N_b = int 1 = A
        assert ms[e].sum() == pytest.approx(-A * th[idx].sum() / 3, rel=1e-9)
    thc = np.full(len(pts), 0.3)
    _, msc, _ = T.element_operators(pts, tri, thc, qx, qz, dxx, dxz, dzz)
    for e in range(len(tri)):
        idx = tri[e]
        _, M2 = _exact_element(pts[idx], thc[idx], qx[idx], qz[idx], dxx[idx], dxz[idx], dzz[idx])
        np.testing.assert_allclose(msc[e], -0.3 * M2, rtol=1e-9)


def test_stiffness_annihilates_constants_and_reaction_and_source_totals():
    pts, tri = _mesh(5, jitter=0.3, seed=4)
    th, qx, qz, dxx, dxz, dzz = _fields(pts, 2)
    st, _, _ = T.element_operators(pts, tri, th, qx, qz, dxx, dxz, dzz)
    S = T.coo_matrix(tri, len(pts), st)
    assert np.abs(S @ np.ones(len(pts))).max() <= 1e-12 * np.abs(S).max()
    # reaction: + mu int theta N N^T adds mu * int theta N_a to each row sum
    mu = 0.37
    st2, ms2, _ = T.element_operators(pts, tri, th, qx, qz, dxx, dxz, dzz, mu=mu)
    np.testing.assert_allclose((st2 - st).sum(axis=2), -mu * ms2.sum(axis=2), rtol=1e-12)
    # source of a coupled species: sum_a src_a = -mu int theta c (quadratic: exact)
    c = np.random.default_rng(9).uniform(0, 2, len(pts))
    _, _, src = T.element_operators(pts, tri, th, qx, qz, dxx, dxz, dzz, mu_src=mu, c_src=c)
    for e in range(len(tri)):
        idx = tri[e]
        _, M2 = _exact_element(pts[idx], th[idx], qx[idx], qz[idx], dxx[idx], dxz[idx], dzz[idx])
        assert src[e].sum() == pytest.approx(-mu * th[idx] @ M2 @ c[idx], rel=1e-9)


def test_edge_terms_closed_forms():
    pts, tri = _mesh(8)
    top = np.nonzero(np.abs(pts[:, 1] - 1.0) < 1e-12)[0]
    n = len(pts)
    L = pts[top[-1], 0] - pts[top[0], 0]
    b = T.neumann(np.zeros(n), -0.05, top, pts)
    assert b.sum() == pytest.approx(0.05 * L, rel=1e-13)
    assert b[top[1]] == pytest.approx(0.05 * (pts[top[2], 0] - pts[top[0], 0]) / 2, rel=1e-12)
    K = np.full(n, 0.2)
    bf = T.freedrain_water(np.zeros(n), K, top, pts)
    assert bf.sum() == pytest.approx(-0.2 * L, rel=1e-13)
    A0 = T.coo_matrix(tri, n, np.zeros((len(tri), 3, 3)))
    A, bc = T.cauchy(A0, np.zeros(n), -0.05, 50.0, top, pts)
    assert A.sum() == pytest.approx(-0.05 * L, rel=1e-13)
    assert bc.sum() == pytest.approx(-0.05 * 50.0 * L, rel=1e-13)
    # consistent 1D mass: diagonal h/3, off-diagonal h/6 (2-point Gauss is exact)
    hseg = pts[top[1], 0] - pts[top[0], 0]
    assert A[top[0], top[0]] == pytest.approx(-0.05 * hseg / 3, rel=1e-12)
    assert A[top[0], top[1]] == pytest.approx(-0.05 * hseg / 6, rel=1e-12)
    th = np.full(n, 0.4)
    An = T.freedrain_nitrogen(A0, K, th, top, pts)
    assert An.sum() == pytest.approx(-(0.2 / 0.4) * L, rel=1e-13)
    # reversed list order flips the sign of every edge length (the reference's x_J - x_I)
    br = T.neumann(np.zeros(n), -0.05, top[::-1], pts)
    np.testing.assert_allclose(br, -b, rtol=1e-13, atol=1e-18)


def test_constant_concentration_is_steady_without_sources():
    """K 1 = 0 and b = M c / dt: with no edge terms and no reactions a constant
    concentration solves the system exactly (solute and lumped forms)."""
    pts, tri = _mesh(6, jitter=0.2, seed=5)
    n = len(pts)
    h = -1.0 - 0.5 * pts[:, 1] + 0.2 * np.sin(3 * pts[:, 0])
    fields = T.fluxes_dispersion(pts, tri, h, PHI, 0.5, 0.1, 0.0)
    c = np.full(n, 0.7)
    A, b = T.solute_system(pts, tri, fields, c, 1e-2, [], -0.1, 1.0)
    np.testing.assert_allclose(A @ c, b, rtol=1e-11, atol=1e-14)
    prm = dict(T.NITROGEN_DEFAULT, mu1=0.0, mu2=0.0, mu3=0.0, rho=0.0)
    K = van_genuchten(h, PHI)[1]
    A3, b3 = T.nitrogen_system(pts, tri, 3, fields, prm, c, np.zeros(n), 1e-2, K, [], [], -0.05, 0.0)
    np.testing.assert_allclose(A3 @ c, b3, rtol=1e-11, atol=1e-14)


def test_lumping_preserves_row_sums():
    pts, tri = _mesh(4, jitter=0.3, seed=8)
    th = np.random.default_rng(2).uniform(0.1, 0.4, len(pts))
    _, ms, _ = T.element_operators(pts, tri, th, *(np.zeros(len(pts)),) * 5)
    M = T.coo_matrix(tri, len(pts), ms)
    ML = T.lump(M)
    np.testing.assert_allclose(ML.diagonal(), np.asarray(M.sum(axis=1)).ravel(), rtol=1e-15)
    assert ML.nnz == len(pts)


@pytest.mark.slow
def test_nitrogen_run_is_finite_and_inflow_enters():
    pts, tri = omesh.structured_tri(10, 1.0)
    r = T.nitrogen_run(pts, np.asarray(tri), PHI, 3, 1e-3)
    for k in ("pressure_head", "NH4", "NO2", "NO3"):
        assert np.isfinite(r[k]).all()
    top = np.nonzero(np.abs(pts[:, 1] - 1) < 1e-10)[0]
    assert r["NH4"][-1][top].mean() > 0.0          # NH4 enters through the Cauchy inlet
