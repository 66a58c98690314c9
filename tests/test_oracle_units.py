"""Pin the oracle's pieces by mathematics (closed forms, identities, brute force)
and by the reference's mesh fixtures.  Each check is chosen so a plausible
mistake (dropped term, wrong sign or index, transposed operand) fails it.

The van Genuchten model rests on these checks alone: the reference commits
no numeric output of its van Genuchten path (its published runs use the
Gardner model), so C, K and theta are pinned here by identities (C = dtheta/dh
by central differences, the saturated limits, the Mualem form at a point
computed by hand) and, for the GPU kernel, by 40-digit arithmetic in
test_gpu_kernels.py -- not by reference vectors."""
import hashlib
import json

import numpy as np
import pytest

from oracle import assembly as oasm
from oracle import krylov as okr
from oracle import mesh as omesh
from oracle import soil as osoil

VG = (0.1, 0.40, 0.06, 2.0, 0.5, 1e-3)
VG_REF = (3.6, 0.43, 0.078, 1.56, 1 - 1 / 1.56, 0.2496)
GARDNER = (0.164, 0.45, 0.15, 0.1)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ----------------------------------------------------------------- meshes ----
def test_generators_equal_reference_fixtures(golden):
    ck = json.loads((golden / "mesh_checksums.json").read_text())
    for name, d in ck.items():
        gen = omesh.structured_tri if d["kind"] == "tri" else omesh.structured_tet
        p, t = gen(d["cells"], d["length"])
        assert list(p.shape) == d["points_shape"] and list(t.shape) == d["elems_shape"], name
        assert sha(p) == d["points_sha256"], name
        assert sha(t) == d["elems_sha256"], name


def test_nnz_formula():
    # structural nnz = n + 2 * edges (t25 -> 4,526; 20^3 Kuhn -> 128,581)
    p, t = omesh.structured_tri(25, 15.24)
    assert omesh.csr_pattern(t, len(p))[1].size == 4526
    p, t = omesh.structured_tet(20, 15.24)
    assert omesh.csr_pattern(t, len(p))[1].size == 128581


def test_boundary_sets():
    p, _ = omesh.structured_tri(25, 15.24)
    top, bot = omesh.boundary_nodes_2d(p, "Test1")
    assert top.size == 26 and bot.size == 26 + 2 * 25
    top3, bot3 = omesh.boundary_nodes_2d(p, "Test3")
    assert bot3.size == 26
    p3, _ = omesh.structured_tet(20, 15.24)
    t3, b3 = omesh.boundary_nodes_3d(p3)
    assert t3.size == 21 * 21 and b3.size == 2041


# ------------------------------------------------------------------- soil ----
@pytest.mark.parametrize("phi", [VG, VG_REF])
def test_vg_capacity_is_dtheta_dh(phi):
    """C = d theta / dh, checked by central differences in extended precision."""
    h = -np.logspace(-2, 3, 60)
    C, K, th = osoil.van_genuchten(h, phi)
    eps = 1e-6 * np.abs(h)
    _, _, tp = osoil.van_genuchten(h + eps, phi)
    _, _, tm = osoil.van_genuchten(h - eps, phi)
    fd = (tp - tm) / (2 * eps)
    assert np.allclose(C, fd, rtol=1e-6)


def test_vg_limits_and_mualem():
    C, K, th = osoil.van_genuchten(np.array([0.0, 1.0, -1e-12]), VG)
    assert np.allclose(th[:2], 0.40) and np.allclose(K[:2], 1e-3) and np.all(C[:2] == 0)
    assert abs(th[2] - 0.40) < 1e-12
    # Mualem: K = Ks sqrt(Se) (1 - (1 - Se^(1/m))^m)^2 with Se from theta; check at one point by hand
    alpha, ts, tr, n, m, ks = VG
    h = -50.0
    se = (1 + (alpha * 50.0) ** n) ** (-m)
    kk = ks * np.sqrt(se) * (1 - (1 - se ** (1 / m)) ** m) ** 2
    _, K1, th1 = osoil.van_genuchten(np.array([h]), VG)
    assert np.isclose(K1[0], kk, rtol=1e-13) and np.isclose(th1[0], tr + (ts - tr) * se, rtol=1e-14)


def test_gardner_identities():
    h = np.linspace(-15.24, 0.0, 50)
    C, K, th = osoil.exponential(h, GARDNER)
    a0, ts, tr, ks = GARDNER
    assert np.allclose(C, a0 * (th - tr), rtol=1e-14)
    assert np.allclose(K, ks * (th - tr) / (ts - tr), rtol=1e-14)
    C2, K2, th2 = osoil.exponential(np.array([0.5]), GARDNER)
    assert C2[0] == 0 and K2[0] == ks and th2[0] == ts


# --------------------------------------------------------------- elements ----
def _one(points, elems, K, C, th_m, th_n, h, dt=1.0):
    return oasm.element_matrices(points, elems, th_m, th_n, h, K, C, dt)


def test_triangle_closed_form():
    """P1 triangle: stiffness = A*Kbar*G^T G (row sums 0), lumped mass = A/3,
    gravity f_i = -A*Kbar*dN_i/dy."""
    pts = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    el = np.array([[0, 1, 2]])
    K = np.array([2.0, 3.0, 5.0])
    zero = np.zeros(3)
    Ae, be = _one(pts, el, K, zero, zero, zero, zero)
    J = np.array([pts[1] - pts[0], pts[2] - pts[0]])
    area = 0.5 * np.linalg.det(J)
    G = np.linalg.inv(J) @ np.array([[-1, 1, 0], [-1, 0, 1.0]])
    kbar = K.mean()
    assert np.allclose(Ae[0], area * kbar * G.T @ G, rtol=1e-13)
    assert np.allclose(Ae[0].sum(axis=1), 0, atol=1e-12)
    assert np.allclose(be[0], -area * kbar * G[1], rtol=1e-13)
    # mass term: C on the diagonal, lumped A/3 (15-digit literals -> 1e-14)
    Cc = np.array([1.0, 2.0, 4.0])
    Ae2, be2 = _one(pts, el, zero, Cc, np.ones(3), zero, zero, dt=0.5)
    assert np.allclose(np.diag(Ae2[0]), area / 3 * Cc / 0.5, rtol=1e-13)
    assert np.allclose(be2[0], -area / 3 / 0.5, rtol=1e-13)


def test_tet_closed_form_with_x6_weights():
    """Tet rule with weights 0.25 (sum 1): every term is 6x the exact integral."""
    pts = np.array([[0, 0, 0], [1.0, 0.1, 0], [0.2, 1.1, 0.1], [0.1, 0.3, 0.9]])
    el = np.array([[0, 1, 2, 3]])
    K = np.array([1.0, 2.0, 3.0, 4.0])
    zero = np.zeros(4)
    Ae, be = _one(pts, el, K, zero, zero, zero, zero)
    J = pts[1:] - pts[0]
    vol = np.linalg.det(J) / 6
    G = np.linalg.inv(J) @ np.hstack([-np.ones((3, 1)), np.eye(3)])
    # conductivity weight of vertex c: sum_q 0.25 N_c(q) with the rule's literals (~0.25 each)
    ck = np.array([0.25 * sum((1 - q.sum()) if c == 0 else q[c - 1] for q in oasm.TET_Q) for c in range(4)])
    assert np.allclose(ck, 0.25, rtol=1e-14)
    kbar = K @ ck
    assert np.allclose(Ae[0], 6 * vol * kbar * G.T @ G, rtol=1e-12)
    assert np.allclose(be[0], -6 * vol * kbar * G[2], rtol=1e-12)
    Ae2, _ = _one(pts, el, zero, np.ones(4), zero, zero, zero)
    assert np.allclose(np.diag(Ae2[0]), 1.5 * vol, rtol=1e-12)


def test_assembly_equals_dense_bruteforce():
    p, t = omesh.structured_tri(4, 1.0)
    n = len(p)
    rng = np.random.default_rng(0)
    args = [rng.uniform(0.1, 1.0, n) for _ in range(5)]
    th_m, th_n, h, K, C = args
    pat = oasm.Pattern(t, n)
    vals, b = oasm.assemble(pat, p, th_m, th_n, h, K, C, 0.3)
    A = pat.to_scipy(vals).toarray()
    Ad, bd = oasm.dense_bruteforce(p, t, n, th_m, th_n, h, K, C, 0.3)
    assert np.allclose(A, Ad, rtol=1e-14, atol=1e-15) and np.allclose(b, bd, rtol=1e-14)
    # stiffness part is symmetric; Dirichlet makes identity rows, later list wins
    vals0, _ = oasm.assemble(pat, p, th_m, th_n, h, K, np.zeros(n), 0.3)
    A0 = pat.to_scipy(vals0).toarray()
    assert np.allclose(A0, A0.T, atol=1e-14)
    v2, b2 = oasm.apply_dirichlet(pat, vals, b, [(np.array([0, 1]), np.array([7.0, 8.0])),
                                                 (np.array([1]), np.array([9.0]))])
    A2 = pat.to_scipy(v2).toarray()
    assert np.array_equal(A2[1], np.eye(n)[1]) and b2[0] == 7.0 and b2[1] == 9.0


# ----------------------------------------------------------------- Krylov ----
def _spd_like(n=60, seed=0, nonsym=False):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * 0.1
    A = M @ M.T + np.diag(rng.uniform(1, 3, n))
    if nonsym:
        A = A + 0.05 * rng.standard_normal((n, n))
    return A, rng.standard_normal(n)


@pytest.mark.parametrize("method", ["cg", "bicgstab", "gmres"])
@pytest.mark.parametrize("jac", [False, True])
def test_krylov_solves(method, jac):
    A, b = _spd_like(nonsym=(method != "cg"))
    dinv = 1.0 / np.diag(A) if jac else None
    x, info = okr.solve(method, lambda v: A @ v, b, np.zeros_like(b), dinv, 1e-12, 0.0, 500, 30)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-9


def test_gmres_breakdown_exact_in_few_steps():
    """A = 2 I: the first Arnoldi vector is exactly invariant, |q| = 0 is a
    happy breakdown after one step and the identity rows left in H give y = 0
    for the unused columns.  A with 3 distinct eigenvalues: one cycle solves."""
    b = np.arange(1.0, 9.0)
    x, info = okr.gmres(lambda v: 2.0 * v, b, np.zeros(8), tol=1e-12, restart=5)
    assert info["arnoldi"] == 1 and info["iters"] == 1
    assert np.allclose(x, b / 2, rtol=1e-15)
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.standard_normal((20, 20)))
    A = Q @ np.diag(np.repeat([1.0, 2.0, 5.0], [7, 7, 6])) @ Q.T
    b = rng.standard_normal(20)
    x, info = okr.gmres(lambda v: A @ v, b, np.zeros(20), tol=1e-12, restart=10)
    assert info["iters"] == 1
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_bicgstab_breakdown_and_truncation_codes():
    A = np.eye(4)
    b = np.ones(4)
    x, info = okr.bicgstab(lambda v: A @ v, b, np.zeros(4))
    assert np.allclose(x, 1.0) and info["iters"] == 1
    # maxiter truncation returns the iterate silently (reference accepts it)
    A, b = _spd_like(80, 3, nonsym=True)
    _, info = okr.bicgstab(lambda v: A @ v, b, np.zeros_like(b), tol=1e-30, maxiter=3)
    assert info["iters"] == 3 and info["k"] == 3


def test_cg_stopping_residual_choice():
    """Without M JAX stops on gamma = r.z = r.r; with M on r.r (not r.z).  A
    scaled Jacobi (M = c D^-1) leaves the CG iterates unchanged but scales r.z
    by c: the stop must not move."""
    A, b = _spd_like(40, 2)
    d = 1.0 / np.diag(A)
    x1, i1 = okr.cg(lambda v: A @ v, b, np.zeros_like(b), tol=1e-8, maxiter=400, dinv=d)
    x2, i2 = okr.cg(lambda v: A @ v, b, np.zeros_like(b), tol=1e-8, maxiter=400, dinv=1e-6 * d)
    assert i1["iters"] == i2["iters"] > 2
    assert np.allclose(x1, x2, rtol=1e-10)


def test_jacobi_zero_diag_maps_to_one():
    p, t = omesh.structured_tri(2, 1.0)
    pat = oasm.Pattern(t, len(p))
    vals = np.zeros(pat.nnz)
    vals[pat.diag_pos[0]] = 4.0
    dinv = oasm.jacobi_inverse(pat, vals)
    assert dinv[0] == 0.25 and np.all(dinv[1:] == 1.0)
