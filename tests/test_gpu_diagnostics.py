"""GPU diagnostics against the oracle: domain integrals (mass balance), element
sizes and Peclet / Courant numbers, through the C ABI.

Bars: element sizes and Peclet / Courant numbers bitwise (the same IEEE
operations in the same order, maxima exact); domain integrals within 1e-13
relative (different summation order).
"""
import numpy as np
import pytest
import torch

from oracle import diagnostics as D
from oracle import mesh as omesh
from oracle import tracy as otracy

pytestmark = pytest.mark.gpu


def _perturbed(cells, seed=0):
    pts, tri = omesh.structured_tri(cells, 1.0)
    pts = pts.copy()
    rng = np.random.default_rng(seed)
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts[inner] += rng.uniform(-0.15, 0.15, size=(inner.sum(), 2)) / cells   # keeps every triangle upright
    return pts, np.asarray(tri)


@pytest.mark.parametrize("cells", [1, 5, 64, 300])
def test_domain_integral(cells):
    from jax_wspm_amd.device import Context
    print(f"[case] domain integral, perturbed {cells}x{cells}", flush=True)
    pts, tri = _perturbed(cells, seed=cells)
    ctx = Context(pts, tri)
    f = np.random.default_rng(cells).normal(size=len(pts))
    got = ctx.domain_integral(torch.tensor(f, device="cuda"))
    assert got == pytest.approx(D.integrate_2d_domain(f, pts, tri), rel=1e-13, abs=1e-13)
    assert ctx.domain_integral(torch.ones(len(pts), dtype=torch.float64, device="cuda")) == pytest.approx(2.0,
                                                                                                        rel=1e-13)
    ctx.close()


def test_element_size_bitwise():
    from jax_wspm_amd.device import Context
    print("[case] element size 40x40 perturbed", flush=True)
    pts, tri = _perturbed(40, seed=7)
    ctx = Context(pts, tri)
    got = ctx.element_size().cpu().numpy()
    assert np.array_equal(got, D.element_size(pts, tri))
    ctx.close()


def _fields(cells=24):
    from jax_wspm_amd.device import Context
    pts, tri = _perturbed(cells, seed=2)
    ctx = Context(pts, tri)
    x, y = pts[:, 0], pts[:, 1]
    h = torch.tensor(-1.0 - 3.0 * y + 0.5 * np.sin(7 * x), device="cuda")
    phi = np.array([0.036, 0.43, 0.078, 1.56, 1 - 1 / 1.56, 0.0104])
    f = ctx.transport_fields(phi, h, 0.5, 0.1, 1e-5)
    return ctx, pts, tri, f


@pytest.mark.parametrize("all_elements", [False, True])
def test_peclet_courant_bitwise(all_elements):
    print(f"[case] peclet/courant all_elements={all_elements}", flush=True)
    ctx, pts, tri, f = _fields()
    qx, qz, _, dxx, _, dzz, th = (v.cpu().numpy() for v in f)
    pe, cr, pe_e, cr_e = ctx.peclet_courant(f[0], f[1], f[6], f[3], f[5], 0.01, all_elements=all_elements,
                                            per_element=True)
    ope, ocr = D.element_peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.01)
    assert np.array_equal(pe_e.cpu().numpy(), ope) and np.array_equal(cr_e.cpu().numpy(), ocr)
    assert (pe, cr) == D.peclet_courant(pts, tri, qx, qz, th, dxx, dzz, 0.01, all_elements=all_elements)
    # without per-element outputs the reference form reads elements 0 and 1 only
    assert ctx.peclet_courant(f[0], f[1], f[6], f[3], f[5], 0.01, all_elements=all_elements) == (pe, cr)
    ctx.close()


def test_peclet_nan_propagates():
    from jax_wspm_amd.device import Context
    print("[case] peclet NaN", flush=True)
    pts, tri = _perturbed(6)
    ctx = Context(pts, tri)
    n = len(pts)
    z = torch.zeros(n, dtype=torch.float64, device="cuda")
    one = torch.ones(n, dtype=torch.float64, device="cuda")
    th = one.clone()
    th[tri[17, 1]] = 0.0                        # q / theta = 0 / 0 at one node of element 17
    pe, cr = ctx.peclet_courant(z, z, th, one, one, 0.1, all_elements=True)
    assert np.isnan(pe) and np.isnan(cr)
    pe, cr = ctx.peclet_courant(z, z, one, one, one, 0.1, all_elements=True)
    assert pe == 0.0 and cr == 0.0
    ctx.close()


def test_mass_balance_records():
    """RichardsSolver(mass_balance=True) on Test1 25x25: the recorded domain
    integrals equal the oracle's integrals of the returned theta and of the
    exact Tracy theta at the recorded times."""
    from jax_wspm_amd.config import SimulationConfig
    from jax_wspm_amd.richards import RichardsSolver
    print("[case] mass balance Test1 25x25, up to 12 steps", flush=True)
    cfg = SimulationConfig()
    cfg.test_case = "Test1"
    cfg.mesh.mesh_size = "25"
    cfg.time.dt_init = 1e-3
    cfg.time.Tmax = 0.25
    sv = RichardsSolver(cfg, mass_balance=True)
    res = sv.solve(max_steps=12)
    sv.ctx.close()
    pts, tri = res["points"], res["triangles"]
    ex = cfg.exponential
    eps0 = np.exp(ex.alpha0 * -15.24)
    d = ex.alpha0 * (ex.thetas - ex.thetar) / ex.Ks
    steps = len(res["times"])                  # adaptive dt: Tmax reached after 9 steps
    assert steps >= 5 and len(res["mass_balance_numerical"]) == steps
    for k in range(steps):
        num = D.integrate_2d_domain(res["theta"][k], pts, tri)
        _, S = otracy.exact_solution(pts[:, 0], pts[:, 1], res["times"][k], 1, 15.24, ex.alpha0, eps0, d)
        exm = D.integrate_2d_domain((ex.thetas - ex.thetar) * S + ex.thetar, pts, tri)
        assert res["mass_balance_numerical"][k] == pytest.approx(num, rel=1e-13)
        assert res["mass_balance_exact"][k] == pytest.approx(exm, rel=1e-12)
    err = D.mass_balance_error(res["mass_balance_numerical"], res["mass_balance_exact"])
    assert np.allclose(res["mass_balance_error"], err, rtol=1e-12, atol=1e-12)
    assert np.all(res["mass_balance_error"] < 5.0)


def test_transport_diagnostics_records():
    from jax_wspm_amd.config import SimulationConfig
    from jax_wspm_amd.transport import NitrogenSolver
    print("[case] nitrogen diagnostics 25x25, 3 steps", flush=True)
    cfg = SimulationConfig()
    cfg.test_case = "NitrogenTest"
    cfg.mesh.mesh_size = "25"
    cfg.time.dt_init = 1e-3
    sv = NitrogenSolver(cfg, diagnostics=True)
    res = sv.solve(max_steps=3)
    pe, cr = res["peclet_max"], res["courant_max"]
    assert len(pe) == 3 and np.all(np.isfinite(pe)) and np.all(pe > 0) and np.all(cr > 0)
    # the last step's value equals the oracle on the last step's fields
    h = torch.tensor(res["pressure_head"][-1], device="cuda")
    nc = cfg.nitrogen
    f = sv.ctx.transport_fields(sv.phi, h, nc.DL, nc.DT, nc.Dm)
    qx, qz, _, dxx, _, dzz, th = (v.cpu().numpy() for v in f)
    want = D.peclet_courant(sv.points, sv.triangles, qx, qz, th, dxx, dzz, res["dt_values"][-1], all_elements=True)
    assert (pe[-1], cr[-1]) == want
    sv.ctx.close()
