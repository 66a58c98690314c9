"""GPU transport path against the oracle: node fields, species systems (every
CSR value and b, with and without edge terms), and coupled time loops.

Bars: fields and systems within 1e-12 (row-scaled) of the oracle; coupled
trajectories within 1e-10 relative L2 with identical Picard iteration counts.
Parity of the oracle itself against the reference is unpinned for this path
(no reference output exists); see tests/test_oracle_transport.py.
"""
import numpy as np
import pytest

from oracle import assembly as oasm
from oracle import mesh as omesh
from oracle import transport as T

from tests._helpers import rel_l2, row_scaled_err

pytestmark = pytest.mark.gpu

PHI = np.array([3.6, 0.43, 0.078, 1.56, 1 - 1 / 1.56, 0.2496])


def _mesh(n, jitter=0.0, seed=0):
    pts, tri = omesh.structured_tri(n, 1.0)
    if jitter:
        rng = np.random.default_rng(seed)
        inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
        pts = pts.copy()
        pts[inner] += jitter / n * rng.uniform(-1, 1, size=(inner.sum(), 2))
    return pts, np.asarray(tri)


def _ctx(pts, tri):
    from jax_wspm_amd.device import Context
    return Context(pts, tri)


def _csr_vals(ctx, S):
    """Oracle scipy matrix -> values on the context's pattern (entries outside
    the pattern must be zero)."""
    rp, ci = ctx.pattern()
    S = S.tocsr()
    S.sum_duplicates()
    out = np.zeros(len(ci))
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    out[:] = np.asarray(S[rows, ci]).ravel()
    assert abs(abs(S).sum() - np.abs(out).sum()) <= 1e-12 * abs(S).sum()
    return out, rp


def test_transport_fields_match_oracle():
    pts, tri = _mesh(12, jitter=0.3, seed=1)
    print("[case] transport fields 12x12 jittered", flush=True)
    h = -1.0 - 0.8 * pts[:, 1] + 0.3 * np.sin(4 * pts[:, 0]) * pts[:, 1]
    ctx = _ctx(pts, tri)
    g = [t.cpu().numpy() for t in ctx.transport_fields(PHI, h, 0.04, 0.0004, 0.288e-4)]
    o = T.fluxes_dispersion(pts, tri, h, PHI, 0.04, 0.0004, 0.288e-4)
    ctx.close()
    for a, b in zip(g, o):
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("species", [1, 2, 3])
def test_nitrogen_species_system_matches_oracle(species):
    pts, tri = _mesh(10, jitter=0.25, seed=2)
    print(f"[case] nitrogen species {species} system 10x10", flush=True)
    n = len(pts)
    h = -1.0 - 0.5 * pts[:, 1] + 0.2 * np.cos(3 * pts[:, 0])
    rng = np.random.default_rng(species)
    c_n = rng.uniform(0, 30, n)
    c_src = rng.uniform(0, 10, n) if species > 1 else None
    top = np.nonzero(np.abs(pts[:, 1] - 1.0) < 1e-10)[0]
    bottom = np.nonzero(np.abs(pts[:, 1]) < 1e-10)[0]
    prm = T.NITROGEN_DEFAULT
    fields = T.fluxes_dispersion(pts, tri, h, PHI, prm["DL"], prm["DT"], prm["Dm"])
    K = T.van_genuchten(h, PHI)[1]
    dt = 2e-3
    c_in = (50.0, 0.0, 20.0)[species - 1]
    A_o, b_o = T.nitrogen_system(pts, tri, species, fields, prm, c_n, c_src, dt, K, top, bottom, -0.05, c_in)
    ctx = _ctx(pts, tri)
    ctx.edge_set("inlet", top)
    ctx.edge_set("outlet", bottom)
    gf = ctx.transport_fields(PHI, h, prm["DL"], prm["DT"], prm["Dm"])
    mus = (prm["mu1"], prm["mu2"], prm["mu3"])
    A, b, dinv = ctx.assemble_transport(gf, c_n, dt, lumped=True, mu=mus[species - 1],
                                        rho_kd=prm["rho"] * prm["Kd"] if species == 1 else 0.0,
                                        mu_src=mus[species - 2] if species > 1 else 0.0, c_src=c_src,
                                        cauchy=(-0.05, c_in), K=K)
    vals_o, rp = _csr_vals(ctx, A_o)
    ctx.close()
    assert row_scaled_err(A.cpu().numpy(), vals_o, rp) <= 1e-12
    assert np.abs(b.cpu().numpy() - b_o).max() <= 1e-12 * np.abs(b_o).max()
    d = A_o.diagonal()
    np.testing.assert_allclose(dinv.cpu().numpy(), 1 / np.where(d != 0, d, 1), rtol=1e-12)


def test_solute_system_consistent_mass_matches_oracle():
    pts, tri = _mesh(9, jitter=0.3, seed=3)
    print("[case] solute system 9x9", flush=True)
    n = len(pts)
    h = -1.3 - 0.4 * pts[:, 1] + 0.1 * pts[:, 0]
    c_n = np.random.default_rng(4).uniform(0, 1, n)
    top = np.nonzero(np.abs(pts[:, 1] - 1.0) < 1e-10)[0][2:5]
    fields = T.fluxes_dispersion(pts, tri, h, PHI, 0.5, 0.1, 0.0)
    A_o, b_o = T.solute_system(pts, tri, fields, c_n, 1e-2, top, -0.1, 1.0)
    ctx = _ctx(pts, tri)
    ctx.edge_set("inlet", top)
    gf = ctx.transport_fields(PHI, h, 0.5, 0.1, 0.0)
    A, b, _ = ctx.assemble_transport(gf, c_n, 1e-2, lumped=False, cauchy=(-0.1, 1.0))
    vals_o, rp = _csr_vals(ctx, A_o)
    ctx.close()
    assert row_scaled_err(A.cpu().numpy(), vals_o, rp) <= 1e-12
    assert np.abs(b.cpu().numpy() - b_o).max() <= 1e-12 * np.abs(b_o).max()


def test_edge_list_matrix_terms_need_mesh_edges():
    """A boundary list whose consecutive nodes are not joined by a mesh edge
    (the reference's bottom lists of the unstructured solute meshes, in index
    order) is accepted for right-hand-side terms; the matrix terms (Cauchy,
    free drainage of nitrogen) have no pattern entries for it and refuse it.
    A repeated node is refused at once."""
    from jax_wspm_amd._lib import WspmError
    pts, tri = _mesh(4)
    n = len(pts)
    ctx = _ctx(pts, tri)
    with pytest.raises(WspmError):
        ctx.edge_set("inlet", [0, 1, 0])       # repeated node
    ctx.edge_set("inlet", [0, 2])              # not joined by a mesh edge
    h = -1.3 - 0.4 * pts[:, 1]
    gf = ctx.transport_fields(PHI, h, 0.5, 0.1, 0.0)
    with pytest.raises(WspmError, match="not joined"):
        ctx.assemble_transport(gf, np.zeros(n), 1e-2, lumped=False, cauchy=(-0.1, 1.0))
    ctx.assemble_transport(gf, np.zeros(n), 1e-2, lumped=False)   # no edge term: fine
    ctx.close()


def _nitrogen_cfg(solver, precond, dt=1e-3, nsteps=3):
    from jax_wspm_amd.config import SimulationConfig, SolverParameters, TimeSteppingParameters
    cfg = SimulationConfig(test_case="NitrogenTest", solver=SolverParameters(solver_type=solver, precond_type=precond),
                           time=TimeSteppingParameters(dt_init=dt, Tmax=1.0))
    return cfg


@pytest.mark.parametrize("solver,precond", [("gmres", "none"), ("bicgstab", "jacobi")])
def test_nitrogen_solver_matches_oracle(solver, precond):
    from jax_wspm_amd.transport import NitrogenSolver
    pts, tri = omesh.structured_tri(10, 1.0)
    tri = np.asarray(tri)
    print(f"[case] NitrogenSolver 10x10 {solver}+{precond} 4 steps", flush=True)
    cfg = _nitrogen_cfg(solver, precond)
    s = NitrogenSolver(cfg, points=pts, triangles=tri)
    g = s.solve(max_steps=4)
    s.ctx.close()
    o = T.nitrogen_run(pts, tri, cfg.van_genuchten.to_array(), 4, 1e-3, solver=solver, precond=precond)
    assert list(g["iterations"]) == list(o["iterations"])
    np.testing.assert_allclose(g["dt_values"], o["dt_values"], rtol=0, atol=0)
    for k in range(4):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        for key, okey in (("NH4_concentration", "NH4"), ("NO2_concentration", "NO2"), ("NO3_concentration", "NO3")):
            assert rel_l2(g[key][k], o[okey][k]) <= 1e-10, (key, k)


def test_coupled_solute_solver_matches_oracle():
    """SoluteTest on a generated [-1, 1] x [0, 2] mesh (inlet |x| <= 0.05 at
    the top, as src/models/boundary.py:499-504)."""
    from jax_wspm_amd.config import SimulationConfig, SolverParameters, TimeSteppingParameters
    from jax_wspm_amd.transport import CoupledSolver, transport_boundary_nodes
    N = 40
    ax = np.round(np.arange(N + 1) / N, 12)
    pts = np.stack([np.round(-1 + 2 * np.tile(ax, N + 1), 12), np.round(2 * np.repeat(ax, N + 1), 12)], axis=1)
    _, tri = omesh.structured_tri(N, 1.0)
    tri = np.asarray(tri)
    top, _ = transport_boundary_nodes(pts, "SoluteTest")
    assert len(top) == 3
    print("[case] CoupledSolver 40x40 bicgstab+jacobi 3 steps", flush=True)
    cfg = SimulationConfig(test_case="SoluteTest",
                           solver=SolverParameters(solver_type="bicgstab", precond_type="jacobi"),
                           time=TimeSteppingParameters(dt_init=1e-3, Tmax=1.0))
    s = CoupledSolver(cfg, points=pts, triangles=tri)
    g = s.solve(max_steps=3)
    s.ctx.close()
    o = T.solute_run(pts, tri, cfg.van_genuchten.to_array(), 3, 1e-3, top, solver="bicgstab", precond="jacobi")
    assert list(g["iterations"]) == list(o["iterations"])
    np.testing.assert_allclose(g["dt_values"], o["dt_values"], rtol=0, atol=0)
    for k in range(3):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        assert rel_l2(g["solute"][k], o["solute"][k]) <= 1e-10, k
