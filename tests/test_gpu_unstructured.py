"""GPU parity on the reference's unstructured solute meshes.

data/mesh/solute/{p,t}_Pinns_{25,1024}.csv (committed as
tests/golden/pinns_meshes.npz) are the meshes src/mesh/loader.py:33-36 loads
for SoluteTest.  Their node numbering is arbitrary and their rows hold up to 9
(Pinns_25) and 10 (Pinns_1024) entries, so the time step runs the LDS-staged
assembly and the CSR operator instead of the register / ELL kernels of the
structured meshes, and the transport assembly its 16-entry row variant.

Bars: trajectories equal the oracle's to <= 1e-10 relative L2 with identical
Picard iteration counts.  ``reorder="rcm"`` runs the same problems on the
reverse Cuthill-McKee renumbering of the mesh (results mapped back to the
reference's numbering, boundary lists kept in the reference's order).
"""
import numpy as np
import pytest

from oracle import picard as opicard
from oracle import transport as T

from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


def _mesh(golden, tag):
    z = np.load(golden / "pinns_meshes.npz")
    return z[f"points_{tag}"], z[f"elements_{tag}"].astype(np.int64)


@pytest.mark.parametrize("solver,precond,reorder", [("gmres", "none", None), ("bicgstab", "jacobi", None),
                                                    ("bicgstab", "jacobi", "rcm")])
def test_solute_pinns25_matches_oracle(golden, solver, precond, reorder):
    """SoluteTest (coupled_solver_v2.py:221-331) on Pinns_25, the CLI's
    default mesh: Richards with the Neumann inlet, then the solute with its
    Cauchy inlet, 3 adaptive steps."""
    from jax_wspm_amd.config import SimulationConfig, SolverParameters, TimeSteppingParameters
    from jax_wspm_amd.transport import CoupledSolver, transport_boundary_nodes
    pts, tri = _mesh(golden, "25")
    top, _ = transport_boundary_nodes(pts, "SoluteTest")
    assert len(top) == 11
    print(f"[case] SoluteTest Pinns_25 {solver}+{precond} reorder={reorder}", flush=True)
    cfg = SimulationConfig(test_case="SoluteTest", solver=SolverParameters(solver_type=solver, precond_type=precond),
                           time=TimeSteppingParameters(dt_init=1e-3, Tmax=1.0))
    s = CoupledSolver(cfg, points=pts, triangles=tri, reorder=reorder)
    assert s.ctx.max_row == 9          # rows beyond the register kernels' 8
    g = s.solve(max_steps=3)
    s.ctx.close()
    o = T.solute_run(pts, tri, cfg.van_genuchten.to_array(), 3, 1e-3, top, solver=solver, precond=precond)
    assert list(g["iterations"]) == list(o["iterations"])
    np.testing.assert_allclose(g["dt_values"], o["dt_values"], rtol=0, atol=0)
    for k in range(3):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        assert rel_l2(g["theta"][k], o["theta"][k]) <= 1e-10, k
        assert rel_l2(g["solute"][k], o["solute"][k]) <= 1e-10, k


def _vg_problem(pts, tri, solver, precond, nsteps=3, dt=1e-4):
    from jax_wspm_amd.problems import Problem
    y = pts[:, 1]
    top = np.nonzero(y == y.max())[0]
    return Problem(name="VGPinns", points=pts, elements=tri, soil_model="vg",
                   phi=(0.1, 0.40, 0.06, 2.0, 0.5, 1e-3), bc_lists=[(top, np.full(top.shape, -75.0))],
                   h0=np.full(len(pts), -1000.0), dt0=dt, nsteps=nsteps, solver=solver, precond=precond)


@pytest.mark.parametrize("tag", ["25", "1024"])
@pytest.mark.parametrize("solver,precond,reorder", [("bicgstab", "jacobi", None), ("cg", "jacobi", None),
                                                    ("gmres", "jacobi", None), ("bicgstab", "jacobi", "rcm")])
def test_richards_pinns_matches_oracle(golden, tag, solver, precond, reorder):
    """The van Genuchten time step (richards_solver_v2.py:107-168) on the
    unstructured meshes: top Dirichlet -75, h0 -1000, 3 fixed steps.  The dry
    interior rows are about 1e-8 the scale of the Dirichlet rows, so without a
    preconditioner GMRES stops (rtol 1e-10 on the unscaled residual) where the
    two summation orders leave the solutions 1.8e-10 apart on Pinns_25; the
    Jacobi-preconditioned solves are compared at the 1e-10 bar."""
    from jax_wspm_amd.richards import RichardsSolver
    from tests._helpers import to_oracle
    pts, tri = _mesh(golden, tag)
    pb = _vg_problem(pts, tri, solver, precond)
    print(f"[case] Richards Pinns_{tag} {solver}+{precond} reorder={reorder}", flush=True)
    s = RichardsSolver(pb, reorder=reorder)
    assert s.ctx.max_row >= 9
    g = s.solve()
    s.ctx.close()
    o = opicard.run(to_oracle(pb))
    assert list(g["iterations"]) == list(o["iterations"])
    for k in range(len(o["times"])):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        assert rel_l2(g["theta"][k], o["theta"][k]) <= 1e-10, k
