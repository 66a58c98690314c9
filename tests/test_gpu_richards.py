"""End-to-end GPU parity of the implicit time step (Picard + Krylov + assembly).

Bars: FP64 trajectories equal the oracle's to <= 1e-10 relative L2 in h and
theta at every recorded step with identical Picard iteration counts; the
reference's published fixtures (3D NPZ trajectories, Test1-3 CSV metrics)
reproduce at their own precision; the mixed-precision form stays within 1e-8
relative L2 of the FP64 form.
"""
import json

import numpy as np
import pytest

from oracle import picard as opicard
from oracle import tracy

from tests._helpers import rel_l2, to_oracle

pytestmark = pytest.mark.gpu


def _cfg(test_case, **kw):
    from jax_wspm_amd.config import PicardParameters, SimulationConfig, SolverParameters, TimeSteppingParameters
    cfg = SimulationConfig(
        time=TimeSteppingParameters(dt_init=kw.get("dt", 1e-3), Tmax=kw.get("tmax", 0.25)),
        solver=SolverParameters(solver_type=kw.get("solver", "cg"), precond_type=kw.get("precond", "none")),
        test_case=test_case, picard=PicardParameters(rule=kw.get("rule", "published")))
    cfg.mesh.mesh_size = kw.get("mesh", "25")
    return cfg


def _run(pb):
    from jax_wspm_amd.richards import RichardsSolver
    s = RichardsSolver(pb)
    out = s.solve()
    s.ctx.close()
    return out


def _compare(g, o, tol=1e-10):
    assert list(g["iterations"]) == list(o["iterations"])
    assert np.allclose(g["times"], o["times"], rtol=0, atol=0)
    for k in range(len(o["times"])):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= tol, k
        assert rel_l2(g["theta"][k], o["theta"][k]) <= tol, k


@pytest.mark.parametrize("tc", ["Test1", "Test2", "Test3"])
def test_tracy_published_metrics(tc, golden):
    """Test1-3 on the 25x25 mesh reproduce resultsRE2D/*.csv (5 significant
    figures, summed Picard iterations) and the oracle trajectory."""
    from jax_wspm_amd.problems import from_config
    print(f"[case] {tc} mesh 25", flush=True)
    pb = from_config(_cfg(tc, solver="cg", precond="none"))
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    _compare(g, o)
    rows = json.loads((golden / "tracy_metrics.json").read_text())
    row = next(r for r in rows if r["test_case"] == tc and r["mesh_size"] == "25" and r["solver"] == "cg"
               and r["preconditioner"] == "none")
    assert int(row["iterations"]) == int(np.sum(g["iterations"]))
    e = tracy.final_errors(pb.points, g["pressure_head"][-1], g["theta"][-1], g["times"][-1], int(tc[-1]))
    for key in ("l2_pressure", "linf_pressure", "l2_saturation", "linf_saturation", "l2_relative_pressure",
                "l2_relative_saturation"):
        assert f"{e[key]:.4e}" == row[key], (key, e[key], row[key])


@pytest.mark.parametrize("solver", ["bicgstab", "gmres"])
def test_test3d_matches_reference_npz(solver, golden):
    """Test3D (20^3 Kuhn mesh, Gardner) against results3D/*.npz of the reference."""
    from jax_wspm_amd.problems import from_config
    print(f"[case] Test3D {solver}", flush=True)
    pb = from_config(_cfg("Test3D", solver=solver, dt=1e-4, tmax=0.1))
    g = _run(pb)
    z = np.load(golden / f"test3d_{solver}.npz")
    assert list(g["iterations"]) == list(z["iterations"])
    assert np.array_equal(g["times"], z["times"]) and np.array_equal(g["dt_values"], z["dt_values"])
    for s, h in zip(z["pressure_head_steps"], z["pressure_head"]):
        assert rel_l2(g["pressure_head"][s], h) <= 1e-10
    assert rel_l2(g["theta"][-1], z["theta_final"]) <= 1e-10
    assert np.allclose(g["errors"], z["errors"], rtol=1e-5)


@pytest.mark.parametrize("solver,precond", [("gmres", "jacobi"), ("bicgstab", "jacobi"), ("cg", "none")])
def test_vg_square_adaptive_matches_oracle(solver, precond):
    """van Genuchten 25x25 unit square, top Dirichlet -75, adaptive dt from 1e-4 to 0.25."""
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print(f"[case] VGSquare 25 {solver} {precond}", flush=True)
    pb = from_config(baseline_vg_config(25, 2, solver=solver, precond=precond, dt=1e-4, tmax=0.25))
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    _compare(g, o)


def test_vg_square_heterogeneous_ks_fixed_steps():
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print("[case] VGSquare 48 Ks field", flush=True)
    pb = from_config(baseline_vg_config(48, 2, solver="bicgstab", precond="jacobi", nsteps=6, dt=1e-4,
                                        ks_field={"sigma": 0.5, "length": 0.05, "seed": 7}))
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    _compare(g, o)


def test_vg_cube_fixed_steps():
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print("[case] VGCube 8", flush=True)
    pb = from_config(baseline_vg_config(8, 3, solver="cg", precond="jacobi", nsteps=4, dt=1e-4))
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    _compare(g, o)


@pytest.mark.parametrize("dim,solver,inner", [(2, "bicgstab", "krylov"), (3, "cg", "krylov"), (2, "gmres", "krylov"),
                                              (2, "bicgstab", "jacobi"), (3, "cg", "jacobi")])
def test_mixed_precision_vs_fp64(dim, solver, inner):
    """FP32 inner solves (the method's Krylov solve, or Jacobi steps) + FP64
    refinement/Picard stays within 1e-8 of FP64."""
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print(f"[case] mixed dim={dim} {solver} inner {inner}", flush=True)
    cells = 40 if dim == 2 else 10
    pb64 = from_config(baseline_vg_config(cells, dim, solver=solver, precond="jacobi", nsteps=5, dt=1e-4))
    pbmx = from_config(baseline_vg_config(cells, dim, solver=solver, precond="jacobi", nsteps=5, dt=1e-4,
                                          precision="mixed"))
    pbmx.inner = inner
    g64 = _run(pb64)
    gmx = _run(pbmx)
    for k in range(5):
        err_h = rel_l2(gmx["pressure_head"][k], g64["pressure_head"][k])
        err_t = rel_l2(gmx["theta"][k], g64["theta"][k])
        print(f"  step {k}: rel L2 h {err_h:.2e} theta {err_t:.2e}")
        assert err_h <= 1e-8 and err_t <= 1e-8


@pytest.mark.parametrize("solver,precond", [("cg", "ilu"), ("bicgstab", "ssor"), ("gmres", "block_jacobi")])
def test_tracy_general_preconditioners(solver, precond, golden):
    """Test1 mesh 25 (published rule) with the preconditioners applied between
    the Krylov kernels: the oracle trajectory, and the converged published
    metrics (53 Picard iterations; resultsRE2D/benchmark_results_cpu_Test1_WA.csv
    holds the same figures for its *,ilu rows)."""
    from jax_wspm_amd.problems import from_config
    print(f"[case] Test1 mesh 25 {solver}+{precond}", flush=True)
    pb = from_config(_cfg("Test1", solver=solver, precond=precond))
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    _compare(g, o)
    assert int(np.sum(g["iterations"])) == 53
    e = tracy.final_errors(pb.points, g["pressure_head"][-1], g["theta"][-1], g["times"][-1], 1)
    assert f"{e['l2_pressure']:.4e}" == "1.7613e-01" and f"{e['l2_saturation']:.4e}" == "5.0311e-03"


@pytest.mark.parametrize("cells,steps,tol,inner", [(16, 5, 1e-10, "krylov"), (16, 5, 1e-13, "krylov"),
                                                 (24, 3, 1e-13, "krylov"), (16, 5, 1e-13, "jacobi"),
                                                 (24, 3, 1e-14, "jacobi")])
def test_mixed_3d_matches_oracle(cells, steps, tol, inner):
    """The 3D configuration with FP32 inner CG (Jacobi) inside FP64 iterative
    refinement, FP64 Picard residual and update (unit cube, Kuhn tets, h0
    -500, top -10, van Genuchten alpha 0.1, n 2), against the FP64 oracle at
    the same Krylov tolerance: the same Picard counts, h and theta within
    1e-8 relative L2 at every step."""
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print(f"[case] mixed 3D {cells}^3 CG+jacobi {steps} steps tol {tol:g} inner {inner}", flush=True)
    pb = from_config(baseline_vg_config(cells, 3, solver="cg", precond="jacobi", nsteps=steps, dt=1e-4,
                                        precision="mixed"))
    pb.tol = tol
    pb.inner = inner
    g = _run(pb)
    o = opicard.run(to_oracle(pb))
    assert list(g["iterations"]) == list(o["iterations"])
    for k in range(steps):
        eh = rel_l2(g["pressure_head"][k], o["pressure_head"][k])
        et = rel_l2(g["theta"][k], o["theta"][k])
        print(f"  step {k}: rel L2 h {eh:.2e} theta {et:.2e}")
        assert eh <= 1e-8 and et <= 1e-8, k


def _dry_instances():
    """Corners of the ensemble's parameter box: alpha 0.01 / 0.2, n 1.3 / 3.0,
    Ks 1e-4 / 1e-2, sides 32 and 48 (alpha 0.2, n 3 at h = -1000 gives
    C ~ 2e-8: interior rows some 1e-8 the scale of the Dirichlet rows)."""
    from jax_wspm_amd import mesh as M
    from jax_wspm_amd.ensemble import Instance
    out = []
    for k, (alpha, n, ks, N) in enumerate([(0.01, 1.3, 1e-4, 32), (0.2, 3.0, 1e-4, 48), (0.2, 1.3, 1e-2, 32),
                                           (0.01, 3.0, 1e-2, 48), (0.2, 3.0, 1e-2, 32), (0.01, 1.3, 1e-2, 48)]):
        pts, el = M.unit_square(N, 1.0)
        top = np.nonzero(pts[:, 1] == 1.0)[0]
        ts, tr = (0.35, 0.1) if k % 2 else (0.5, 0.02)
        out.append(Instance(points=pts, elements=el, phi=(alpha, ts, tr, n, 1.0 - 1.0 / n, ks),
                            bc_lists=[(top, np.full(top.shape, -75.0))], h0=np.full(len(pts), -1000.0), cells=N))
    return out


@pytest.mark.parametrize("inner", ["krylov", "jacobi"])
def test_mixed_ensemble_dry_instances_within_1e8(inner):
    """Batched mixed precision on the driest, worst-conditioned corners of the
    ensemble's parameter box: refined until the FP64 residual is 1e-14 of
    |b| (where the FP64 solution itself is converged), every instance stays
    within 1e-8 relative L2 of the FP64 ensemble at the same tolerance and of
    the oracle, with the same Picard counts, over 10 steps."""
    from jax_wspm_amd.ensemble import Ensemble
    insts = _dry_instances()
    steps, tol = 10, 1e-14
    print(f"[case] mixed ensemble, dry corners, tol 1e-14, inner {inner}", flush=True)
    res = {}
    for prec in ("mixed", "fp64"):
        ens = Ensemble(insts)
        ens.richards_init("vg", "bicgstab", "jacobi", precision=prec, tol=tol, inner_tol=1e-6, inner=inner)
        its = [ens.step(1e-4, per_instance=True)["picard_iters_i"].copy() for _ in range(steps)]
        h, th, _ = ens.state()
        res[prec] = (np.array(its), ens.split(h), ens.split(th))
        ens.close()
    (im, hm, tm), (i6, h6, t6) = res["mixed"], res["fp64"]
    assert np.array_equal(im, i6)
    for i, inst in enumerate(insts):
        eh, et = rel_l2(hm[i], h6[i]), rel_l2(tm[i], t6[i])
        print(f"  instance {i} alpha {inst.phi[0]} n {inst.phi[3]}: rel L2 h {eh:.2e} theta {et:.2e}")
        assert eh <= 1e-8 and et <= 1e-8, i
        o = opicard.run(opicard.Problem(points=inst.points, elems=inst.elements, soil_model="vg",
                                        phi=tuple(inst.phi), bc_lists=inst.bc_lists, h0=inst.h0, dt0=1e-4,
                                        nsteps=steps, adaptive=False, solver="bicgstab", precond="jacobi", tol=tol,
                                        maxiter=100))
        assert list(o["iterations"]) == [int(x) for x in im[:, i]], i
        assert rel_l2(hm[i], o["pressure_head"][-1]) <= 1e-8, i


def _trajectory(pb, steps):
    from jax_wspm_amd.richards import RichardsSolver
    s = RichardsSolver(pb, history="last")
    s.init_state()
    its = [s.ctx.richards_step(pb.dt0)["picard_iters"] for _ in range(steps)]
    h, th, _ = s.ctx.richards_state()
    windows = s.ctx.assembly_windows
    s.ctx.close()
    return its, h.cpu().numpy(), th.cpu().numpy(), windows


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_windowed_fan_assembly_bitwise(monkeypatch, precision):
    """The fan assembly that stages each row block's node window in LDS,
    against the fan assembly that gathers the records from global memory
    (WSPM_ASM_WIN=0): bitwise the same trajectories and Picard counts, on a
    300 x 300 grid with per-element Ks (a row block then touches three
    separate stretches of nodes: its own rows and the grid rows either side)
    and on an ensemble of uneven instances."""
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print(f"[case] windowed fan assembly bitwise, {precision}", flush=True)
    steps = 3
    res = {}
    for win in ("1", "0"):
        monkeypatch.setenv("WSPM_ASM_WIN", win)
        pb = from_config(baseline_vg_config(300, 2, solver="bicgstab", precond="jacobi", nsteps=steps, dt=1e-4,
                                            precision=precision,
                                            ks_field={"sigma": 0.5, "length": 0.05, "seed": 3}))
        if precision == "mixed":
            pb.tol = 1e-13
            pb.inner = "jacobi"
        res[win] = _trajectory(pb, steps)
    assert res["1"][3] and not res["0"][3]
    assert res["1"][0] == res["0"][0]
    assert np.array_equal(res["1"][1], res["0"][1]) and np.array_equal(res["1"][2], res["0"][2])
    insts = vg_ensemble(n_inst=5, sides=(16, 40, 64, 100, 300), seed=6)
    ens_res = {}
    for win in ("1", "0"):
        monkeypatch.setenv("WSPM_ASM_WIN", win)
        ens = Ensemble(insts)
        ens.richards_init("vg", "bicgstab", "jacobi", precision=precision, tol=1e-13, inner="jacobi")
        its = [ens.step(1e-4, per_instance=True)["picard_iters_i"].copy() for _ in range(steps)]
        ens_res[win] = (np.array(its), ens.state()[0].cpu().numpy(), ens.assembly_windows)
        ens.close()
    assert ens_res["1"][2] and not ens_res["0"][2]
    assert np.array_equal(ens_res["1"][0], ens_res["0"][0])
    assert np.array_equal(ens_res["1"][1], ens_res["0"][1])
