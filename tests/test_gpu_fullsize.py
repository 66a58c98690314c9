"""Parity at the benchmark sizes themselves, against the oracle on the same
inputs:

* 2D 512x512 (263,169 nodes, 524,288 elements), heterogeneous log-normal Ks,
  Jacobi-BiCGSTAB, 2 fixed steps of dt 1e-4;
* 3D 64^3 Kuhn mesh (274,625 nodes, 1,572,864 tetrahedra), Jacobi-CG, 1 step;
* the 256-instance ensemble (7.2 M nodes), 1 step: every instance converged,
  and a sample of 16 instances (the four largest and twelve seeded others)
  against the oracle run of each instance alone.

Bar: h and theta within 1e-10 relative L2 of the oracle's, with the same
Picard iteration counts.  The oracle runs on the CPU (about 1.4 s per 2D step
at 512x512).
"""
import numpy as np
import pytest

from oracle import picard as opicard

from tests._helpers import rel_l2, to_oracle

pytestmark = pytest.mark.gpu


def _run_gpu(pb):
    from jax_wspm_amd.richards import RichardsSolver
    s = RichardsSolver(pb)
    out = s.solve()
    s.ctx.close()
    return out


@pytest.mark.parametrize("dim,cells,solver,steps", [(2, 512, "bicgstab", 2), (3, 64, "cg", 1)])
def test_fullsize_matches_oracle(dim, cells, solver, steps):
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    ks = {"sigma": 0.5, "length": 0.05, "seed": 0} if dim == 2 else None
    print(f"[case] full size {dim}D {cells}, {solver}+jacobi, {steps} steps", flush=True)
    pb = from_config(baseline_vg_config(cells, dim, solver=solver, precond="jacobi", ks_field=ks, nsteps=steps,
                                        dt=1e-4))
    g = _run_gpu(pb)
    o = opicard.run(to_oracle(pb))
    assert list(g["iterations"]) == list(o["iterations"])
    for k in range(steps):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        assert rel_l2(g["theta"][k], o["theta"][k]) <= 1e-10, k


def test_fullsize_ensemble_sample_matches_oracle():
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(256, seed=0)
    print("[case] full ensemble 256 instances, 1 step", flush=True)
    ens = Ensemble(insts)
    ens.richards_init("vg", "bicgstab", "jacobi")
    st = ens.step(1e-4, per_instance=True)
    assert bool(np.all(st["converged_i"])), "an instance did not converge"
    h = ens.split(ens.state()[0])
    its = st["picard_iters_i"]
    ens.close()
    sizes = np.array([len(i.h0) for i in insts])
    rng = np.random.default_rng(3)
    sample = list(np.argsort(-sizes)[:4]) + list(rng.choice(len(insts), size=12, replace=False))
    for i in sorted(set(int(j) for j in sample)):
        inst = insts[i]
        o = opicard.run(opicard.Problem(points=inst.points, elems=inst.elements, soil_model="vg",
                                        phi=tuple(inst.phi), bc_lists=inst.bc_lists, h0=inst.h0, dt0=1e-4, nsteps=1,
                                        adaptive=False, picard_tol=1e-6, picard_maxit=100, picard_mode="rel",
                                        solver="bicgstab", precond="jacobi", tol=1e-10, maxiter=100))
        assert int(its[i]) == int(o["iterations"][0]), i
        assert rel_l2(h[i], o["pressure_head"][0]) <= 1e-10, i


@pytest.mark.parametrize("inner", ["krylov", "jacobi"])
def test_fullsize_mixed_3d_matches_oracle(inner):
    """The 64^3 mixed-precision configuration (FP32 inner CG or Jacobi steps,
    FP64 Picard residual, update and refinement to 1e-13), one step, against
    the FP64 oracle at the same tolerance: h and theta within 1e-8, the same
    Picard count."""
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    print(f"[case] full size 3D 64 mixed, cg+jacobi, 1 step, inner {inner}", flush=True)
    pb = from_config(baseline_vg_config(64, 3, solver="cg", precond="jacobi", nsteps=1, dt=1e-4, precision="mixed"))
    pb.tol = 1e-13
    pb.inner = inner
    g = _run_gpu(pb)
    o = opicard.run(to_oracle(pb))
    assert list(g["iterations"]) == list(o["iterations"])
    assert rel_l2(g["pressure_head"][0], o["pressure_head"][0]) <= 1e-8
    assert rel_l2(g["theta"][0], o["theta"][0]) <= 1e-8
