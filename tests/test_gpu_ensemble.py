"""Ensemble of uneven instances: every member against the oracle run of the
same instance alone.

Bar: each instance's h and theta after every step within 1e-10 relative L2
of the oracle's, with the same Picard iteration count per step, while the
other members (different sizes and soil parameters) converge on their own
schedules in the same launches.
"""
import numpy as np
import pytest

from oracle import picard as opicard

from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


def _oracle(inst, steps, dt, solver, precond, Ks=None):
    pb = opicard.Problem(points=inst.points, elems=inst.elements, soil_model="vg", phi=tuple(inst.phi),
                         bc_lists=inst.bc_lists, h0=inst.h0, dt0=dt, nsteps=steps, adaptive=False,
                         picard_tol=1e-6, picard_maxit=100, picard_mode="rel", solver=solver, precond=precond,
                         tol=1e-10, maxiter=100, Ks_elem=Ks)
    return opicard.run(pb)


def _run(ens, steps, dt):
    hs, ths, its = [], [], []
    for _ in range(steps):
        st = ens.step(dt, per_instance=True)
        h, th, _ = ens.state()
        hs.append(ens.split(h))
        ths.append(ens.split(th))
        its.append(st["picard_iters_i"].copy())
        assert st["picard_iters"] == int(st["picard_iters_i"].max())
        assert st["krylov_iters"] == int(st["krylov_iters_i"].sum())
    return hs, ths, its


# Unpreconditioned CG is left out here: on these van Genuchten instances the
# oracle itself turns a 1e-15 perturbation of h0 into 2e-11 after one step
# (roundoff amplification of the solve, not a property of the GPU path).
@pytest.mark.parametrize("solver,precond", [("bicgstab", "jacobi"), ("cg", "jacobi"), ("bicgstab", "none")])
def test_ensemble_matches_oracle_per_instance(solver, precond):
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=5, sides=(8, 12, 16, 20, 24), seed=7)
    print(f"[case] ensemble 5 {solver}+{precond} sides {[i.cells for i in insts]}", flush=True)
    ens = Ensemble(insts)
    ens.richards_init("vg", solver, precond)
    steps, dt = 3, 1e-4
    hs, ths, its = _run(ens, steps, dt)
    ens.close()
    for i, inst in enumerate(insts):
        o = _oracle(inst, steps, dt, solver, precond)
        assert [int(its[k][i]) for k in range(steps)] == list(o["iterations"]), i
        for k in range(steps):
            assert rel_l2(hs[k][i], o["pressure_head"][k]) <= 1e-10, (i, k)
            assert rel_l2(ths[k][i], o["theta"][k]) <= 1e-10, (i, k)


def test_ensemble_with_element_ks_and_single_member():
    """Per-element Ks fields, and an ensemble of one (table with one instance)."""
    from jax_wspm_amd import mesh as M
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=3, sides=(10, 14, 18), seed=11)
    for k, inst in enumerate(insts):
        inst.Ks_elem = M.correlated_log_ks(inst.points, inst.elements, inst.cells, mean_ks=1e-3, sigma=0.5,
                                           length=0.1, seed=k)
    for members in (insts, insts[1:2]):
        print(f"[case] ensemble Ks {len(members)} members", flush=True)
        ens = Ensemble(members)
        ens.richards_init("vg", "bicgstab", "jacobi")
        hs, ths, its = _run(ens, 2, 1e-4)
        ens.close()
        for i, inst in enumerate(members):
            o = _oracle(inst, 2, 1e-4, "bicgstab", "jacobi", Ks=inst.Ks_elem)
            assert [int(its[k][i]) for k in range(2)] == list(o["iterations"]), i
            for k in range(2):
                assert rel_l2(hs[k][i], o["pressure_head"][k]) <= 1e-10, (i, k)
                assert rel_l2(ths[k][i], o["theta"][k]) <= 1e-10, (i, k)


def test_ensemble_equals_single_context_runs():
    """Each member against the single-instance GPU path on the same inputs."""
    from jax_wspm_amd.device import Context
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=4, sides=(16, 32, 24, 40), seed=5)
    print("[case] ensemble vs single contexts", flush=True)
    ens = Ensemble(insts)
    ens.richards_init("vg", "bicgstab", "jacobi")
    for _ in range(4):
        ens.step(1e-4)
    h_ens = ens.split(ens.state()[0])
    ens.close()
    for i, inst in enumerate(insts):
        ctx = Context(inst.points, inst.elements)
        ctx.set_dirichlet(inst.bc_lists)
        ctx.richards_init("vg", inst.phi, "bicgstab", "jacobi", inst.h0)
        for _ in range(4):
            ctx.richards_step(1e-4)
        h = ctx.richards_state()[0].cpu().numpy()
        ctx.close()
        assert rel_l2(h_ens[i], h) <= 1e-12, i


def test_ensemble_rejects_single_instance_calls():
    from jax_wspm_amd._lib import WspmError
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    ens = Ensemble(vg_ensemble(n_inst=2, sides=(8,), seed=1))
    with pytest.raises(WspmError):
        ens._ck(ens.lib.wspm_krylov(ens.h, 1, None, None, None, None, None, 1e-10, 0.0, 10, 30, None))
    with pytest.raises(ValueError):
        ens.richards_init("vg", "gmres", "jacobi")
    ens.close()


@pytest.mark.parametrize("solver,inner", [("bicgstab", "krylov"), ("cg", "krylov"), ("bicgstab", "jacobi"),
                                          ("cg", "jacobi")])
def test_mixed_ensemble_within_1e8_of_fp64(solver, inner):
    """FP32 inner / FP64 outer ensemble: FP32 Krylov corrections of the FP64
    residual, refined per instance until that residual meets the FP64
    tolerance.  Every instance's h and theta stay within 1e-8 relative L2 of
    the FP64 ensemble (and of the oracle) over the steps, with the same
    Picard iteration counts."""
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=6, sides=(8, 12, 16, 20, 24, 32), seed=11)
    print(f"[case] mixed ensemble 6 {solver}+jacobi inner {inner}", flush=True)
    steps, dt = 4, 1e-4
    runs = {}
    for prec in ("fp64", "mixed"):
        ens = Ensemble(insts)
        ens.richards_init("vg", solver, "jacobi", precision=prec, inner=inner)
        runs[prec] = _run(ens, steps, dt)
        ens.close()
    (h64, th64, it64), (h32, th32, it32) = runs["fp64"], runs["mixed"]
    for i, inst in enumerate(insts):
        o = _oracle(inst, steps, dt, solver, "jacobi")
        assert [int(it32[k][i]) for k in range(steps)] == [int(it64[k][i]) for k in range(steps)], i
        for k in range(steps):
            assert rel_l2(h32[k][i], h64[k][i]) <= 1e-8, (i, k, rel_l2(h32[k][i], h64[k][i]))
            assert rel_l2(th32[k][i], th64[k][i]) <= 1e-8, (i, k)
            assert rel_l2(h32[k][i], o["pressure_head"][k]) <= 1e-8, (i, k)


def test_grouped_ensemble_equals_single_context():
    """EnsembleGroups (instances split over ensemble contexts on separate
    streams, stepped concurrently by host threads) against one context, on
    the default (autotuned) launch shapes: every instance's h, theta and
    per-step Picard / Krylov counts bitwise equal.  Each context derives its
    SpMV layout from the row lengths and tunes only its launch grid."""
    from jax_wspm_amd.ensemble import Ensemble, EnsembleGroups, vg_ensemble
    insts = vg_ensemble(n_inst=7, sides=(16, 40, 24, 32, 20, 36), seed=9)
    print("[case] grouped ensemble", flush=True)
    one = Ensemble(insts)
    one.richards_init("vg", "bicgstab", "jacobi")
    grp = EnsembleGroups(insts, EnsembleGroups.split_slowest([3, 9, 1, 4, 1, 8, 2], 2))
    assert grp.groups == [[0, 2, 3, 4, 6], [1, 5]]
    grp.richards_init("vg", "bicgstab", "jacobi")
    for _ in range(5):
        a = one.step(1e-4, per_instance=True)
        b = grp.step(1e-4, per_instance=True)
        for key in ("picard_iters_i", "krylov_iters_i"):
            assert np.array_equal(a[key], b[key]), key
        assert a["picard_iters"] == b["picard_iters"] and a["krylov_iters"] == b["krylov_iters"]
    for x, y in zip(one.state(), grp.state()):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
    one.close()
    grp.close()


def _ens_run(insts, steps, solver="bicgstab", precision="fp64"):
    from jax_wspm_amd.ensemble import Ensemble
    ens = Ensemble(insts)
    prec, _, inner = precision.partition("-")
    ens.richards_init("vg", solver, "jacobi", precision=prec, inner=inner or "krylov")
    its = [ens.step(1e-4, per_instance=True)["picard_iters_i"].copy() for _ in range(steps)]
    h = ens.state()[0].cpu().numpy().copy()
    shape = ens.shape
    ens.close()
    return h, np.array(its), shape


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-jacobi"])
def test_ensemble_repeatable_across_runs_and_grids(monkeypatch, precision):
    """Fresh contexts of the same instances, each autotuned on its own, and
    contexts with the launch grid forced to 1, 3 and 64 workgroups: bitwise
    the same trajectories and Picard counts (the grid decides which
    workgroup works an entry, never what is summed)."""
    from jax_wspm_amd.ensemble import vg_ensemble
    insts = vg_ensemble(n_inst=5, sides=(16, 40, 24, 32, 48), seed=4)
    print(f"[case] ensemble repeatable {precision}", flush=True)
    monkeypatch.delenv("WSPM_ENS_GRID", raising=False)
    h0, it0, sh0 = _ens_run(insts, 4, precision=precision)
    assert sh0["format"] == "ell" and sh0["grids_timed"] >= 1 and 1 <= sh0["grid"] <= sh0["entries"]
    h1, it1, _ = _ens_run(insts, 4, precision=precision)
    assert np.array_equal(it0, it1) and np.array_equal(h0, h1)
    for g in ("1", "3", "64"):
        monkeypatch.setenv("WSPM_ENS_GRID", g)
        h2, it2, sh2 = _ens_run(insts, 4, precision=precision)
        assert sh2["grid"] == min(int(g), sh2["entries"]) and sh2["grids_timed"] == 0
        assert np.array_equal(it0, it2), g
        assert np.array_equal(h0, h2), g


def test_mixed_jacobi_stall_hands_over_to_krylov(monkeypatch):
    """A Jacobi refinement sweep that does not bring the FP64 residual below
    the stall ratio times the residual before it hands the context over to
    Krylov inner solves.  Forced here with a stall ratio of 1e-30, so the
    first sweep that leaves an instance unconverged stalls: the context
    reports the hand-over and every instance still matches the FP64 ensemble
    within 1e-8 with the same Picard counts.  With the default ratio the
    same instances never stall."""
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=5, sides=(8, 16, 24, 32), seed=5)
    print("[case] mixed ensemble, Jacobi inner steps, forced stall", flush=True)
    steps, dt = 3, 1e-4
    runs, shapes = {}, {}
    for key, prec, stall in (("fp64", "fp64", None), ("jac", "mixed", None), ("stall", "mixed", "1e-30")):
        if stall:
            monkeypatch.setenv("WSPM_JAC_STALL", stall)
        else:
            monkeypatch.delenv("WSPM_JAC_STALL", raising=False)
        ens = Ensemble(insts)
        ens.richards_init("vg", "bicgstab", "jacobi", precision=prec, tol=1e-13, inner="jacobi")
        runs[key] = _run(ens, steps, dt)
        shapes[key] = ens.shape
        ens.close()
    assert not shapes["jac"]["jacobi_switched_off"]
    assert shapes["stall"]["jacobi_switched_off"]
    h64, th64, it64 = runs["fp64"]
    for key in ("jac", "stall"):
        h, th, it = runs[key]
        for k in range(steps):
            assert np.array_equal(it[k], it64[k]), (key, k)
            for i in range(len(insts)):
                assert rel_l2(h[k][i], h64[k][i]) <= 1e-8, (key, i, k)
                assert rel_l2(th[k][i], th64[k][i]) <= 1e-8, (key, i, k)


def test_mixed_jacobi_needs_jacobi_preconditioner():
    """Jacobi inner steps without the Jacobi preconditioner, and out-of-range
    step counts, are refused at initialisation."""
    from jax_wspm_amd._lib import WspmError
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=2, sides=(8, 12), seed=1)
    ens = Ensemble(insts)
    with pytest.raises(WspmError, match="Jacobi preconditioner"):
        ens.richards_init("vg", "bicgstab", None, precision="mixed", inner="jacobi")
    with pytest.raises(WspmError, match="inner_iters"):
        ens.richards_init("vg", "bicgstab", "jacobi", precision="mixed", inner="jacobi", inner_iters=17)
    ens.richards_init("vg", "bicgstab", "jacobi", precision="mixed", inner="jacobi", inner_iters=16)
    ens.step(1e-4)
    ens.close()


def test_timing_hooks_count_working_launches():
    """The benchmark's timing hooks: every timed SpMV launch that worked is
    counted once with the rows it worked on, the launches over a row
    threshold are a subset of them (read without resetting), and a read
    resets the slots."""
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=4, sides=(8, 16, 24, 32), seed=2)
    ens = Ensemble(insts)
    ens.richards_init("vg", "bicgstab", "jacobi")
    ens.step(1e-4)
    ens.timing(True, assembly=True)
    ens.step(1e-4)
    over_all = ens.timing_read_over(1)
    over_big = ens.timing_read_over(ens.n // 2)
    allr = ens.timing_read()
    asm = ens.timing_read_asm()
    assert over_all[0] == allr[0] and over_all[2] == allr[2] and abs(over_all[1] - allr[1]) < 1e-9
    assert allr[0] >= 1 and 0 < allr[2] <= allr[0] * ens.n
    assert over_big[0] <= allr[0] and over_big[2] >= over_big[0] * (ens.n // 2)
    assert asm[0] >= 1 and asm[2] <= asm[0] * ens.n
    assert ens.timing_read() == (0, 0.0, 0)
    ens.timing(False)
    ens.close()


def test_launch_count_follows_steps():
    """wspm_launch_count: a step enqueues at least an assembly, an SpMV and a
    Picard update per Picard round, and the rounds an ensemble step runs are
    the most Picard iterations any of its instances took; one system counts
    its Picard iterations."""
    from jax_wspm_amd.device import Context
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    insts = vg_ensemble(n_inst=4, sides=(8, 16, 24, 32), seed=2)
    ens = Ensemble(insts)
    ens.richards_init("vg", "bicgstab", "jacobi")
    l0, r0 = ens.launch_count()
    st = ens.step(1e-4)
    l1, r1 = ens.launch_count()
    assert r1 - r0 == st["picard_iters"] >= 1
    assert l1 - l0 >= 3 * (r1 - r0)
    ens.close()
    inst = insts[1]
    ctx = Context(inst.points, inst.elements)
    ctx.set_dirichlet(inst.bc_lists)
    ctx.richards_init("vg", inst.phi, "bicgstab", "jacobi", inst.h0)
    l0, r0 = ctx.launch_count()
    st = ctx.richards_step(1e-4)
    l1, r1 = ctx.launch_count()
    assert r1 - r0 == st["picard_iters"] >= 1
    assert l1 - l0 >= 3 * (r1 - r0)
    ctx.close()
