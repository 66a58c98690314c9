"""Launch modes of the time step: HIP-graph replay of whole Picard
iterations, the spin-wait read-back and the finish-free Krylov iterations.

The time step replays each Picard iteration (assembly + solver start, the
expected Krylov iterations, the Picard error, the state publication) from a
captured HIP graph and reads the device state back through a sequence word in
pinned host memory.  Without graphs, the second Picard iteration of a step
is enqueued before the first one's read-back, behind a device-side gate
(WSPM_AHEAD=0: enqueued after it).  WSPM_GRAPHS=0 launches every kernel on the stream and
WSPM_POLL=sync reads back by copy + stream synchronisation.  Ensemble
Krylov lists are rebuilt by the finish kernel of each iteration's last
reduction (entries in arrival order); WSPM_ENS_LIST=kernel and graph replay
rebuild them with a separate kernel, in instance order.  Every kernel is
deterministic, so all modes must give bitwise identical trajectories and
iteration counts; each graph is replayed many times, with the host's own
launches (Krylov follow-ups, the theta_n copy) and synchronisations between
the replays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [("1", "spin"), ("0", "spin"), ("0", "sync")]  # (WSPM_GRAPHS, WSPM_POLL)


def _single(monkeypatch, graphs, poll, solver, precond, dts, krylov="folded", precision="fp64"):
    from jax_wspm_amd.device import Context
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    monkeypatch.setenv("WSPM_GRAPHS", graphs)
    monkeypatch.setenv("WSPM_POLL", poll)
    monkeypatch.setenv("WSPM_KRYLOV", krylov)
    pb = from_config(baseline_vg_config(40, 2, solver=solver, precond=precond, nsteps=1, dt=1e-4,
                                        ks_field={"sigma": 0.5, "length": 0.05, "seed": 3}))
    ctx = Context(pb.points, pb.elements)
    ctx.set_dirichlet(pb.bc_lists)
    ctx.richards_init("vg", pb.phi, solver, precond, pb.h0, Ks_elem=pb.Ks_elem, precision=precision)
    hs, its = [], []
    for dt in dts:
        st = ctx.richards_step(dt)
        its.append((st["picard_iters"], st["krylov_iters"]))
        hs.append(ctx.richards_state()[0].cpu().numpy().copy())
    ctx.close()
    return hs, its


@pytest.mark.parametrize("solver,precond", [("bicgstab", "jacobi"), ("cg", "jacobi"), ("bicgstab", "none")])
def test_single_context_modes_bitwise(monkeypatch, solver, precond):
    # repeated and alternating dt values: graphs are looked up by dt, by the
    # buffer holding the Picard iterate and by the Krylov chunk
    dts = [1e-4] * 5 + [2e-4, 1e-4, 2e-4, 2e-4, 1e-4]
    print(f"[case] single context {solver}+{precond}", flush=True)
    ref_h, ref_it = _single(monkeypatch, "0", "sync", solver, precond, dts)
    for graphs, poll in MODES[:2]:
        h, it = _single(monkeypatch, graphs, poll, solver, precond, dts)
        assert it == ref_it, (graphs, poll)
        for k in range(len(dts)):
            assert np.array_equal(h[k], ref_h[k]), (graphs, poll, k)


@pytest.mark.parametrize("solver,precond,precision", [("bicgstab", "jacobi", "fp64"), ("cg", "jacobi", "fp64"),
                                                     ("bicgstab", "none", "fp64"), ("cg", "none", "fp64"),
                                                     ("bicgstab", "jacobi", "mixed"), ("cg", "jacobi", "mixed")])
def test_finish_free_krylov_bitwise(monkeypatch, solver, precond, precision):
    """CG / BiCGSTAB iterations without finish kernels (each kernel re-reduces
    its producer's partial rows; scalars in an iteration history) against one
    finish kernel per reduction (WSPM_KRYLOV=unfolded): the same sums in the
    same order, so bitwise the same trajectories and iteration counts."""
    dts = [1e-4] * 4 + [2e-4, 2e-4]
    print(f"[case] finish-free {solver}+{precond} {precision}", flush=True)
    ref_h, ref_it = _single(monkeypatch, "0", "spin", solver, precond, dts, "unfolded", precision)
    h, it = _single(monkeypatch, "0", "spin", solver, precond, dts, "folded", precision)
    assert it == ref_it
    for k in range(len(dts)):
        assert np.array_equal(h[k], ref_h[k]), k


def _ensemble(monkeypatch, graphs, poll, solver, steps, lists="fused", precision="fp64"):
    from jax_wspm_amd.ensemble import Ensemble, vg_ensemble
    monkeypatch.setenv("WSPM_GRAPHS", graphs)
    monkeypatch.setenv("WSPM_POLL", poll)
    monkeypatch.setenv("WSPM_ENS_LIST", lists)
    insts = vg_ensemble(n_inst=6, sides=(16, 32, 24, 40, 20, 36), seed=11)
    ens = Ensemble(insts)
    ens.richards_init("vg", solver, "jacobi", precision=precision)
    hs, its = [], []
    for _ in range(steps):
        st = ens.step(1e-4, per_instance=True)
        its.append((st["picard_iters_i"].copy(), st["krylov_iters_i"].copy()))
        hs.append(ens.state()[0].cpu().numpy().copy())
    ens.close()
    return hs, its


@pytest.mark.parametrize("solver", ["bicgstab", "cg"])
def test_ensemble_modes_bitwise(monkeypatch, solver):
    print(f"[case] ensemble {solver}", flush=True)
    ref_h, ref_it = _ensemble(monkeypatch, "0", "sync", solver, 6)
    for graphs, poll in MODES[:2]:
        h, it = _ensemble(monkeypatch, graphs, poll, solver, 6)
        for k in range(len(h)):
            assert np.array_equal(it[k][0], ref_it[k][0]) and np.array_equal(it[k][1], ref_it[k][1]), (graphs, k)
            assert np.array_equal(h[k], ref_h[k]), (graphs, poll, k)


@pytest.mark.parametrize("solver,precision", [("bicgstab", "fp64"), ("cg", "fp64"), ("bicgstab", "mixed"),
                                              ("cg", "mixed")])
def test_ensemble_list_rebuild_bitwise(monkeypatch, solver, precision):
    """Krylov lists rebuilt inside the finish kernel (entry order varies from
    run to run) against a separate compaction kernel (WSPM_ENS_LIST=kernel):
    every entry is computed on its own, so bitwise the same trajectories and
    per-instance iteration counts, run after run."""
    print(f"[case] ensemble list rebuild {solver} {precision}", flush=True)
    ref_h, ref_it = _ensemble(monkeypatch, "0", "spin", solver, 6, "kernel", precision)
    for _ in range(2):
        h, it = _ensemble(monkeypatch, "0", "spin", solver, 6, "fused", precision)
        for k in range(len(h)):
            assert np.array_equal(it[k][0], ref_it[k][0]) and np.array_equal(it[k][1], ref_it[k][1]), k
            assert np.array_equal(h[k], ref_h[k]), k


def _single_ahead(monkeypatch, ahead, dts, **kw):
    from jax_wspm_amd.device import Context
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    monkeypatch.setenv("WSPM_GRAPHS", "0")
    monkeypatch.setenv("WSPM_POLL", "spin")
    monkeypatch.setenv("WSPM_AHEAD", ahead)
    pb = from_config(baseline_vg_config(40, 2, solver="bicgstab", precond="jacobi", nsteps=1, dt=1e-4,
                                        ks_field={"sigma": 0.5, "length": 0.05, "seed": 3}))
    ctx = Context(pb.points, pb.elements)
    ctx.set_dirichlet(pb.bc_lists)
    ctx.richards_init("vg", pb.phi, "bicgstab", "jacobi", pb.h0, Ks_elem=pb.Ks_elem, **kw)
    hs, its = [], []
    for dt in dts:
        st = ctx.richards_step(dt)
        its.append((st["picard_iters"], st["krylov_iters"], st["converged"]))
        hs.append(ctx.richards_state()[0].cpu().numpy().copy())
    ctx.close()
    return hs, its


@pytest.mark.parametrize("case,kw", [("default", {}),
                                     ("converges_in_one", {"picard_tol": 1.0}),
                                     ("two_allowed", {"picard_maxit": 2}),
                                     ("krylov_cut", {"maxiter": 1}),
                                     ("absolute", {"picard_mode": "abs", "picard_tol": 1e-3})])
def test_picard_iteration_ahead_bitwise(monkeypatch, case, kw):
    """The second Picard iteration enqueued ahead of the first one's
    read-back, behind a gate the first one's finish opens, against the plain
    order (WSPM_AHEAD=0): bitwise the same trajectories and counts, also where
    the gate stays shut (converged after one iteration, Krylov solve cut
    short, a limit of two iterations)."""
    dts = [1e-4] * 3 + [5e-4, 1e-4, 2e-3, 1e-4]
    print(f"[case] picard ahead {case}", flush=True)
    ref_h, ref_it = _single_ahead(monkeypatch, "0", dts, **kw)
    h, it = _single_ahead(monkeypatch, "1", dts, **kw)
    assert it == ref_it
    for k in range(len(dts)):
        assert np.array_equal(h[k], ref_h[k]), k


@pytest.mark.parametrize("case,kw,dts", [
    ("default", {}, [1e-4] * 3 + [5e-4, 1e-4]),
    # many BiCGSTAB iterations per Picard iteration: the chunk the host
    # enqueues first is too short, so the solve is continued after the read-back
    ("long_solves", {"tol": 1e-14, "maxiter": 400}, [2e-2, 5e-2, 1e-1])])
def test_picard_ahead_blocking_readback_bitwise(monkeypatch, case, kw, dts):
    """The host blocks on the stream instead of spinning (WSPM_SPIN_US=0), so
    the stream drains past the first Picard iteration's publication and the
    second iteration, enqueued ahead, has published too by the time the host
    reads the sequence word: the read-back must take the first iteration's
    state from its own mirror and go on.  Bitwise equal to the plain order."""
    print(f"[case] picard ahead, blocking read-back, {case}", flush=True)
    monkeypatch.setenv("WSPM_SPIN_US", "0")
    ref_h, ref_it = _single_ahead(monkeypatch, "0", dts, **kw)
    h, it = _single_ahead(monkeypatch, "1", dts, **kw)
    assert it == ref_it
    assert any(p >= 2 for p, _, _ in it)      # the gate opened
    for k in range(len(dts)):
        assert np.array_equal(h[k], ref_h[k]), k
