"""Benchmark of the implicit FEM time step on one MI355X.

Workload (default): plain FP64 2D Richards at scale -- 512 x 512 structured
triangular mesh of the unit square (263,169 nodes, 524,288 elements), van
Genuchten alpha=0.1, n=2, m=0.5, theta_r=0.06, theta_s=0.40, per-element
Ks = exp(N(ln 1e-3, 0.5^2)) spatially correlated (Gaussian kernel, length
0.05, seed 0), h0 = -1000, top Dirichlet -75, fixed dt = 1e-4,
Jacobi-BiCGSTAB (rtol 1e-10, maxiter 100), Picard relative 1e-6.

A step is one implicit time step (Picard iterations to convergence, each an
assembly + a Krylov solve).  value = steps/s over the timed steps, inputs
resident in HBM.  The SpMV launches inside the timed region are bracketed by
HIP events on the solver's stream; their mean duration gives the roofline
line.  The CPU baseline runs the oracle (NumPy/SciPy FP64 restatement of the
reference, one thread) on the first steps of the same run.

Workload c4: the discrete adjoint of the flow at 256 x 256 (200 steps of dt
1e-3, theta misfit at 64 nodes, checkpoint every 20 steps); value = time
steps of forward + backward per second; the gradient is checked on 16 random
directions by central differences.

Workload c3: the batched ensemble -- 256 independent 2D instances, side N
drawn from {32, 48, ..., 256}, per-instance van Genuchten parameters
(jax_wspm_amd.ensemble.vg_ensemble, seed 0), 100 steps of dt 1e-4,
Jacobi-BiCGSTAB FP64, all instances in one ensemble context.  value =
instance time steps/s (instances x steps / s); the plain single-instance GPU
path is timed on every instance for comparison.

    python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--workload c1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time steps/s + SpMV HBM GB/s (% peak): batched/mixed vs plain FP64"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (description, cells, dim, solver, precision, ks_field, dt)
    "c1": ("2D Richards 512x512 P1 triangles, heterogeneous log-normal Ks, fixed dt 1e-4, "
           "Jacobi-BiCGSTAB, FP64", 512, 2, "bicgstab", "fp64", {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-4),
    "c2": ("3D Richards 64^3 Kuhn tetrahedra, fixed dt 1e-4, Jacobi-CG, FP64", 64, 3, "cg", "fp64", None, 1e-4),
    "c2-mixed": ("3D Richards 64^3 Kuhn tetrahedra, fixed dt 1e-4, Jacobi-CG, FP32 inner / FP64 outer",
                 64, 3, "cg", "mixed", None, 1e-4),
    "c1-mixed": ("2D Richards 512x512, heterogeneous Ks, Jacobi-BiCGSTAB, FP32 inner / FP64 outer",
                 512, 2, "bicgstab", "mixed", {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-4),
    "c3": ("batched ensemble: 256 independent 2D van Genuchten instances, side N in {32,48,...,256}, "
           "per-instance alpha, n, Ks, theta_r, theta_s, fixed dt 1e-4, Jacobi-BiCGSTAB, FP64",
           None, 2, "bicgstab", "fp64", None, 1e-4),
    "c3-mixed": ("batched ensemble: 256 independent 2D van Genuchten instances, side N in {32,48,...,256}, "
                 "per-instance alpha, n, Ks, theta_r, theta_s, fixed dt 1e-4, Jacobi-BiCGSTAB, FP32 inner / "
                 "FP64 outer (per-instance iterative refinement to the FP64 tolerance)",
                 None, 2, "bicgstab", "mixed", None, 1e-4),
    "c4": ("discrete adjoint of 2D Richards flow + NH4 -> NO2 -> NO3 transport: 256x256 triangles "
           "(131,072 elements), log-normal Ks, rates 0.2/0.5/0.05, rho Kd 1.5 x 3.5 (NH4), dispersivities 0.5/0.1, "
           "200 steps dt 1e-3, squared misfit of theta and NO3 at 64 seeded nodes, gradient w.r.t. per-element "
           "log Ks, alpha, n, checkpoint every 20 steps, FP64", 256, 2, "bicgstab", "fp64",
           {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-3),
}
DEFAULT_STEPS = {"c3": 100, "c3-mixed": 100, "c4": 200}

# HBM-side bytes per working launch of the dominant SpMV (k_spmv<8,double,BicgQ>),
# from two rocprofv3 --pmc runs of this script with --calibrate-copy: FETCH_SIZE
# (KiB, x2.000 by the 256 MiB copy calibration) + WRITE_SIZE (KiB, x1.000).
# Counter values are not readable from inside the process, so the figure of
# the profiled run is reported here.
MEASURED_TRAFFIC = {"c1": 28.59e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; c3: 100)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--no-timing", action="store_true", help="do not bracket SpMV launches with events")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device-to-device (a kernel of known traffic for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE scaling)")
    ap.add_argument("--slow-group", type=lambda v: [int(k) for k in v.split(",")], default=[8],
                    help="c3: instances (the slowest Picard loops of a probe) stepped in a second context and "
                         "stream; a,b: the a slowest and the next b in two more; 0 = one context")
    ap.add_argument("--skip-errors", action="store_true",
                    help="c3-mixed: skip the untimed error runs against the FP64 ensemble")
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="c3-mixed: relative tolerance of the FP32 inner solves (each refinement sweep)")
    ap.add_argument("--fd-dirs", type=int, default=16, help="c4: random directions of the finite-difference check")
    ap.add_argument("--fd-eps", type=float, default=None,
                    help="c4: step of the difference quotients (default 5e-4 at order 6, 1e-4 at orders 2 and 4)")
    ap.add_argument("--fd-order", type=int, default=6, choices=(2, 4, 6),
                    help="c4: order p of the central-difference stencil (p forward runs per direction)")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="c3: steps per instance of the plain single-instance path (0 = skip)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = DEFAULT_STEPS.get(args.workload, 200)
    return args


def spmv_bytes(n, nnz, value_bytes=8):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row pointers
    + x gather (once per entry of x) + y store."""
    return (value_bytes + 4) * nnz + 4 * (n + 1) + 2 * value_bytes * n


def cpu_baseline(pb, steps):
    """Oracle restatement of the reference path, timed on this host (1 thread)."""
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(1)
    except Exception:  # pragma: no cover
        limiter = None
    from oracle import picard as opicard
    opb = opicard.Problem(points=pb.points, elems=pb.elements, soil_model=pb.soil_model, phi=pb.phi,
                          bc_lists=pb.bc_lists, h0=pb.h0, dt0=pb.dt0, nsteps=steps, adaptive=False,
                          picard_tol=pb.picard_tol, picard_maxit=pb.picard_maxit, picard_mode=pb.picard_mode,
                          solver=pb.solver, precond=pb.precond, tol=pb.tol, maxiter=pb.maxiter,
                          restart=pb.restart, Ks_elem=pb.Ks_elem)
    st = opicard.Stepper(opb)
    import numpy as np
    h = pb.h0.astype(np.float64).copy()
    t0 = time.perf_counter()
    for _ in range(steps):
        h, _, _, _ = st.step(h, pb.dt0)
    el = time.perf_counter() - t0
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    return steps / el, el, h


def roofline(n, nnz, launches, spmv_ms, rows, traffic=None, value_bytes=8, fmt="csr"):
    """Mean HBM rate of the timed SpMV launches: algorithmic bytes of the rows
    each launch worked on (the full operator's bytes per row) over their
    summed duration.  value_bytes 4: the FP32 SpMVs of a mixed-precision run."""
    if not launches:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None}
    full = spmv_bytes(n, nnz, value_bytes)
    moved = full * (rows / n)
    achieved = moved / (spmv_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": (f"{'k_spmv_ell (ELL' if fmt == 'ell' else 'k_spmv (CSR'} SpMV, "
                       f"{'FP32' if value_bytes == 4 else 'FP64'}, fused Krylov epilogues)"),
            "bytes_per_launch": round(moved / launches),
            "full_operator_bytes": full, "launches": launches,
            "avg_launch_us": round(spmv_ms / launches * 1e3, 3)}


def spmv_timing_region(timing, timing_read, step, dt, steps, fp32):
    """Run `steps` more steps with start/stop HIP events attached to every
    SpMV dispatch on the solver's stream (the kernel's own execution time);
    returns (working launches, their summed ms, rows worked on, wall seconds
    of the region)."""
    import torch
    if steps <= 0:
        return 0, 0.0, 0, 0.0
    torch.cuda.synchronize()
    timing(True, fp32=fp32)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    launches, spmv_ms, rows = timing_read()
    timing(False)
    return launches, spmv_ms, rows, el


def launch_note():
    graphs = os.environ.get("WSPM_GRAPHS") == "1"
    poll = "copy + stream synchronisation" if os.environ.get("WSPM_POLL") == "sync" else \
        "a device-written sequence word in pinned memory"
    return (("Picard iterations replayed from HIP graphs" if graphs else "stream launches") +
            f"; the host reads the solver state back through {poll}")


def timing_note(steps, el):
    return (f"roofline from {steps} further steps of the same run, every SpMV launched with start/stop HIP "
            f"events attached to its dispatch (hipExtLaunchKernelGGL; {el / max(steps, 1) * 1e3:.4f} ms/step there; "
            "ensemble groups stepped one after another there)")


def run_ensemble(args):
    """c3: the whole ensemble (one context, or two on separate streams with the
    slowest instances in the second, --slow-group); the plain single-instance
    path and the oracle are timed on the same instances for comparison."""
    import numpy as np
    import torch

    from jax_wspm_amd.device import Context
    from jax_wspm_amd.ensemble import Ensemble, EnsembleGroups, vg_ensemble

    desc, _, _, solver, precision, _, dt = WORKLOADS[args.workload]
    insts = vg_ensemble(256, seed=0)
    mixed = precision == "mixed"
    groups = None
    if sum(args.slow_group) > 0:
        # the instances with the most Picard iterations over a few probe steps
        # (untimed, from the initial state) get a context and stream of their own
        probe = Ensemble(insts)
        probe.richards_init("vg", solver, "jacobi", precision=precision, inner_tol=args.inner_tol)
        counts = np.zeros(len(insts))
        for _ in range(max(args.warmup, 1)):
            counts += probe.step(dt, per_instance=True)["picard_iters_i"]
        probe.close()
        del probe
        groups = EnsembleGroups.split_slowest(counts, args.slow_group)
        ens = EnsembleGroups(insts, groups)
    else:
        ens = Ensemble(insts)
    ens.richards_init("vg", solver, "jacobi", precision=precision, inner_tol=args.inner_tol)
    print(f"[bench] {args.workload}: instances={ens.n_inst} n={ens.n} E={ens.n_elems} nnz={ens.nnz} lanes={ens.lanes}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ens.step(dt)
    torch.cuda.synchronize()
    picard = kry = 0
    conv = True
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = ens.step(dt)
        picard += st["picard_iters"]
        kry += st["krylov_iters"]
        conv &= st["converged"]
        if (k + 1) % 25 == 0:
            print(f"[bench] c3 step {k + 1}/{args.steps}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # per-launch timing: groups stepped one after another, so that launches of
    # different streams do not share the GPU while they are timed
    step_seq = (lambda d: ens.step(d, concurrent=False)) if groups else ens.step
    launches, spmv_ms, rows, el_ev = spmv_timing_region(ens.timing, ens.timing_read, step_seq, dt,
                                                         0 if args.no_timing else args.steps, mixed)
    h = ens.state()[0]
    assert bool(torch.isfinite(h).all()), "non-finite pressure head"
    value = ens.n_inst * args.steps / el
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "time steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (structured meshes, seeded per-instance soil parameters generated in-process)",
        "config": {"workload": desc, "instances": ens.n_inst, "nodes": ens.n, "elements": ens.n_elems,
                   "nnz": ens.nnz, "dt": dt, "solver": f"{solver}+jacobi", "precision": precision,
                   **({"inner_tol": args.inner_tol} if mixed else {}),
                   "value_counts": "instance time steps (each timed step advances every instance once)",
                   "ensemble_steps_per_s": round(args.steps / el, 3),
                   "max_picard_iters_per_step": round(picard / args.steps, 3),
                   "krylov_iters_per_step_summed": round(kry / args.steps, 1),
                   "all_converged": bool(conv), "spmv_format": ens.spmv_format,
                   "launch": launch_note(),
                   "timing_region": timing_note(args.steps, el_ev),
                   "contexts": ([len(g) for g in groups] if groups else [len(insts)]),
                   "grouping": (f"instances ranked by Picard iterations over {max(args.warmup, 1)} probe "
                                f"steps; the slowest {'/'.join(map(str, args.slow_group))} in ensemble contexts of "
                                "their own, each on its own stream stepped by its own host thread"
                                if groups else "one context")},
        "roofline": roofline(ens.n, ens.nnz, launches, spmv_ms, rows, value_bytes=4 if mixed else 8,
                             fmt=ens.spmv_format),
    }
    if mixed and not args.skip_errors:
        # error against the FP64 ensemble: both forms rerun untimed for the
        # same number of steps with per-instance Picard counts recorded, plus an
        # FP64 run with the Krylov tolerance at 1e-13 ("tight").  Both forms
        # stop their solves at a 1e-10 relative residual along different
        # iterates, so on an ill-conditioned instance (dry soil: C ~ 0) their
        # solutions differ by up to cond x 1e-10; the tight run shows how far
        # each form is from the converged solution.
        ens.close()
        res = {}
        for prec, ktol in (("mixed", 1e-10), ("fp64", 1e-10), ("fp64", 1e-13), ("mixed", 1e-12), ("fp64", 1e-12)):
            e2 = Ensemble(insts)
            e2.richards_init("vg", solver, "jacobi", precision=prec, tol=ktol, inner_tol=args.inner_tol)
            its = [e2.step(dt, per_instance=True)["picard_iters_i"].copy() for _ in range(args.warmup + args.steps)]
            h2, th2, _ = e2.state()
            res[(prec, ktol)] = (np.array(its), e2.split(h2), e2.split(th2))
            e2.close()
            del e2
        (it32, h32, th32), (it64, h64, th64) = res[("mixed", 1e-10)], res[("fp64", 1e-10)]
        _, hT, _ = res[("fp64", 1e-13)]

        def rel(a, b):
            return np.array([np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(a, b)])
        same = np.all(it32 == it64, axis=0)
        eh, et = rel(h32, h64), rel(th32, th64)
        e32T, e64T = rel(h32, hT), rel(h64, hT)
        cfg = out["config"]
        cfg["instances_with_identical_picard_histories"] = int(same.sum())
        cfg["max_rel_l2_vs_fp64_h"] = float(eh.max())
        cfg["max_rel_l2_vs_fp64_theta"] = float(et.max())
        cfg["median_rel_l2_vs_fp64_h"] = float(np.median(eh))
        cfg["instances_within_1e-8_of_fp64_h"] = int((eh <= 1e-8).sum())
        cfg["max_rel_l2_h_vs_tight_fp64"] = {"mixed": float(e32T.max()), "fp64": float(e64T.max())}
        cfg["median_rel_l2_h_vs_tight_fp64"] = {"mixed": float(np.median(e32T)), "fp64": float(np.median(e64T))}
        cfg["error_sample"] = f"every instance after {args.warmup + args.steps} steps; tight = FP64, Krylov tol 1e-13"
        # both forms with the Krylov tolerance at 1e-12: the dry instances'
        # conditioning no longer hides the FP32 inner solves' effect
        (i32b, h32b, _), (i64b, h64b, _) = res[("mixed", 1e-12)], res[("fp64", 1e-12)]
        sameb = np.all(i32b == i64b, axis=0)
        ehb = rel(h32b, h64b)
        cfg["krylov_tol_1e-12"] = {
            "max_rel_l2_mixed_vs_fp64_h": float(ehb.max()),
            "median_rel_l2_mixed_vs_fp64_h": float(np.median(ehb)),
            "instances_within_1e-8": int((ehb <= 1e-8).sum()),
            "instances_with_identical_picard_histories": int(sameb.sum()),
            "max_rel_l2_h_identical_histories": float(ehb[sameb].max()) if sameb.any() else None}
    else:
        ens.close()
    del ens
    if args.plain_steps > 0:
        # the plain FP64 single-instance path on every instance (one context each)
        tot = 0.0
        for i, inst in enumerate(insts):
            ctx = Context(inst.points, inst.elements)
            ctx.set_dirichlet(inst.bc_lists)
            ctx.richards_init("vg", inst.phi, solver, "jacobi", inst.h0)
            ctx.richards_step(dt)            # warm-up step (first Krylov chunk sizing)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.plain_steps):
                ctx.richards_step(dt)
            torch.cuda.synchronize()
            tot += time.perf_counter() - t1
            ctx.close()
            if (i + 1) % 64 == 0:
                print(f"[bench] c3 plain path {i + 1}/256", file=sys.stderr, flush=True)
        plain = len(insts) * args.plain_steps / tot
        out["config"]["plain_fp64_single_instance_steps_per_s"] = round(plain, 2)
        out["config"]["speedup_vs_plain_fp64"] = round(value / plain, 2)
        out["config"]["plain_sample"] = (f"steps 2..{1 + args.plain_steps} of each of the 256 instances, "
                                         "one context per instance, run one after another")
    if args.cpu_steps > 0:
        from oracle import picard as opicard
        try:
            from threadpoolctl import threadpool_limits
            threadpool_limits(1)
        except Exception:  # pragma: no cover
            pass
        sample = insts[:8]
        t1 = time.perf_counter()
        for inst in sample:
            opb = opicard.Problem(points=inst.points, elems=inst.elements, soil_model="vg", phi=inst.phi,
                                  bc_lists=inst.bc_lists, h0=inst.h0, dt0=dt, nsteps=args.cpu_steps,
                                  adaptive=False, solver=solver, precond="jacobi")
            stp = opicard.Stepper(opb)
            hh = inst.h0.astype(np.float64).copy()
            for _ in range(args.cpu_steps):
                hh, _, _, _ = stp.step(hh, dt)
        cel = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(len(sample) * args.cpu_steps / cel, 4), "unit": "time steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"first {args.cpu_steps} steps of instances 0-7 "
                                         f"(sides {[i.cells for i in sample]}), {cel:.1f} s, NumPy/SciPy FP64 "
                                         f"oracle, 1 thread; host has {os.cpu_count()} logical CPUs"}
    print(json.dumps(out))


# c4 species settings: rates per species, bulk density x Kd of NH4,
# dispersivities, Cauchy inlet on the top boundary (NH4 and NO3 in the inflow)
C4_SPECIES = dict(DL=0.5, DT=0.1, Dm=0.0, mu=(0.2, 0.5, 0.05), rho_kd=1.5 * 3.5, cauchy_flux=-0.05,
                  c_inlet=(1.0, 0.0, 0.4))


def run_adjoint(args):
    """c4: coupled flow + nitrogen forward and the checkpointed adjoint
    backward of L = sum (theta - theta_obs)^2 + sum (NO3 - NO3_obs)^2, the
    observations coming from a run at perturbed parameters.  The gradient is
    checked on random directions by central differences of the same forward
    (tight Picard / Krylov tolerances so the differences resolve the
    converged-step model)."""
    import numpy as np
    import torch

    from jax_wspm_amd import mesh as M
    from jax_wspm_amd.adjoint import NitrogenAdjointModel

    desc, cells, _, _, _, ksf, dt = WORKLOADS["c4"]
    steps = args.steps
    pts, tri = M.unit_square(cells, 1.0)
    top = np.nonzero(pts[:, 1] == 1.0)[0]
    log_ks = np.log(M.correlated_log_ks(pts, tri, cells, mean_ks=1e-2, sigma=ksf["sigma"], length=ksf["length"],
                                        seed=ksf["seed"]))
    rng = np.random.default_rng(0)
    upper = np.nonzero((pts[:, 1] >= 0.85) & (pts[:, 1] < 1.0))[0]
    obs = np.sort(rng.choice(upper, size=64, replace=False))
    obs_steps = list(range(20, steps + 1, 20)) or [steps]
    model = NitrogenAdjointModel(pts, tri, [(top, np.full(top.shape, -5.0))], np.full(len(pts), -100.0), dt, steps,
                                 obs, obs_steps, 0.40, 0.06, inlet_nodes=top, checkpoint=20, tol=1e-13, maxiter=400,
                                 picard_tol=1e-12, picard_maxit=300, adj_tol=1e-12, tr_tol=1e-13, **C4_SPECIES)
    lk = torch.tensor(log_ks, device="cuda")
    alpha, n = 0.1, 2.0
    print(f"[bench] c4: n={model.n} E={model.E} steps={steps} obs={len(obs)}x{len(obs_steps)}", file=sys.stderr,
          flush=True)
    # synthetic observations: the same model at perturbed parameters
    y_obs = model.forward(lk + 0.2 * torch.tensor(np.sin(8 * pts[tri].mean(axis=1)[:, 0]), device="cuda"),
                          alpha * 1.1, n * 0.95).detach()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = model.forward(lk, alpha, n)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    dldy = 2.0 * (y - y_obs)
    t0 = time.perf_counter()
    g_ks, g_a, g_n = model.backward(dldy)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    print(f"[bench] c4 forward {t_fwd:.2f} s backward {t_bwd:.2f} s", file=sys.stderr, flush=True)
    g = np.concatenate([g_ks.cpu().numpy(), [float(g_a), float(g_n)]])
    misfit = {"theta": float(((y[0] - y_obs[0]) ** 2).sum()), "NO3": float(((y[1] - y_obs[1]) ** 2).sum())}
    # element Peclet / Courant numbers of the last step's fields
    h_last = model.ctx.richards_state()[0]
    ks = torch.exp(lk)
    f = model.ctx.transport_fields(model.phi(alpha, n), h_last, model.DL, model.DT, model.Dm, Ks_elem=ks)
    pe, cr = model.ctx.peclet_courant(f[0], f[1], f[6], f[3], f[5], dt, all_elements=True)

    def loss(p):
        with torch.no_grad():
            yy = model.forward(torch.tensor(p[:-2], device="cuda"), float(p[-2]), float(p[-1]))
            return float(((yy - y_obs) ** 2).sum())

    p0 = np.concatenate([log_ks, [alpha, n]])
    scale = np.concatenate([np.ones(len(log_ks)), [alpha, n]])
    errs, nerrs = [], []
    drng = np.random.default_rng(1)
    for k in range(args.fd_dirs):
        d = drng.normal(size=p0.size) * scale
        # error budget of the difference quotient: truncation ~eps^p for the
        # order-p stencil plus the forward's converged-tolerance noise (about
        # 1e-11 in the loss) divided by eps.  Measured on these directions:
        # order 2 at eps 1e-4, up to 1.6e-6 truncation; order 4 at 1e-4,
        # noise of up to 1e-6 in the direction with the smallest derivative
        # (g.d 0.41); order 4 at 5e-4, 8.9e-7 truncation in the most
        # nonlinear direction.  Order 6 at 5e-4 keeps both terms below 1e-7.
        eps = args.fd_eps if args.fd_eps else {2: 1e-4, 4: 1e-4, 6: 5e-4}[args.fd_order]

        def dl(k):
            return loss(p0 + k * eps * d) - loss(p0 - k * eps * d)

        if args.fd_order == 2:
            fd = dl(1) / (2 * eps)
        elif args.fd_order == 4:
            fd = (8.0 * dl(1) - dl(2)) / (12 * eps)
        else:
            fd = (45.0 * dl(1) - 9.0 * dl(2) + dl(3)) / (60 * eps)
        errs.append(abs(g @ d - fd) / max(abs(fd), 1e-300))
        nerrs.append(abs(g @ d - fd) / (np.linalg.norm(g) * np.linalg.norm(d)))
        print(f"[bench] c4 direction {k + 1}/{args.fd_dirs}: g.d {g @ d:.6e} fd {fd:.6e} rel err {errs[-1]:.2e}",
              file=sys.stderr, flush=True)
    model.close()
    out = {
        "metric": METRIC, "value": round(steps / (t_fwd + t_bwd), 3), "unit": "time steps/s", "n_gpus": 1,
        "steps": steps, "warmup": 1, "ms_per_step": round((t_fwd + t_bwd) / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (structured mesh, seeded Ks field and observation nodes, observations from a run at "
                "perturbed parameters)",
        "config": {"workload": desc, "nodes": model.n, "elements": model.E,
                   "value_counts": "time steps of forward + adjoint backward (one gradient) per second",
                   "forward_s": round(t_fwd, 3), "backward_s": round(t_bwd, 3),
                   "backward_over_forward": round(t_bwd / t_fwd, 2),
                   "misfit": misfit,
                   "peclet_max": pe, "courant_max": cr,
                   "flow_adjoint_bicgstab_iters_per_step": round(float(np.mean(model.adj_iters)), 2),
                   "species_adjoint_bicgstab_iters_per_step": round(float(np.mean(model.adj_tr_iters)), 2),
                   "fd_directions": args.fd_dirs,
                   "fd_stencil": f"central, order {args.fd_order}, eps {eps:g}",
                   "fd_max_rel_err": float(max(errs)) if errs else None,
                   "fd_median_rel_err": float(np.median(errs)) if errs else None,
                   "fd_max_err_over_norms": float(max(nerrs)) if nerrs else None,
                   "tolerances": "Picard rel 1e-12, Krylov 1e-13 (flow and species), adjoint solves 1e-12"},
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": f"this build runs on one GPU; --gpus {args.gpus} is not supported "
                                   "(multi-GPU sharding of meshes and ensembles is a design, not built)"}))
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload in ("c3", "c3-mixed"):
        run_ensemble(args)
        return
    if args.workload == "c4":
        run_adjoint(args)
        return
    import numpy as np
    import torch

    from jax_wspm_amd.problems import baseline_vg_config, from_config
    from jax_wspm_amd.richards import RichardsSolver

    desc, cells, dim, solver, precision, ksf, dt = WORKLOADS[args.workload]
    cfg = baseline_vg_config(cells, dim, solver=solver, precond="jacobi", precision=precision, ks_field=ksf,
                             nsteps=args.warmup + args.steps, dt=dt)
    pb = from_config(cfg)
    rs = RichardsSolver(pb, history="last")
    ctx = rs.ctx
    rs.init_state()
    print(f"[bench] {args.workload}: n={ctx.n} E={ctx.n_elems} nnz={ctx.nnz} lanes={ctx.lanes}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        ctx.richards_step(dt)
    torch.cuda.synchronize()
    picard = kry = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.richards_step(dt)
        picard += st["picard_iters"]
        kry += st["krylov_iters"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # SpMV launch durations: the next steps of the same run with HIP events
    # attached to every SpMV dispatch (graph replay is off while they are on)
    launches, spmv_ms, rows, el_ev = spmv_timing_region(ctx.timing, ctx.timing_read, ctx.richards_step, dt,
                                                         0 if args.no_timing else args.steps, precision == "mixed")
    h, th, _ = ctx.richards_state()
    assert bool(torch.isfinite(h).all()), "non-finite pressure head"
    value = args.steps / el
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32+f64",
        "data": "synthetic (structured mesh + seeded correlated Ks field generated in-process)",
        "config": {"workload": desc, "nodes": ctx.n, "elements": ctx.n_elems, "nnz": ctx.nnz, "dt": dt,
                   "solver": f"{solver}+jacobi", "precision": precision,
                   "picard_iters_per_step": round(picard / args.steps, 3),
                   "krylov_iters_per_step": round(kry / args.steps, 3),
                   "launch": launch_note(),
                   "timing_region": timing_note(args.steps, el_ev)},
        # mixed precision: the FP32 SpMVs of the inner solves, at FP32 bytes
        "roofline": roofline(ctx.n, ctx.nnz, launches, spmv_ms, rows, MEASURED_TRAFFIC.get(args.workload),
                             value_bytes=4 if precision == "mixed" else 8, fmt=ctx.spmv_format),
    }
    if args.cpu_steps > 0:
        pb_cpu = from_config(baseline_vg_config(cells, dim, solver=solver, precond="jacobi", ks_field=ksf,
                                                nsteps=args.cpu_steps, dt=dt))
        cps, cel, _ = cpu_baseline(pb_cpu, args.cpu_steps)
        out["cpu_baseline"] = {"value": round(cps, 4), "unit": "time steps/s", "cores": 1, "kind": "port",
                               "sample": f"first {args.cpu_steps} steps of the same run ({cel:.1f} s), "
                                         f"NumPy/SciPy FP64 oracle, 1 thread; host has {os.cpu_count()} "
                                         "logical CPUs"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
