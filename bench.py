"""Benchmark of the implicit FEM time step on one MI355X.

Default workload (c3, the configuration the metric is quoted on: batched /
mixed vs plain FP64): 256 independent 2D van Genuchten instances on
structured triangular meshes, side N drawn from {32, 48, ..., 256}, alpha ~
U(0.01, 0.2), n ~ U(1.3, 3.0), Ks ~ logU(1e-4, 1e-2), theta_r ~ U(0.02,
0.1), theta_s ~ U(0.35, 0.5) (jax_wspm_amd.ensemble.vg_ensemble, seed 0),
h0 = -1000, top Dirichlet -75, 100 fixed steps of dt 1e-4,
Jacobi-BiCGSTAB.  One run times, on the same instances:

* the batched mixed-precision form (``value``): FP32 BiCGSTAB corrections of
  the FP64 residual per instance, refined until that residual is below 1e-14
  of |b|, where the FP64 solution is itself converged (the FP64 solve at the
  reference's 1e-10 is up to 1.6e-3 away from it on the driest instances);
* the batched FP64 form at the same tolerance and at the reference's 1e-10;
* the plain FP64 single-instance path (one context per instance, one after
  another) at both tolerances.

Every form is 100 timed steps after the warm-up (the plain path: 3 steps per
instance).  The line reports each form's instance-steps/s, the speed-ups, and
per-instance relative L2 errors of the mixed form against the FP64 form after
the same steps.  ``roofline`` prices the dominant kernel (the fused assembly)
at its compulsory bytes, ``roofline_spmv`` the SpMV of the mixed form's inner
solves; both from HIP events around every launch in a further timed region of
the same run.  ``cpu_baseline`` runs the oracle (NumPy/SciPy, one thread) on
8 instances over the steps right after the warm-up, from the GPU's state
there.

Other workloads: c1 (512^2 FP64, heterogeneous Ks), c1-mixed, c2 (64^3 tets,
FP32 inner / FP64 outer, error against the FP64 form), c2-fp64, c3-fp64 /
c3-mixed (one ensemble form, for profiling), c4 (discrete adjoint of flow +
nitrogen at 256^2 with a finite-difference check).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]
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
DT = 1e-4
# Krylov tolerance at which the FP64 solution is converged on every instance
# of the ensemble (the reference's own 1e-10 leaves the driest instances up to
# 1.6e-3 away from it); the mixed form refines to it, and is compared with the
# FP64 form at the same tolerance
TIGHT_TOL = 1e-14
MIXED_INNER_TOL = 1e-6      # FP32 inner solves: relative tolerance per refinement sweep
MIXED_INNER_ITERS = 4       # and at most this many BiCGSTAB iterations

ENS_DESC = ("batched ensemble: 256 independent 2D van Genuchten instances, side N in {32,48,...,256}, "
            "per-instance alpha, n, Ks, theta_r, theta_s, h0 -1000, top Dirichlet -75, fixed dt 1e-4, "
            "Jacobi-BiCGSTAB")
WORKLOADS = {
    # name: (description, cells, dim, solver, precision, ks_field, dt)
    "c1": ("2D Richards 512x512 P1 triangles, heterogeneous log-normal Ks, fixed dt 1e-4, "
           "Jacobi-BiCGSTAB, FP64", 512, 2, "bicgstab", "fp64", {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-4),
    "c1-mixed": ("2D Richards 512x512, heterogeneous Ks, Jacobi-BiCGSTAB, FP32 inner / FP64 outer",
                 512, 2, "bicgstab", "mixed", {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-4),
    "c2": ("3D Richards 64^3 Kuhn tetrahedra, h0 -500, top Dirichlet -10, fixed dt 1e-4, Jacobi-CG, FP32 inner "
           "CG / FP64 Picard residual, update and iterative refinement", 64, 3, "cg", "mixed", None, 1e-4),
    "c2-fp64": ("3D Richards 64^3 Kuhn tetrahedra, fixed dt 1e-4, Jacobi-CG, FP64", 64, 3, "cg", "fp64", None, 1e-4),
    "c3": (ENS_DESC + "; batched mixed (FP32 inner / FP64 outer) vs batched FP64 vs plain FP64 single-instance",
           None, 2, "bicgstab", "mixed", None, DT),
    "c3-fp64": (ENS_DESC + ", FP64", None, 2, "bicgstab", "fp64", None, DT),
    "c3-mixed": (ENS_DESC + ", FP32 inner / FP64 outer", None, 2, "bicgstab", "mixed", None, DT),
    "c4": ("discrete adjoint of 2D Richards flow + NH4 -> NO2 -> NO3 transport: 256x256 triangles "
           "(131,072 elements), log-normal Ks, rates 0.2/0.5/0.05, rho Kd 1.5 x 3.5 (NH4), dispersivities 0.5/0.1, "
           "200 steps dt 1e-3, squared misfit of theta and NO3 at 64 seeded nodes, gradient w.r.t. per-element "
           "log Ks, alpha, n, checkpoint every 20 steps, FP64", 256, 2, "bicgstab", "fp64",
           {"sigma": 0.5, "length": 0.05, "seed": 0}, 1e-3),
}
DEFAULT_STEPS = {"c2": 100, "c2-fp64": 100, "c3": 100, "c3-fp64": 100, "c3-mixed": 100, "c4": 200}
# oracle steps of the CPU baseline: about 10-20 s of one host core each
DEFAULT_CPU_STEPS = {"c1": 10, "c1-mixed": 10, "c2": 3, "c2-fp64": 3, "c3": 15}
DEFAULT_INNER_ITERS = {"c1-mixed": 2}
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 100; c1: 200; c4: 200)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps for the CPU baseline (0 = skip; default: about 10-20 s of CPU work)")
    ap.add_argument("--no-timing", action="store_true", help="skip the event-timed roofline region")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device-to-device (a kernel of known traffic for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE scaling)")
    ap.add_argument("--slow-group", type=lambda v: [int(k) for k in v.split(",")], default=[8],
                    help="ensembles: instances (the slowest Picard loops of a probe) stepped in a second context "
                         "and stream; a,b: the a slowest and the next b in two more; 0 = one context")
    ap.add_argument("--inner", choices=["krylov", "jacobi"], default="jacobi",
                    help="mixed forms: FP32 inner solves of each refinement sweep (jacobi: Jacobi-preconditioned "
                         "Richardson steps; krylov: the method's BiCGSTAB / CG to the inner tolerance)")
    ap.add_argument("--inner-iters", type=int, default=None,
                    help="--inner jacobi: steps per refinement sweep (default 2 for c1-mixed, else 3: measured)")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="c3: timed steps per instance of the plain single-instance path (0 = skip)")
    ap.add_argument("--fd-dirs", type=int, default=16, help="c4: random directions of the finite-difference check")
    ap.add_argument("--fd-eps", type=float, default=None,
                    help="c4: step of the difference quotients (default 5e-4 at order 6, 1e-4 at orders 2 and 4)")
    ap.add_argument("--fd-order", type=int, default=6, choices=(2, 4, 6),
                    help="c4: order p of the central-difference stencil (p forward runs per direction)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = DEFAULT_STEPS.get(args.workload, 200)
    if args.cpu_steps is None:
        args.cpu_steps = DEFAULT_CPU_STEPS.get(args.workload, 0)
    if args.inner_iters is None:
        # Jacobi steps per sweep, measured per workload on the final kernels:
        # 512^2 one system 2 (5,923-6,214 steps/s against 5,692-5,784 at 3),
        # 64^3 and the ensemble 3 (2: 1,262-1,389 against 1,437-1,455 steps/s,
        # and 67.1k against 70.9k instance-steps/s)
        args.inner_iters = DEFAULT_INNER_ITERS.get(args.workload, 3)
    return args


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, value_bytes=8):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row pointers
    + x gather (once per entry of x) + y store."""
    return (value_bytes + 4) * nnz + 4 * (n + 1) + 2 * value_bytes * n


def assembly_bytes(n, n_elems, nnz, dim, ks_elem):
    """Compulsory bytes of one Picard iteration's assembly: coordinates,
    connectivity, h_m and theta_n, [per-element Ks,] the values of A and b."""
    return 8 * dim * n + 4 * (dim + 1) * n_elems + 16 * n + (8 * n_elems if ks_elem else 0) + 8 * nnz + 8 * n


def measured_traffic(kernel):
    """HBM bytes per row of a kernel from the committed PMC summary
    (profiles/pmc_traffic.json, written from two rocprofv3 --pmc runs of this
    script, FETCH_SIZE then WRITE_SIZE, scaled by a copy of known size), with
    the commit it was measured at; None if absent."""
    try:
        with open(PROFILE_TRAFFIC) as fh:
            return json.load(fh).get(kernel)
    except (OSError, ValueError):
        return None


def roofline(kernel, bytes_full, n, launches, ms, rows, traffic_key=None):
    """Mean HBM rate of the timed launches of one kernel: algorithmic bytes of
    the rows each launch worked on (bytes_full / n per row) over their summed
    duration (HIP events around each launch on the solver's stream)."""
    if not launches:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": kernel}
    moved = bytes_full * (rows / n)
    achieved = moved / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
           "bytes_per_launch": round(moved / launches), "launches": launches,
           "avg_launch_us": round(ms / launches * 1e3, 3), "rows_per_launch": round(rows / launches)}
    tr = measured_traffic(traffic_key) if traffic_key else None
    if tr:
        out["traffic"] = round(tr["bytes_per_row"] * rows / launches)
        out["traffic_source"] = (f"{tr['file']} (commit {tr['commit']}): {tr['bytes_per_row']:.1f} HBM bytes per "
                                 f"row x {round(rows / launches)} rows per launch")
    return out


def timed_region(obj, step, dt, steps, fp32, assembly=True):
    """`steps` more steps with HIP events around every SpMV (FP32 ones when
    fp32) and assembly launch; returns (spmv launches, ms, rows),
    (assembly launches, ms, rows), wall seconds of the region, and
    (launches, ms, rows) of the SpMV launches over at least a quarter of the
    operator's rows."""
    import torch
    if steps <= 0:
        return (0, 0.0, 0), (0, 0.0, 0), 0.0, (0, 0.0, 0)
    torch.cuda.synchronize()
    obj.timing(True, fp32=fp32, assembly=assembly)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    big = obj.timing_read_over(max(1, obj.n // 4))
    sp = obj.timing_read()
    asm = obj.timing_read_asm() if assembly else (0, 0.0, 0)
    obj.timing(False)
    return sp, asm, el, big


def launch_note():
    graphs = os.environ.get("WSPM_GRAPHS") == "1"
    poll = "copy + stream synchronisation" if os.environ.get("WSPM_POLL") == "sync" else \
        "a device-written sequence word in pinned memory"
    return (("Picard iterations replayed from HIP graphs" if graphs else "stream launches") +
            f"; the host reads the solver state back through {poll}")


# ---------------------------------------------------------------------------------------------
# ensembles
# ---------------------------------------------------------------------------------------------

def _ens_groups(insts, args):
    """Instance groups: the slowest Picard loops of a few FP64 probe steps in
    contexts of their own (EnsembleGroups), or one context."""
    import numpy as np
    from jax_wspm_amd.ensemble import Ensemble, EnsembleGroups
    if sum(args.slow_group) <= 0:
        return None
    probe = Ensemble(insts)
    probe.richards_init("vg", "bicgstab", "jacobi")
    counts = np.zeros(len(insts))
    for _ in range(max(args.warmup, 1)):
        counts += probe.step(DT, per_instance=True)["picard_iters_i"]
    probe.close()
    return EnsembleGroups.split_slowest(counts, args.slow_group)


def _ens_form(insts, groups, precision, tol, args, timing=True, inner=None):
    """One batched form: warm-up, timed steps, final state, then the
    event-timed roofline region.  Returns a dict of its figures."""
    import torch
    from jax_wspm_amd.ensemble import Ensemble, EnsembleGroups
    os.environ["WSPM_ENS_INNER"] = str(MIXED_INNER_ITERS)
    inner = inner or args.inner
    ens = EnsembleGroups(insts, groups) if groups else Ensemble(insts)
    ens.richards_init("vg", "bicgstab", "jacobi", precision=precision, tol=tol, inner_tol=MIXED_INNER_TOL,
                      inner=inner, inner_iters=args.inner_iters if inner == "jacobi" else 0)
    for _ in range(args.warmup):
        ens.step(DT)
    h_warm = ens.state()[0].cpu().numpy()
    torch.cuda.synchronize()
    picard = kry = 0
    conv = True
    lc0 = _launch_counts(ens)
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = ens.step(DT)
        picard += st["picard_iters"]
        kry += st["krylov_iters"]
        conv &= st["converged"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lc1 = _launch_counts(ens)
    h, th, _ = ens.state()
    members = ens.members if groups else [ens]
    out = {"rate": ens.n_inst * args.steps / el, "elapsed": el, "picard": picard / args.steps,
           "krylov": kry / args.steps, "converged": bool(conv), "h": ens.split(h), "theta": ens.split(th),
           "h_warm": h_warm, "n": ens.n, "n_elems": ens.n_elems, "nnz": ens.nnz, "format": ens.spmv_format,
           "shape": ens.shape, "launches": _launch_summary(lc0, lc1, args.steps),
           "windows": all(m.assembly_windows for m in members)}
    assert all(bool(torch.isfinite(x).all()) for x in (h, th)), "non-finite state"
    log(f"{precision}{' (' + inner + ' inner)' if precision == 'mixed' else ''} tol {tol:g}: {out['rate']:.0f} "
        f"instance-steps/s, {out['krylov']:.1f} summed Krylov its/step")
    if timing and not args.no_timing:
        seq = (lambda d: ens.step(d, concurrent=False)) if groups else ens.step
        sp, asm, el_ev, big = timed_region(ens, seq, DT, args.steps, precision == "mixed")
        out.update(spmv=sp, asm=asm, el_ev=el_ev, spmv_big=big)
    ens.close()
    return out


def _asm_kernel(windows):
    """The time-step assembly kernel's name."""
    return "k_assemble_win" if windows else "k_assemble_reg"


def _launch_counts(ens):
    """(launches, Picard rounds) per context of an Ensemble or EnsembleGroups."""
    lc = ens.launch_count()
    return lc if isinstance(lc, list) else [lc]


def _launch_summary(lc0, lc1, steps):
    """Kernel launches of the timed steps: per step over all contexts, and per
    Picard round of each context (its launches over the rounds it ran)."""
    dl = [b[0] - a[0] for a, b in zip(lc0, lc1)]
    dr = [b[1] - a[1] for a, b in zip(lc0, lc1)]
    return {"per_step": round(sum(dl) / steps, 1),
            "per_picard_round_by_context": [round(l / r, 1) if r else None for l, r in zip(dl, dr)],
            "picard_rounds_per_step_by_context": [round(r / steps, 2) for r in dr],
            "what": "kernel launches enqueued (device copies and graph replays one each), from wspm_launch_count"}


def _plain_rate(insts, tol, steps):
    """The plain FP64 single-instance path on every instance, one context each,
    one after another: steps 2..steps+1 of each timed."""
    import torch
    from jax_wspm_amd.device import Context
    tot = 0.0
    for i, inst in enumerate(insts):
        ctx = Context(inst.points, inst.elements)
        ctx.set_dirichlet(inst.bc_lists)
        ctx.richards_init("vg", inst.phi, "bicgstab", "jacobi", inst.h0, tol=tol)
        ctx.richards_step(DT)             # warm-up step (first Krylov chunk sizing)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ctx.richards_step(DT)
        torch.cuda.synchronize()
        tot += time.perf_counter() - t1
        ctx.close()
    return len(insts) * steps / tot


def _rel(a, b):
    import numpy as np
    return np.array([np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(a, b)])


def _err_summary(a, b, a_th=None, b_th=None):
    import numpy as np
    eh = _rel(a, b)
    out = {"max_rel_l2_h": float(eh.max()), "median_rel_l2_h": float(np.median(eh)),
           "instances_within_1e-8_h": int((eh <= 1e-8).sum())}
    if a_th is not None:
        et = _rel(a_th, b_th)
        out["max_rel_l2_theta"] = float(et.max())
        out["instances_within_1e-8_theta"] = int((et <= 1e-8).sum())
    return out


def _cpu_ensemble_baseline(insts, h_warm, node_off, args):
    """The oracle (1 thread) on instances 0-7 over the 3 steps that follow the
    GPU's warm-up, started from the GPU's state there."""
    import numpy as np
    from oracle import picard as opicard
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:  # pragma: no cover
        pass
    sample = list(range(8))
    its = kry = 0
    t1 = time.perf_counter()
    for i in sample:
        inst = insts[i]
        opb = opicard.Problem(points=inst.points, elems=inst.elements, soil_model="vg", phi=inst.phi,
                              bc_lists=inst.bc_lists, h0=inst.h0, dt0=DT, nsteps=args.cpu_steps, adaptive=False,
                              solver="bicgstab", precond="jacobi")
        stp = opicard.Stepper(opb)
        hh = np.asarray(h_warm[node_off[i]:node_off[i + 1]], dtype=np.float64).copy()
        for _ in range(args.cpu_steps):
            hh, _, it, _ = stp.step(hh, DT)
            its += it
        kry += int(np.sum(stp.krylov_iters))
    cel = time.perf_counter() - t1
    ns = len(sample) * args.cpu_steps
    return {"value": round(ns / cel, 4), "unit": "time steps/s", "cores": 1, "kind": "port",
            "sample": (f"instances 0-7 (sides {[insts[i].cells for i in sample]}), steps {args.warmup + 1}-"
                       f"{args.warmup + args.cpu_steps} started from the GPU's state after the warm-up (the first "
                       f"steps of the GPU's timed window), FP64 Krylov tol 1e-10; {cel:.1f} s, NumPy/SciPy FP64 "
                       f"oracle, 1 thread; host has {os.cpu_count()} logical CPUs"),
            "picard_iters_per_instance_step": round(its / ns, 3),
            "krylov_iters_per_instance_step": round(kry / ns, 3)}


def _big(form, bytes_full, n):
    """The SpMV roofline over the timed launches that worked on at least a
    quarter of the operator's rows (the launches that stream; the others are
    the short tails of instances still solving)."""
    lb, mb, rb = form["spmv_big"]
    if not lb:
        return None
    achieved = bytes_full * (rb / n) / (mb * 1e-3) / 1e9
    return {"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4), "launches": lb,
            "avg_launch_us": round(mb / lb * 1e3, 3), "rows_per_launch": round(rb / lb)}


def _inner_desc(args, inner=None):
    inner = inner or args.inner
    if inner == "jacobi":
        return {"method": "jacobi", "steps_per_sweep": args.inner_iters,
                "what": "Jacobi-preconditioned Richardson steps d <- d + M (r - A d) in FP32 storage from d = 0, "
                        "no dot products; sweeps until the FP64 residual meets the Krylov tolerance"}
    return {"method": "krylov", "inner_tol": MIXED_INNER_TOL, "max_iters_per_sweep": MIXED_INNER_ITERS,
            "what": "the method's FP32 Krylov solve per refinement sweep"}


def run_headline(args):
    """c3: batched mixed vs batched FP64 vs plain FP64 on the 256-instance ensemble."""
    import numpy as np
    from jax_wspm_amd.ensemble import vg_ensemble

    insts = vg_ensemble(256, seed=0)
    groups = _ens_groups(insts, args)
    log(f"c3: 256 instances, groups {[len(g) for g in groups] if groups else [256]}")
    mixed = _ens_form(insts, groups, "mixed", TIGHT_TOL, args)
    # the other inner method, for comparison (no timing region)
    other = "krylov" if args.inner == "jacobi" else "jacobi"
    mixed_o = _ens_form(insts, groups, "mixed", TIGHT_TOL, args, timing=False, inner=other)
    f64t = _ens_form(insts, groups, "fp64", TIGHT_TOL, args, timing=False)
    f64r = _ens_form(insts, groups, "fp64", 1e-10, args)
    plain_r = plain_t = None
    if args.plain_steps > 0:
        plain_r = _plain_rate(insts, 1e-10, args.plain_steps)
        log(f"plain FP64 tol 1e-10: {plain_r:.0f} instance-steps/s")
        plain_t = _plain_rate(insts, TIGHT_TOL, args.plain_steps)
        log(f"plain FP64 tol {TIGHT_TOL:g}: {plain_t:.0f} instance-steps/s")
    n, ne, nnz = mixed["n"], mixed["n_elems"], mixed["nnz"]
    value = mixed["rate"]
    node_off = np.concatenate([[0], np.cumsum([len(i.points) for i in insts])])
    cfg = {
        "workload": WORKLOADS["c3"][0], "instances": len(insts), "nodes": n, "elements": ne, "nnz": nnz, "dt": DT,
        "value_counts": "instance time steps of the batched mixed-precision form (each timed step advances every "
                        "instance once)",
        "mixed": {"krylov_tol": TIGHT_TOL, "inner": _inner_desc(args),
                  "instance_steps_per_s": round(mixed["rate"], 1),
                  "max_picard_iters_per_step": round(mixed["picard"], 3),
                  "fp32_inner_iters_per_step_summed": round(mixed["krylov"], 1),
                  "all_converged": mixed["converged"], "launches": mixed["launches"]},
        f"mixed_{other}_inner": {"inner": _inner_desc(args, other),
                                 "instance_steps_per_s": round(mixed_o["rate"], 1),
                                 "fp32_inner_iters_per_step_summed": round(mixed_o["krylov"], 1),
                                 "all_converged": mixed_o["converged"],
                                 "vs_fp64_same_tol": _err_summary(mixed_o["h"], f64t["h"], mixed_o["theta"],
                                                                  f64t["theta"])},
        "batched_fp64": {"krylov_tol_1e-10_instance_steps_per_s": round(f64r["rate"], 1),
                         f"krylov_tol_{TIGHT_TOL:g}_instance_steps_per_s": round(f64t["rate"], 1),
                         "krylov_iters_per_step_summed_tol_1e-10": round(f64r["krylov"], 1),
                         f"krylov_iters_per_step_summed_tol_{TIGHT_TOL:g}": round(f64t["krylov"], 1),
                         "launches_tol_1e-10": f64r["launches"]},
        "errors_after_steps": args.warmup + args.steps,
        "mixed_vs_fp64_same_tol": _err_summary(mixed["h"], f64t["h"], mixed["theta"], f64t["theta"]),
        "mixed_vs_fp64_tol_1e-10": _err_summary(mixed["h"], f64r["h"]),
        "fp64_tol_1e-10_vs_fp64_same_tol": _err_summary(f64r["h"], f64t["h"]),
        "speedup_batched_mixed_vs_batched_fp64_same_tol": round(mixed["rate"] / f64t["rate"], 3),
        "speedup_batched_fp64_vs_plain_fp64_tol_1e-10": round(f64r["rate"] / plain_r, 2) if plain_r else None,
        "spmv_format": mixed["format"], "launch_shape": mixed["shape"],
        "launch": launch_note(),
        "contexts": [len(g) for g in groups] if groups else [len(insts)],
        "grouping": (f"instances ranked by Picard iterations over {max(args.warmup, 1)} FP64 probe steps; the "
                     f"slowest {'/'.join(map(str, args.slow_group))} in ensemble contexts of their own, each on its "
                     "own stream stepped by its own host thread" if groups else "one context"),
    }
    if plain_r:
        cfg["plain_fp64_single_instance"] = {
            "krylov_tol_1e-10_instance_steps_per_s": round(plain_r, 1),
            f"krylov_tol_{TIGHT_TOL:g}_instance_steps_per_s": round(plain_t, 1),
            "sample": f"steps 2..{1 + args.plain_steps} of each of the 256 instances, one context per instance, "
                      "run one after another"}
        cfg["speedup_vs_plain_fp64"] = round(value / plain_r, 2)
        cfg["speedup_vs_plain_fp64_same_tol"] = round(value / plain_t, 2)
    out = {"metric": METRIC, "value": round(value, 2), "unit": "time steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(mixed["elapsed"] / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (structured meshes, seeded per-instance soil parameters generated in-process)",
           "config": cfg}
    if "asm" in mixed:
        la, ma, ra = mixed["asm"]
        out["roofline"] = roofline(_asm_kernel(mixed["windows"]) + " (fused P1 assembly + Dirichlet rows + Jacobi "
                                   "+ refinement residual, one thread per row)", assembly_bytes(n, ne, nnz, 2, False),
                                   n, la, ma, ra, "assembly_c3_mixed")
        ls, ms, rs = mixed["spmv"]
        epi = "Jacobi-step epilogues" if args.inner == "jacobi" else "BiCGSTAB epilogues"
        out["roofline_spmv"] = roofline(f"k_spmv_{mixed['format']} (FP32 SpMV of the inner solves, fused {epi})",
                                        spmv_bytes(n, nnz, 4), n, ls, ms, rs, "spmv_c3_fp32")
        out["roofline_spmv"]["launches_over_quarter_operator"] = _big(mixed, spmv_bytes(n, nnz, 4), n)
        ls, ms, rs = f64r["spmv"]
        out["roofline_spmv_fp64"] = roofline(f"k_spmv_{f64r['format']} (FP64 SpMV, batched FP64 form at tol 1e-10)",
                                             spmv_bytes(n, nnz, 8), n, ls, ms, rs, "spmv_c3_fp64")
        out["roofline_spmv_fp64"]["launches_over_quarter_operator"] = _big(f64r, spmv_bytes(n, nnz, 8), n)
        cfg["timing_region"] = (f"{args.steps} further steps of the mixed form with HIP events around every "
                                f"assembly and FP32 SpMV launch ({mixed['el_ev'] / args.steps * 1e3:.3f} ms/step "
                                "there, groups stepped one after another)")
    if args.cpu_steps > 0:
        out["cpu_baseline"] = _cpu_ensemble_baseline(insts, f64r["h_warm"], node_off, args)
        cfg["gpu_window_picard_max_per_step"] = round(f64r["picard"], 3)
        cfg["gpu_window_krylov_iters_per_instance_step_tol_1e-10"] = round(f64r["krylov"] / len(insts), 3)
    print(json.dumps(out))


def run_ensemble_form(args):
    """c3-fp64 / c3-mixed: one batched form alone (for profiling)."""
    from jax_wspm_amd.ensemble import vg_ensemble
    precision = "mixed" if args.workload == "c3-mixed" else "fp64"
    tol = TIGHT_TOL if precision == "mixed" else 1e-10
    insts = vg_ensemble(256, seed=0)
    groups = _ens_groups(insts, args)
    r = _ens_form(insts, groups, precision, tol, args)
    n, ne, nnz = r["n"], r["n_elems"], r["nnz"]
    out = {"metric": METRIC, "value": round(r["rate"], 2), "unit": "time steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if precision == "mixed" else "f64",
           "data": "synthetic (structured meshes, seeded per-instance soil parameters generated in-process)",
           "config": {"workload": WORKLOADS[args.workload][0], "krylov_tol": tol, "instances": len(insts),
                      "nodes": n, "elements": ne, "nnz": nnz, "krylov_iters_per_step_summed": round(r["krylov"], 1),
                      "max_picard_iters_per_step": round(r["picard"], 3), "launch_shape": r["shape"],
                      "launch": launch_note(), "launches": r["launches"]}}
    if "asm" in r:
        out["roofline"] = roofline(_asm_kernel(r["windows"]) + " (fused assembly)",
                                   assembly_bytes(n, ne, nnz, 2, False), n,
                                   *r["asm"], "assembly_c3_mixed" if precision == "mixed" else "assembly_c3")
        out["roofline_spmv"] = roofline(f"k_spmv_{r['format']}", spmv_bytes(n, nnz, 4 if precision == "mixed" else 8),
                                        n, *r["spmv"], None if precision == "mixed" else "spmv_c3_fp64")
    print(json.dumps(out))


# ---------------------------------------------------------------------------------------------
# one mesh: c1, c1-mixed, c2, c2-fp64
# ---------------------------------------------------------------------------------------------

def cpu_baseline(pb, h_start, steps):
    """Oracle restatement of the reference path, timed on this host (1 thread),
    from h_start."""
    import numpy as np
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:  # pragma: no cover
        pass
    from oracle import picard as opicard
    opb = opicard.Problem(points=pb.points, elems=pb.elements, soil_model=pb.soil_model, phi=pb.phi,
                          bc_lists=pb.bc_lists, h0=pb.h0, dt0=pb.dt0, nsteps=steps, adaptive=False,
                          picard_tol=pb.picard_tol, picard_maxit=pb.picard_maxit, picard_mode=pb.picard_mode,
                          solver=pb.solver, precond=pb.precond, tol=pb.tol, maxiter=pb.maxiter,
                          restart=pb.restart, Ks_elem=pb.Ks_elem)
    st = opicard.Stepper(opb)
    h = np.asarray(h_start, dtype=np.float64).copy()
    its = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        h, _, it, _ = st.step(h, pb.dt0)
        its += it
    el = time.perf_counter() - t0
    return steps / el, el, its / steps, float(np.sum(st.krylov_iters)) / steps


def run_single(args):
    import numpy as np
    import torch

    from jax_wspm_amd.problems import baseline_vg_config, from_config
    from jax_wspm_amd.richards import RichardsSolver

    desc, cells, dim, solver, precision, ksf, dt = WORKLOADS[args.workload]
    mixed = precision == "mixed"
    tol = TIGHT_TOL if mixed else 1e-10

    def make(prec, ktol):
        cfg = baseline_vg_config(cells, dim, solver=solver, precond="jacobi", precision=prec, ks_field=ksf,
                                 nsteps=args.warmup + args.steps, dt=dt)
        pb = from_config(cfg)
        pb.tol = ktol
        if prec == "mixed":
            pb.inner_tol = MIXED_INNER_TOL
            pb.inner = args.inner
            pb.inner_iters = args.inner_iters if args.inner == "jacobi" else 0
        rs = RichardsSolver(pb, history="last")
        rs.init_state()
        return pb, rs.ctx

    pb, ctx = make(precision, tol)
    log(f"{args.workload}: n={ctx.n} E={ctx.n_elems} nnz={ctx.nnz} lanes={ctx.lanes}")
    for _ in range(args.warmup):
        ctx.richards_step(dt)
    h_warm = ctx.richards_state()[0].cpu().numpy()
    torch.cuda.synchronize()
    picard = kry = 0
    lc0 = _launch_counts(ctx)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.richards_step(dt)
        picard += st["picard_iters"]
        kry += st["krylov_iters"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    launches = _launch_summary(lc0, _launch_counts(ctx), args.steps)
    h, th, _ = ctx.richards_state()
    h, th = h.cpu().numpy(), th.cpu().numpy()
    assert np.all(np.isfinite(h)), "non-finite pressure head"
    sp, asm, el_ev, _ = timed_region(ctx, ctx.richards_step, dt, 0 if args.no_timing else args.steps, mixed)
    n, ne, nnz = ctx.n, ctx.n_elems, ctx.nnz
    fmt = ctx.spmv_format
    asm_name = _asm_kernel(ctx.assembly_windows) if dim == 2 else "k_assemble"
    ctx.close()
    value = args.steps / el
    cfgd = {"workload": desc, "nodes": n, "elements": ne, "nnz": nnz, "dt": dt, "solver": f"{solver}+jacobi",
            "precision": precision, "krylov_tol": tol, "picard_iters_per_step": round(picard / args.steps, 3),
            "krylov_iters_per_step": round(kry / args.steps, 3), "launch": launch_note(), "launches": launches,
            "timing_region": (f"{args.steps} further steps of the same run with HIP events around every "
                              f"assembly and {'FP32 ' if mixed else ''}SpMV launch "
                              f"({el_ev / max(args.steps, 1) * 1e3:.4f} ms/step there)")}
    if mixed:
        # the FP64 form on the same problem, same steps, same tolerance
        _, c64 = make("fp64", tol)
        for _ in range(args.warmup):
            c64.richards_step(dt)
        torch.cuda.synchronize()
        t64 = time.perf_counter()
        for _ in range(args.steps):
            c64.richards_step(dt)
        torch.cuda.synchronize()
        rate64 = args.steps / (time.perf_counter() - t64)
        h64, th64, _ = (x.cpu().numpy() for x in c64.richards_state())
        c64.close()
        cfgd["fp64_same_tol_steps_per_s"] = round(rate64, 3)
        cfgd["speedup_vs_fp64_same_tol"] = round(value / rate64, 3)
        cfgd["inner"] = _inner_desc(args)
        cfgd["rel_l2_vs_fp64_h"] = float(np.linalg.norm(h - h64) / np.linalg.norm(h64))
        cfgd["rel_l2_vs_fp64_theta"] = float(np.linalg.norm(th - th64) / np.linalg.norm(th64))
        cfgd["error_sample"] = (f"after {args.warmup + args.steps} steps, against the FP64 form at the same Krylov "
                                "tolerance (itself parity-tested against the oracle)")
    out = {"metric": METRIC, "value": round(value, 3), "unit": "time steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if precision == "fp64" else "f32+f64",
           "data": "synthetic (structured mesh" + (" + seeded correlated Ks field" if ksf else "") +
                   " generated in-process)",
           "config": cfgd,
           "roofline": roofline(f"{asm_name} (fused P1 assembly + Dirichlet rows + Jacobi + solver start)", assembly_bytes(n, ne, nnz, dim, ksf is not None), n,
                                *asm, "assembly_c1" if args.workload == "c1" else None),
           "roofline_spmv": roofline(f"k_spmv_{fmt} ({'FP32' if mixed else 'FP64'} SpMV, fused Krylov epilogues)",
                                     spmv_bytes(n, nnz, 4 if mixed else 8), n, *sp,
                                     "spmv_c1" if args.workload == "c1" else None)}
    if args.cpu_steps > 0:
        pb_cpu = from_config(baseline_vg_config(cells, dim, solver=solver, precond="jacobi", ks_field=ksf,
                                                nsteps=args.cpu_steps, dt=dt))
        cps, cel, cit, ckr = cpu_baseline(pb_cpu, h_warm, args.cpu_steps)
        out["cpu_baseline"] = {"value": round(cps, 4), "unit": "time steps/s", "cores": 1, "kind": "port",
                               "sample": f"steps {args.warmup + 1}-{args.warmup + args.cpu_steps} of the same run, "
                                         f"from the GPU's state after the warm-up ({cel:.1f} s), NumPy/SciPy FP64 "
                                         f"oracle at Krylov tol 1e-10, 1 thread; host has {os.cpu_count()} logical "
                                         "CPUs",
                               "picard_iters_per_step": round(cit, 3), "krylov_iters_per_step": round(ckr, 3)}
    print(json.dumps(out))


# ---------------------------------------------------------------------------------------------
# c4: adjoint
# ---------------------------------------------------------------------------------------------

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
    log(f"c4: n={model.n} E={model.E} steps={steps} obs={len(obs)}x{len(obs_steps)}")
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
    log(f"c4 forward {t_fwd:.2f} s backward {t_bwd:.2f} s")
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
    eps = args.fd_eps if args.fd_eps else {2: 1e-4, 4: 1e-4, 6: 5e-4}[args.fd_order]
    for k in range(args.fd_dirs):
        d = drng.normal(size=p0.size) * scale
        # error budget of the difference quotient: truncation ~eps^p for the
        # order-p stencil plus the forward's converged-tolerance noise (about
        # 1e-11 in the loss) divided by eps.  Measured on these directions:
        # order 2 at eps 1e-4, up to 1.6e-6 truncation; order 4 at 1e-4,
        # noise of up to 1e-6 in the direction with the smallest derivative
        # (g.d 0.41); order 4 at 5e-4, 8.9e-7 truncation in the most
        # nonlinear direction.  Order 6 at 5e-4 keeps both terms below 1e-7.

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
        log(f"c4 direction {k + 1}/{args.fd_dirs}: g.d {g @ d:.6e} fd {fd:.6e} rel err {errs[-1]:.2e}")
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
    if args.workload == "c3":
        run_headline(args)
    elif args.workload in ("c3-fp64", "c3-mixed"):
        run_ensemble_form(args)
    elif args.workload == "c4":
        run_adjoint(args)
    else:
        run_single(args)


if __name__ == "__main__":
    main()
