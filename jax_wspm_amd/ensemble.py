"""Ensembles of independent Richards instances on one GPU (``wspm_ens_*``).

The reference solves one instance per process (src/cli.py:87-150, one
``RichardsSolver`` per run).  An :class:`Ensemble` time-steps many instances
of the same implicit step (src/solvers/richards_solver_v2.py:107-168) at once:
every instance has its own mesh, van Genuchten / Gardner parameters, Dirichlet
rows and initial state; each one's Krylov solve and Picard loop stop on their
own, and each one's result equals that of the single-instance path on the same
inputs (its dot products cover its own rows only).

:func:`vg_ensemble` builds the uneven-size van Genuchten ensemble: side N of
each unit-square instance drawn from {32, 48, ..., 256}, alpha ~ U(0.01, 0.2),
n ~ U(1.3, 3.0) (m = 1 - 1/n), Ks ~ logU(1e-4, 1e-2), theta_r ~ U(0.02, 0.1),
theta_s ~ U(0.35, 0.5); h0 = -1000, h = -75 on the top edge.

``richards_init(precision="mixed")`` solves each instance's systems with FP32
Krylov corrections of the FP64 residual (refined per instance until that
residual meets the FP64 tolerance).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import mesh as M
from ._lib import check, ptr
from .device import _f64, _require_gpu, _stream


@dataclass
class Instance:
    """One member: mesh, soil parameters (phi layout of the C ABI), BCs, h0."""
    points: np.ndarray
    elements: np.ndarray
    phi: Tuple[float, ...]
    bc_lists: List[Tuple[np.ndarray, np.ndarray]]
    h0: np.ndarray
    Ks_elem: Optional[np.ndarray] = None
    cells: int = 0


def vg_ensemble(n_inst: int = 256, sides: Sequence[int] = tuple(range(32, 257, 16)), seed: int = 0,
                h_init: float = -1000.0, h_top: float = -75.0) -> List[Instance]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n_inst):
        N = int(rng.choice(np.asarray(sides)))
        alpha = rng.uniform(0.01, 0.2)
        n = rng.uniform(1.3, 3.0)
        ks = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        tr = rng.uniform(0.02, 0.1)
        ts = rng.uniform(0.35, 0.5)
        pts, el = M.unit_square(N, 1.0)
        top = np.nonzero(pts[:, 1] == 1.0)[0]
        out.append(Instance(points=pts, elements=el, phi=(alpha, ts, tr, n, 1.0 - 1.0 / n, ks),
                            bc_lists=[(top, np.full(top.shape, h_top))], h0=np.full(len(pts), h_init),
                            cells=N))
    return out


class Ensemble:
    """Instances resident on the GPU in one ensemble context."""

    def __init__(self, instances: Sequence[Instance]):
        _require_gpu()
        self.lib = _lib.load()
        self.instances = list(instances)
        if not self.instances:
            raise ValueError("empty ensemble")
        dims = {inst.points.shape[1] for inst in self.instances}
        if dims != {2}:
            raise ValueError("ensembles are 2D")
        self.n_inst = len(self.instances)
        nc = np.array([inst.points.shape[0] for inst in self.instances], dtype=np.int32)
        ec = np.array([inst.elements.shape[0] for inst in self.instances], dtype=np.int32)
        self.node_off = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
        self.elem_off = np.concatenate([[0], np.cumsum(ec)]).astype(np.int64)
        conn = np.ascontiguousarray(np.concatenate([inst.elements for inst in self.instances]), dtype=np.int32)
        coords = np.ascontiguousarray(np.concatenate([inst.points for inst in self.instances]), dtype=np.float64)
        h = C.c_void_p()
        check(self.lib.wspm_ens_create(2, self.n_inst, ptr(nc), ptr(ec), ptr(conn), ptr(coords), _stream(),
                                       C.byref(h)), None)
        self.h = h
        info = (C.c_int64 * 8)()
        check(self.lib.wspm_ctx_info(self.h, info), self.h)
        self.n, self.n_elems, self.nnz = int(info[0]), int(info[1]), int(info[2])
        self.lanes = int(info[6])
        self.device = torch.device("cuda")
        nodes, vals = [], []
        for i, inst in enumerate(self.instances):
            for nd, v in inst.bc_lists:
                nd = np.asarray(nd, dtype=np.int64).ravel()
                nodes.append((nd + self.node_off[i]).astype(np.int32))
                vals.append(np.broadcast_to(np.asarray(v, dtype=np.float64), nd.shape))
        nodes = np.ascontiguousarray(np.concatenate(nodes) if nodes else np.zeros(0, np.int32))
        vals = np.ascontiguousarray(np.concatenate(vals) if vals else np.zeros(0))
        check(self.lib.wspm_bc_set(self.h, int(nodes.size), ptr(nodes), ptr(vals)), self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.wspm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        check(rc, self.h)

    def richards_init(self, soil_model: str = "vg", method: str = "bicgstab", precond: Optional[str] = "jacobi", *,
                      tol=1e-10, atol=0.0, maxiter=100, picard_tol=1e-6, picard_maxit=100, picard_mode="rel",
                      precision: str = "fp64", inner_tol: float = 1e-6, inner: str = "krylov",
                      inner_iters: int = 0):
        """precision "mixed": FP32 corrections of the FP64 residual, refined per
        instance until that residual meets ``tol``; inner "krylov": the
        method's FP32 solve to inner_tol per sweep, "jacobi": inner_iters
        Jacobi-preconditioned Richardson steps per sweep (0: 3)."""
        if method not in ("cg", "bicgstab"):
            raise ValueError(f"ensembles solve with cg or bicgstab, not {method!r}")
        if precond not in _lib.PRECOND:
            raise NotImplementedError(f"preconditioner {precond!r} is not available (supported: none, jacobi)")
        prm = _lib.RichardsParams()
        prm.soil_model = _lib.SOIL[soil_model]
        prm.method = _lib.METHOD[method]
        prm.precond = _lib.PRECOND[precond]
        prm.precision = _lib.PRECISION[precision]
        prm.inner_tol = float(inner_tol)
        prm.inner_method = _lib.INNER[inner]
        prm.inner_iters = int(inner_iters)
        prm.tol, prm.atol, prm.maxiter, prm.restart = float(tol), float(atol), int(maxiter), 30
        prm.picard_tol, prm.picard_maxit = float(picard_tol), int(picard_maxit)
        prm.picard_mode = _lib.PICARD[picard_mode]
        phi = np.zeros((self.n_inst, 6))
        for i, inst in enumerate(self.instances):
            pa = np.asarray(inst.phi, dtype=np.float64)
            phi[i, : pa.size] = pa
        phi = np.ascontiguousarray(phi)
        h0 = _f64(np.concatenate([np.asarray(inst.h0, dtype=np.float64) for inst in self.instances]))
        has_ks = [inst.Ks_elem is not None for inst in self.instances]
        if any(has_ks) and not all(has_ks):
            raise ValueError("Ks_elem must be given for every instance or for none")
        ks = _f64(np.concatenate([inst.Ks_elem for inst in self.instances])) if all(has_ks) else None
        self._ck(self.lib.wspm_ens_richards_init(self.h, C.byref(prm), ptr(phi), ptr(ks), ptr(h0)))

    def step(self, dt: float, per_instance: bool = False):
        st = _lib.StepStats()
        self._ck(self.lib.wspm_richards_step(self.h, float(dt), C.byref(st)))
        out = {"picard_iters": st.picard_iters, "krylov_iters": st.krylov_iters, "converged": bool(st.converged),
               "err": st.err}
        if per_instance:
            out.update(self.step_stats())
        return out

    def step_stats(self):
        ip = np.zeros(self.n_inst, dtype=np.int32)
        ik = np.zeros(self.n_inst, dtype=np.int32)
        ic = np.zeros(self.n_inst, dtype=np.int32)
        er = np.zeros(self.n_inst)
        self._ck(self.lib.wspm_ens_step_stats(self.h, ptr(ip), ptr(ik), ptr(ic), ptr(er)))
        return {"picard_iters_i": ip, "krylov_iters_i": ik, "converged_i": ic.astype(bool), "err_i": er}

    def state(self):
        """Concatenated (h, theta, theta at the start of the last step) on the GPU."""
        h = torch.empty(self.n, dtype=torch.float64, device=self.device)
        th = torch.empty_like(h)
        thn = torch.empty_like(h)
        self._ck(self.lib.wspm_richards_state(self.h, ptr(h), ptr(th), ptr(thn)))
        return h, th, thn

    def split(self, x) -> List[np.ndarray]:
        """Per-instance pieces of a concatenated nodal array."""
        x = x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
        return [x[self.node_off[i]:self.node_off[i + 1]] for i in range(self.n_inst)]

    @property
    def assembly_windows(self) -> bool:
        """True if the time-step assembly stages each entry's node window in LDS."""
        info = (C.c_int64 * 8)()
        check(self.lib.wspm_ctx_info(self.h, info), self.h)
        return info[7] == 3

    @property
    def spmv_format(self) -> str:
        """Operator layout of the time-step solves since the last Richards
        init: "ell" or "csr"."""
        info = (C.c_int64 * 8)()
        check(self.lib.wspm_ctx_info(self.h, info), self.h)
        return "ell" if info[7] else "csr"

    @property
    def shape(self) -> dict:
        """Launch shape (wspm_ens_shape): operator format and width, the
        tuned number of workgroups walking the entry lists, rows per entry,
        entries, grids timed by the autotuner and the chosen grid's SpMV time
        over the whole operator (us), and whether stalled Jacobi inner steps
        have handed over to Krylov inner solves."""
        out = (C.c_int32 * 8)()
        check(self.lib.wspm_ens_shape(self.h, out), self.h)
        return {"format": "ell" if out[0] else "csr", "width": int(out[1]), "grid": int(out[2]),
                "rows_per_entry": int(out[3]), "entries": int(out[4]), "grids_timed": int(out[5]),
                "tuned_spmv_us": out[6] / 1000.0, "jacobi_switched_off": bool(out[7])}

    def timing(self, on: bool, fp32: bool = False, assembly: bool = False):
        """Time the FP64 SpMV launches (or, fp32=True, the FP32 ones of the
        mixed-precision inner solves); assembly=True: the assembly launches
        too (timing_read_asm)."""
        self._ck(self.lib.wspm_timing_enable(self.h, ((2 if fp32 else 1) | (4 if assembly else 0)) if on else 0))

    def timing_read_asm(self):
        """(working assembly launches, their summed duration in ms, rows they assembled)."""
        n = C.c_int64()
        ms = C.c_double()
        rows = C.c_int64()
        self._ck(self.lib.wspm_timing_read_asm(self.h, C.byref(n), C.byref(ms), C.byref(rows)))
        return int(n.value), float(ms.value), int(rows.value)

    def launch_count(self):
        """(kernel launches enqueued since creation, Picard rounds run)."""
        n = C.c_int64()
        r = C.c_int64()
        self._ck(self.lib.wspm_launch_count(self.h, C.byref(n), C.byref(r)))
        return int(n.value), int(r.value)

    def timing_read_over(self, min_rows: int):
        """timing_read's figures over the launches of at least min_rows rows,
        without resetting the timing slots (read these first)."""
        n = C.c_int64()
        ms = C.c_double()
        rows = C.c_int64()
        self._ck(self.lib.wspm_timing_read_over(self.h, int(min_rows), C.byref(n), C.byref(ms), C.byref(rows)))
        return int(n.value), float(ms.value), int(rows.value)

    def timing_read(self):
        """(working SpMV launches, their summed duration in ms, rows they worked on)."""
        n = C.c_int64()
        ms = C.c_double()
        rows = C.c_int64()
        self._ck(self.lib.wspm_timing_read(self.h, C.byref(n), C.byref(ms), C.byref(rows)))
        return int(n.value), float(ms.value), int(rows.value)


class EnsembleGroups:
    """Instances split over several ensemble contexts, each on its own HIP
    stream and stepped by its own host thread (the library's calls release the
    GIL), so that the kernels of different groups run concurrently.

    Instances are independent, and every context derives its SpMV layout from
    the instances' row lengths (its autotuner only times the launch grid,
    which changes no bit), so every instance's results are bitwise those of
    any other grouping.  The point of a split is the shape of a time step:
    the few instances whose Picard loops take the most iterations (stiff, dry
    soils) keep a step going for many short, latency-bound rounds after the
    others have converged.  In a group of their own (``split_slowest``) those
    rounds overlap the other instances' work instead of following it.
    """

    def __init__(self, instances: Sequence[Instance], groups: Sequence[Sequence[int]]):
        from concurrent.futures import ThreadPoolExecutor
        self.instances = list(instances)
        self.groups = [[int(i) for i in g] for g in groups if len(g)]
        seen = sorted(i for g in self.groups for i in g)
        if seen != list(range(len(self.instances))):
            raise ValueError("groups must hold every instance exactly once")
        self.n_inst = len(self.instances)
        self.streams = [torch.cuda.Stream() for _ in self.groups]
        self.members: List[Ensemble] = []
        for g, s in zip(self.groups, self.streams):
            with torch.cuda.stream(s):
                self.members.append(Ensemble([self.instances[i] for i in g]))
        self.n = sum(m.n for m in self.members)
        self.n_elems = sum(m.n_elems for m in self.members)
        self.nnz = sum(m.nnz for m in self.members)
        self.lanes = self.members[0].lanes
        self.device = torch.device("cuda")
        self._pool = ThreadPoolExecutor(max_workers=len(self.members))

    @staticmethod
    def split_slowest(picard_iters, n_slow) -> List[List[int]]:
        """Groups by Picard iterations (e.g. summed over a few probe steps):
        n_slow an int, the n_slow slowest instances and the rest; a sequence
        (a, b, ...), the a slowest, the next b, ..., and the rest.  The rest
        comes first."""
        sizes = [int(n_slow)] if np.isscalar(n_slow) else [int(k) for k in n_slow]
        order = [int(i) for i in np.argsort(-np.asarray(picard_iters), kind="stable")]
        tiers, pos = [], 0
        for k in sizes:
            tiers.append(sorted(order[pos:pos + k]))
            pos += k
        rest = sorted(order[pos:])
        return [g for g in [rest] + tiers[::-1] if g]

    def _each(self, fn):
        def call(k):
            with torch.cuda.stream(self.streams[k]):
                return fn(self.members[k])
        futures = [self._pool.submit(call, k) for k in range(len(self.members))]
        return [f.result() for f in futures]

    def richards_init(self, *args, **kw):
        self._each(lambda m: m.richards_init(*args, **kw))

    def step(self, dt: float, per_instance: bool = False, concurrent: bool = True):
        """One time step of every instance; concurrent False steps the groups
        one after another (same results; kernels of different groups then do
        not overlap, e.g. for per-launch timing)."""
        if concurrent:
            res = self._each(lambda m: m.step(dt, per_instance))
        else:
            res = []
            for m, s in zip(self.members, self.streams):
                with torch.cuda.stream(s):
                    res.append(m.step(dt, per_instance))
        errs = [r["err"] for r in res]
        out = {"picard_iters": max(r["picard_iters"] for r in res),
               "krylov_iters": sum(r["krylov_iters"] for r in res),
               "converged": all(r["converged"] for r in res),
               "err": float("nan") if any(np.isnan(errs)) else max(errs)}
        if per_instance:
            for key in ("picard_iters_i", "krylov_iters_i", "converged_i", "err_i"):
                arr = np.zeros(self.n_inst, dtype=res[0][key].dtype)
                for g, r in zip(self.groups, res):
                    arr[g] = r[key]
                out[key] = arr
        return out

    def state(self):
        """(h, theta, theta_n) concatenated in the original instance order."""
        parts = self._each(lambda m: [m.split(x) for x in m.state()])
        out = []
        for q in range(3):
            pieces = [None] * self.n_inst
            for g, p in zip(self.groups, parts):
                for j, i in enumerate(g):
                    pieces[i] = p[q][j]
            out.append(torch.as_tensor(np.concatenate(pieces), device=self.device))
        return tuple(out)

    def split(self, x) -> List[np.ndarray]:
        x = x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
        off = np.concatenate([[0], np.cumsum([inst.points.shape[0] for inst in self.instances])])
        return [x[off[i]:off[i + 1]] for i in range(self.n_inst)]

    @property
    def spmv_format(self) -> str:
        fmts = {m.spmv_format for m in self.members}
        return fmts.pop() if len(fmts) == 1 else "+".join(sorted(fmts))

    def timing(self, on: bool, fp32: bool = False, assembly: bool = False):
        self._each(lambda m: m.timing(on, fp32, assembly))

    def timing_read_over(self, min_rows: int):
        res = self._each(lambda m: m.timing_read_over(min_rows))
        return sum(r[0] for r in res), sum(r[1] for r in res), sum(r[2] for r in res)

    def timing_read(self):
        res = self._each(lambda m: m.timing_read())
        return sum(r[0] for r in res), sum(r[1] for r in res), sum(r[2] for r in res)

    def timing_read_asm(self):
        res = self._each(lambda m: m.timing_read_asm())
        return sum(r[0] for r in res), sum(r[1] for r in res), sum(r[2] for r in res)

    def launch_count(self):
        """launch_count of every member context, in member order."""
        return self._each(lambda m: m.launch_count())

    @property
    def shape(self) -> dict:
        """Launch shape of every member context (wspm_ens_shape)."""
        return {"contexts": [m.shape for m in self.members]}

    def close(self):
        for m in self.members:
            m.close()
        self._pool.shutdown(wait=True)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
