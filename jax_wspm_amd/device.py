"""Device context over the C ABI: one mesh, its CSR pattern and its stream.

PyTorch is used only for device memory (tensors whose ``data_ptr`` is handed
to the C ABI) and for the current HIP stream.  Every computation runs in
libwspm.so; if the library or a GPU is missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("jax_wspm_amd needs a HIP GPU (torch.cuda.is_available() is False)")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _f64(t, device="cuda"):
    if isinstance(t, torch.Tensor):
        return t.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(t, dtype=np.float64), device=device)


class Context:
    """A mesh resident on the GPU (wraps ``wspm_ctx``)."""

    def __init__(self, points: np.ndarray, elements: np.ndarray):
        _require_gpu()
        self.lib = _lib.load()
        points = np.ascontiguousarray(points, dtype=np.float64)
        conn = np.ascontiguousarray(elements, dtype=np.int32)
        self.dim = points.shape[1]
        if conn.shape[1] != self.dim + 1:
            raise ValueError("elements must have dim+1 vertices")
        self.n, self.n_elems = points.shape[0], conn.shape[0]
        h = C.c_void_p()
        rc = self.lib.wspm_ctx_create(self.dim, self.n, self.n_elems, ptr(conn), ptr(points), _stream(),
                                      C.byref(h))
        check(rc, None)
        self.h = h
        info = (C.c_int64 * 8)()
        check(self.lib.wspm_ctx_info(self.h, info), self.h)
        self.nnz = int(info[2])
        self.max_row = int(info[4])
        self.lanes = int(info[6])
        self.device = torch.device("cuda")

    # -- lifetime -------------------------------------------------------------
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

    # -- pattern ----------------------------------------------------------------
    def pattern(self) -> Tuple[np.ndarray, np.ndarray]:
        rp = np.empty(self.n + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        self._ck(self.lib.wspm_get_pattern(self.h, ptr(rp), ptr(ci)))
        return rp, ci

    def set_dirichlet(self, bc_lists: Sequence[Tuple[np.ndarray, np.ndarray]]):
        """Ordered (nodes, values) lists; a later list wins (top then bottom)."""
        nodes, vals = [], []
        for nd, v in bc_lists:
            nd = np.asarray(nd, dtype=np.int32).ravel()
            nodes.append(nd)
            vals.append(np.broadcast_to(np.asarray(v, dtype=np.float64), nd.shape))
        nodes = np.ascontiguousarray(np.concatenate(nodes) if nodes else np.zeros(0, np.int32))
        vals = np.ascontiguousarray(np.concatenate(vals) if vals else np.zeros(0))
        self._ck(self.lib.wspm_bc_set(self.h, int(nodes.size), ptr(nodes), ptr(vals)))

    # -- kernels ------------------------------------------------------------------
    def soil(self, model: str, params, h: torch.Tensor):
        h = _f64(h)
        p = np.zeros(6)
        pa = np.asarray(params, dtype=np.float64)
        p[: pa.size] = pa
        Cc, K, th = (torch.empty_like(h) for _ in range(3))
        self._ck(self.lib.wspm_soil(self.h, _lib.SOIL[model], ptr(p), int(h.numel()), ptr(h), ptr(Cc), ptr(K),
                                    ptr(th)))
        return Cc, K, th

    def assemble_richards(self, h_m, C_m, K_m, theta_m, theta_n, dt, Ks_elem=None, want_dinv=True):
        args = [_f64(a) for a in (h_m, C_m, K_m, theta_m, theta_n)]
        ks = _f64(Ks_elem) if Ks_elem is not None else None
        A = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        b = torch.empty(self.n, dtype=torch.float64, device=self.device)
        dinv = torch.empty(self.n, dtype=torch.float64, device=self.device) if want_dinv else None
        self._ck(self.lib.wspm_assemble_richards(self.h, ptr(ks), *(ptr(a) for a in args), float(dt),
                                                  ptr(A), ptr(b), ptr(dinv)))
        return A, b, dinv

    def spmv(self, A: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        x = _f64(x)
        y = torch.empty_like(x)
        self._ck(self.lib.wspm_spmv(self.h, ptr(A), ptr(x), ptr(y)))
        return y

    def krylov_pc(self, method: str, precond: Optional[str], A, b, x0, tol=1e-10, atol=0.0, maxiter=100,
                  restart=30):
        """Krylov solve with any preconditioner of the reference built from A:
        none, jacobi, block_jacobi, ssor, ilu."""
        if method not in _lib.STEP_METHOD:
            raise ValueError(f"Unsupported solver type: {method}")
        if precond not in _lib.PRECOND:
            raise ValueError(f"Unsupported preconditioner: {precond}")
        b, x0 = _f64(b), _f64(x0)
        x = torch.empty_like(x0)
        st = _lib.KrylovStats()
        self._ck(self.lib.wspm_krylov_pc(self.h, _lib.METHOD[method], _lib.PRECOND[precond], ptr(A), ptr(b), ptr(x0),
                                         ptr(x), float(tol), float(atol), int(maxiter), int(restart), C.byref(st)))
        return x, {"iters": st.iters, "arnoldi": st.arnoldi, "code": st.code, "refinements": st.refinements,
                   "res_norm": st.res_norm, "b_norm": st.b_norm}

    def direct(self, A, b):
        """x = A^{-1} b by the banded LU with partial pivoting (wspm_direct)."""
        b = _f64(b)
        x = torch.empty_like(b)
        info = C.c_int32()
        self._ck(self.lib.wspm_direct(self.h, ptr(A), ptr(b), ptr(x), C.byref(info)))
        return x

    def direct_info(self):
        """(kl, ku, reordered) of the direct solver's band."""
        kl, ku, r = C.c_int32(), C.c_int32(), C.c_int32()
        self._ck(self.lib.wspm_direct_info(self.h, C.byref(kl), C.byref(ku), C.byref(r)))
        return kl.value, ku.value, bool(r.value)

    def krylov(self, method: str, A, b, x0, dinv=None, tol=1e-10, atol=0.0, maxiter=100, restart=30,
               precision="fp64", inner_tol=1e-6, max_refine=50):
        b, x0 = _f64(b), _f64(x0)
        x = torch.empty_like(x0)
        st = _lib.KrylovStats()
        if method not in _lib.METHOD:
            raise ValueError(f"Unsupported solver type: {method}")
        if precision == "mixed":
            rc = self.lib.wspm_krylov_mixed(self.h, _lib.METHOD[method], ptr(A), ptr(dinv), ptr(b), ptr(x0),
                                            ptr(x), float(tol), float(atol), int(maxiter), int(restart),
                                            float(inner_tol), int(max_refine), C.byref(st))
        else:
            rc = self.lib.wspm_krylov(self.h, _lib.METHOD[method], ptr(A), ptr(dinv), ptr(b), ptr(x0), ptr(x),
                                      float(tol), float(atol), int(maxiter), int(restart), C.byref(st))
        self._ck(rc)
        return x, {"iters": st.iters, "arnoldi": st.arnoldi, "code": st.code, "refinements": st.refinements,
                   "res_norm": st.res_norm, "b_norm": st.b_norm}

    # -- device-resident time stepping --------------------------------------------
    def richards_init(self, soil_model: str, phi, method: str, precond: Optional[str], h0, *,
                      Ks_elem=None, tol=1e-10, atol=0.0, maxiter=100, restart=30, picard_tol=1e-6,
                      picard_maxit=100, picard_mode="rel", precision="fp64", inner_tol=1e-6,
                      inner="krylov", inner_iters=0):
        """inner (mixed precision): "krylov", the method's FP32 solve to inner_tol per
        refinement sweep, or "jacobi", inner_iters Jacobi-preconditioned Richardson
        steps per sweep (0: 3; needs precond "jacobi")."""
        if method not in _lib.STEP_METHOD:
            raise ValueError(f"Unsupported solver type: {method}")
        if precond not in _lib.PRECOND:
            raise NotImplementedError(f"preconditioner {precond!r} is not available on the MI355X path "
                                      "(supported: none, jacobi, block_jacobi, ssor, ilu)")
        prm = _lib.RichardsParams()
        prm.soil_model = _lib.SOIL[soil_model]
        prm.method = _lib.STEP_METHOD[method]
        prm.precond = _lib.PRECOND[precond]
        prm.precision = _lib.PRECISION[precision]
        pa = np.asarray(phi, dtype=np.float64)
        for k in range(6):
            prm.phi[k] = float(pa[k]) if k < pa.size else 0.0
        prm.tol, prm.atol, prm.maxiter, prm.restart = float(tol), float(atol), int(maxiter), int(restart)
        prm.picard_tol, prm.picard_maxit = float(picard_tol), int(picard_maxit)
        prm.picard_mode = _lib.PICARD[picard_mode]
        prm.inner_tol = float(inner_tol)
        prm.inner_method = _lib.INNER[inner]
        prm.inner_iters = int(inner_iters)
        h0 = _f64(h0)
        ks = _f64(Ks_elem) if Ks_elem is not None else None
        self._ck(self.lib.wspm_richards_init(self.h, C.byref(prm), ptr(ks), ptr(h0)))

    def richards_step(self, dt: float):
        st = _lib.StepStats()
        self._ck(self.lib.wspm_richards_step(self.h, float(dt), C.byref(st)))
        return {"picard_iters": st.picard_iters, "krylov_iters": st.krylov_iters,
                "converged": bool(st.converged), "err": st.err}

    def richards_state(self):
        h = torch.empty(self.n, dtype=torch.float64, device=self.device)
        th = torch.empty_like(h)
        thn = torch.empty_like(h)
        self._ck(self.lib.wspm_richards_state(self.h, ptr(h), ptr(th), ptr(thn)))
        return h, th, thn

    # -- transport ---------------------------------------------------------------------
    def edge_set(self, which: str, nodes):
        """Boundary node list ("inlet" / "outlet") for edge terms, in list order."""
        nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).ravel())
        self._ck(self.lib.wspm_edge_set(self.h, _lib.EDGE[which], int(nd.size), ptr(nd)))

    def richards_edges(self, neumann_flux: Optional[float] = None, freedrain: bool = False):
        """Neumann flux on the inlet list and/or free drainage on the outlet list
        in every following Richards step."""
        self._ck(self.lib.wspm_richards_edges(self.h, 0 if neumann_flux is None else 1,
                                              float(neumann_flux or 0.0), 1 if freedrain else 0))

    def transport_fields(self, phi, h, DL, DT, Dm, Ks_elem=None):
        """(qx, qz, |q|, Dxx, Dxz, Dzz, theta) at the nodes (device tensors).
        Ks_elem: optional per-element Ks; K at a node is then the soil K times
        the mean Ks of the node's elements."""
        h = _f64(h)
        ks = _f64(Ks_elem) if Ks_elem is not None else None
        p = np.zeros(6)
        pa = np.asarray(phi, dtype=np.float64)
        p[: pa.size] = pa
        out = [torch.empty_like(h) for _ in range(7)]
        self._ck(self.lib.wspm_transport_fields(self.h, ptr(p), ptr(h), ptr(ks), float(DL), float(DT), float(Dm),
                                                *(ptr(o) for o in out)))
        return tuple(out)

    def domain_integral(self, f) -> float:
        """sum_e |det J_e| * mean_q f over the 3-point rule (twice the integral,
        as richards_mass_balance-checkpoint.py:25-65)."""
        f = _f64(f)
        out = C.c_double()
        self._ck(self.lib.wspm_domain_integral(self.h, ptr(f), C.byref(out)))
        return out.value

    def element_size(self) -> torch.Tensor:
        out = torch.empty(self.n_elems, dtype=torch.float64, device=self.device)
        self._ck(self.lib.wspm_element_size(self.h, ptr(out)))
        return out

    def peclet_courant(self, qx, qz, theta, dxx, dzz, dt, *, all_elements=False, per_element=False):
        """(Pe, Cr) of transport_utilities.py:206-242; see wspm_peclet_courant.
        per_element: also return the per-element arrays."""
        args = [_f64(v) for v in (qx, qz, theta, dxx, dzz)]
        pe = cr = None
        if per_element:
            pe = torch.empty(self.n_elems, dtype=torch.float64, device=self.device)
            cr = torch.empty_like(pe)
        out = (C.c_double * 2)()
        self._ck(self.lib.wspm_peclet_courant(self.h, *(ptr(a) for a in args), float(dt), 1 if all_elements else 0,
                                              ptr(pe), ptr(cr), out))
        return (out[0], out[1], pe, cr) if per_element else (out[0], out[1])

    def assemble_transport(self, fields, c_n, dt, *, lumped, mu=0.0, rho_kd=0.0, mu_src=0.0, c_src=None,
                           cauchy=None, K=None, want_dinv=True):
        """One species system on the context's pattern.  fields: the 7-tuple of
        transport_fields; cauchy: (flux, c_inlet) on the inlet list; K: enables
        free drainage on the outlet list."""
        qx, qz, _, dxx, dxz, dzz, th = (_f64(f) for f in fields)
        tp = _lib.TransportParams()
        tp.lumped = 1 if lumped else 0
        tp.inlet_cauchy = 1 if cauchy is not None else 0
        tp.outlet_freedrain = 1 if K is not None else 0
        tp.dt, tp.mu, tp.rho_kd, tp.mu_src = float(dt), float(mu), float(rho_kd), float(mu_src)
        if cauchy is not None:
            tp.cauchy_flux, tp.c_inlet = float(cauchy[0]), float(cauchy[1])
        c_n = _f64(c_n)
        cs = _f64(c_src) if c_src is not None else None
        Kd = _f64(K) if K is not None else None
        A = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        b = torch.empty(self.n, dtype=torch.float64, device=self.device)
        dinv = torch.empty(self.n, dtype=torch.float64, device=self.device) if want_dinv else None
        self._ck(self.lib.wspm_assemble_transport(self.h, C.byref(tp), ptr(th), ptr(qx), ptr(qz), ptr(dxx),
                                                  ptr(dxz), ptr(dzz), ptr(Kd), ptr(c_n), ptr(cs), ptr(A), ptr(b),
                                                  ptr(dinv)))
        return A, b, dinv

    # -- timing hooks ---------------------------------------------------------------
    @property
    def operator_order(self) -> str:
        """Order of each row's columns in the time-step operator: "fan" (the
        node, then its neighbours around it) or "sorted"."""
        info = (C.c_int64 * 8)()
        check(self.lib.wspm_ctx_info(self.h, info), self.h)
        return "fan" if info[7] >= 2 else "sorted"

    @property
    def assembly_windows(self) -> bool:
        """True if the time-step assembly stages each row block's node window
        in LDS (fan order, every window within the kernel's capacity)."""
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
