"""Gradients of Richards-flow observations by the hand-written discrete adjoint.

The reference has no gradient code (its solvers are JAX but never
differentiated); this module adds d(observations)/d(parameters) for the
implicit 2D van Genuchten step of src/solvers/richards_solver_v2.py:107-168:

* parameters: per-element log Ks, and the global van Genuchten alpha and n
  (m = 1 - 1/n; theta_s, theta_r fixed);
* observations: y[j] = theta(h_{k_j})[obs] at the requested steps k_j;
* forward: ``wspm_richards_step`` at fixed dt, h stored every ``checkpoint``
  steps;
* backward: per checkpoint segment (last first) the steps are recomputed from
  the stored h (the forward is deterministic, so the recomputed states equal
  the forward's bit for bit) and ``wspm_adj_step`` runs backward through them:
  a Jacobi-BiCGSTAB solve with the transposed step Jacobian and the parameter
  gradient kernels.

``RichardsObservations.apply(log_ks, alpha, n, model)`` is the
torch.autograd.Function.  Memory: one n-vector per checkpoint plus
``checkpoint + 1`` during the backward sweep.

The adjoint differentiates the converged step (F_k = 0).  A forward stopped
at Picard tolerance tau differs from that model by O(tau), so the gradient
matches finite differences of the forward to about tau / (FD step); run the
forward with a tight ``picard_tol`` when comparing.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import ptr
from .device import Context, _f64


class RichardsAdjointModel:
    def __init__(self, points, elements, bc_lists, h0, dt: float, nsteps: int, obs_nodes: Sequence[int],
                 obs_steps: Sequence[int], theta_s: float, theta_r: float, *, checkpoint: int = 20,
                 solver: str = "bicgstab", precond: Optional[str] = "jacobi", tol: float = 1e-10,
                 maxiter: int = 100, picard_tol: float = 1e-6, picard_maxit: int = 100, adj_tol: float = 1e-12,
                 adj_maxiter: int = 1000):
        if np.asarray(points).shape[1] != 2:
            raise ValueError("the adjoint is 2D")
        self.ctx = Context(points, elements)
        self.ctx.set_dirichlet(bc_lists)
        self.n, self.E = self.ctx.n, self.ctx.n_elems
        self.dev = self.ctx.device
        self.h0 = _f64(h0)
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.obs = torch.as_tensor(np.asarray(obs_nodes, dtype=np.int32), device=self.dev)
        self.obs_steps = [int(k) for k in obs_steps]
        if any(k < 1 or k > self.nsteps for k in self.obs_steps) or len(set(self.obs_steps)) != len(self.obs_steps):
            raise ValueError("obs_steps must be distinct steps in [1, nsteps]")
        self.ts, self.tr = float(theta_s), float(theta_r)
        self.checkpoint = max(1, int(checkpoint))
        self.kw = dict(tol=tol, maxiter=maxiter, picard_tol=picard_tol, picard_maxit=picard_maxit, picard_mode="rel")
        self.solver, self.precond = solver, precond
        self.adj_tol, self.adj_maxiter = float(adj_tol), int(adj_maxiter)
        self._saved: Dict[str, object] = {}

    def phi(self, alpha: float, n: float) -> np.ndarray:
        return np.array([alpha, self.ts, self.tr, n, 1.0 - 1.0 / n, 1.0])

    def _init(self, h, ks, alpha, n):
        self.ctx.richards_init("vg", self.phi(alpha, n), self.solver, self.precond, h, Ks_elem=ks, **self.kw)

    def forward(self, log_ks: torch.Tensor, alpha: float, n: float) -> torch.Tensor:
        ks = torch.exp(log_ks.to(device=self.dev, dtype=torch.float64)).contiguous()
        self._init(self.h0, ks, alpha, n)
        ckpt = {0: self.h0.clone()}
        ys = {}
        for k in range(1, self.nsteps + 1):
            self.ctx.richards_step(self.dt)
            if k in self.obs_steps or k % self.checkpoint == 0:
                h, th, _ = self.ctx.richards_state()
                if k in self.obs_steps:
                    ys[k] = th.index_select(0, self.obs.long())
                if k % self.checkpoint == 0 and k < self.nsteps:
                    ckpt[k] = h
        self._saved = dict(ks=ks, alpha=float(alpha), n=float(n), ckpt=ckpt)
        return torch.stack([ys[k] for k in self.obs_steps]) if self.obs_steps else torch.zeros(0, device=self.dev)

    def backward(self, dldy: torch.Tensor):
        S = self._saved
        ks, alpha, n = S["ks"], S["alpha"], S["n"]
        dldy = dldy.to(device=self.dev, dtype=torch.float64).contiguous()
        row = {k: j for j, k in enumerate(self.obs_steps)}
        g_ks = torch.zeros(self.E, dtype=torch.float64, device=self.dev)
        g_an = torch.zeros(2, dtype=torch.float64, device=self.dev)
        lam = [torch.zeros(self.n, dtype=torch.float64, device=self.dev) for _ in range(2)]
        lam_next = None
        prm = _lib.AdjParams()
        prm.dt = self.dt
        for k2, v in enumerate(self.phi(alpha, n)):
            prm.phi[k2] = float(v)
        prm.tol, prm.maxiter = self.adj_tol, self.adj_maxiter
        lib = self.ctx.lib
        stats = _lib.KrylovStats()
        self.adj_iters = []
        starts = sorted(S["ckpt"])
        for s in reversed(starts):
            e = min(s + self.checkpoint, self.nsteps)
            self._init(S["ckpt"][s], ks, alpha, n)
            hs = [S["ckpt"][s]]
            for _ in range(s + 1, e + 1):
                self.ctx.richards_step(self.dt)
                hs.append(self.ctx.richards_state()[0])
            for k in range(e, s, -1):
                out = lam[k % 2]
                j = row.get(k)
                prm.n_obs = int(self.obs.numel()) if j is not None else 0
                dy = dldy[j].contiguous() if j is not None else None
                rc = lib.wspm_adj_step(self.ctx.h, C.byref(prm), ptr(ks), ptr(hs[k - s]), ptr(hs[k - s - 1]),
                                       ptr(self.obs) if j is not None else None, ptr(dy), ptr(lam_next), None,
                                       ptr(out), ptr(g_ks), ptr(g_an), C.byref(stats))
                self.ctx._ck(rc)
                self.adj_iters.append(int(stats.iters))
                lam_next = out
        return g_ks * 1.0, g_an[0].clone(), g_an[1].clone()

    def close(self):
        self.ctx.close()


class NitrogenAdjointModel(RichardsAdjointModel):
    """Richards flow followed by the nitrogen species chain NH4 -> NO2 -> NO3
    every step (src/solvers/nitrogen_solver.py:148-238 at fixed dt: lumped
    mass, first-order rates ``mu``, sorption ``rho_kd`` of NH4, Cauchy inlet
    with ``cauchy_flux`` and ``c_inlet`` on ``inlet_nodes``), with the flux
    and dispersion fields of each step's h and the per-element Ks.

    Observations at ``obs_steps``: theta and NO3 at ``obs_nodes``; forward
    returns y [2, n_obs_steps, n_obs] (theta, NO3).  The backward sweep runs,
    per step and last step first, the species adjoint (``wspm_adj_transport_step``:
    NO3, NO2, NH4 transposed solves, then the fields' sensitivity carried to
    h_k, alpha, n and log Ks) and the Richards adjoint (``wspm_adj_step``)
    with that dL/dh_k added.  Memory: h and the three species per checkpoint,
    ``checkpoint + 1`` of each during the sweep."""

    def __init__(self, points, elements, bc_lists, h0, dt, nsteps, obs_nodes, obs_steps, theta_s, theta_r, *,
                 inlet_nodes, DL=0.5, DT=0.1, Dm=0.0, mu=(0.2, 0.5, 0.05), rho_kd=0.0, cauchy_flux=-0.05,
                 c_inlet=(1.0, 0.0, 0.4), c_init=(0.0, 0.0, 0.0), tr_tol=1e-12, tr_maxiter=400, **kw):
        super().__init__(points, elements, bc_lists, h0, dt, nsteps, obs_nodes, obs_steps, theta_s, theta_r, **kw)
        self.inlet = np.asarray(inlet_nodes, dtype=np.int32)
        self.ctx.edge_set("inlet", self.inlet)
        self.DL, self.DT, self.Dm = float(DL), float(DT), float(Dm)
        self.mu = tuple(float(v) for v in mu)
        self.rho_kd = float(rho_kd)
        self.cauchy_flux = float(cauchy_flux)
        self.c_inlet = tuple(float(v) for v in c_inlet)
        self.c0 = torch.cat([torch.full((self.n,), float(v), dtype=torch.float64, device=self.dev) for v in c_init])
        self.tr_tol, self.tr_maxiter = float(tr_tol), int(tr_maxiter)

    def _species_step(self, h, c_prev, ks, alpha, n):
        """The three species solves of one step; c_prev, result: [3n]."""
        N = self.n
        fields = self.ctx.transport_fields(self.phi(alpha, n), h, self.DL, self.DT, self.Dm, Ks_elem=ks)
        new = []
        for s in range(3):
            A, b, dinv = self.ctx.assemble_transport(
                fields, c_prev[s * N:(s + 1) * N], self.dt, lumped=True, mu=self.mu[s],
                rho_kd=self.rho_kd if s == 0 else 0.0, mu_src=self.mu[s - 1] if s > 0 else 0.0,
                c_src=new[-1] if s > 0 else None,
                cauchy=(self.cauchy_flux, self.c_inlet[s]) if self.cauchy_flux != 0.0 else None)
            x, _ = self.ctx.krylov("bicgstab", A, b, c_prev[s * N:(s + 1) * N], dinv=dinv, tol=self.tr_tol,
                                   maxiter=self.tr_maxiter)
            new.append(x)
        return torch.cat(new), fields[6]

    def forward(self, log_ks: torch.Tensor, alpha: float, n: float) -> torch.Tensor:
        ks = torch.exp(log_ks.to(device=self.dev, dtype=torch.float64)).contiguous()
        self._init(self.h0, ks, alpha, n)
        ckpt = {0: (self.h0.clone(), self.c0.clone())}
        c = self.c0
        ys = {}
        N = self.n
        obs = self.obs.long()
        for k in range(1, self.nsteps + 1):
            self.ctx.richards_step(self.dt)
            h = self.ctx.richards_state()[0]
            c, th = self._species_step(h, c, ks, alpha, n)
            if k in self.obs_steps:
                ys[k] = torch.stack([th.index_select(0, obs), c[2 * N:].index_select(0, obs)])
            if k % self.checkpoint == 0 and k < self.nsteps:
                ckpt[k] = (h, c)
        self._saved = dict(ks=ks, alpha=float(alpha), n=float(n), ckpt=ckpt)
        if not self.obs_steps:
            return torch.zeros(2, 0, self.obs.numel(), dtype=torch.float64, device=self.dev)
        return torch.stack([ys[k] for k in self.obs_steps], dim=1)

    def backward(self, dldy: torch.Tensor):
        """dldy [2, n_obs_steps, n_obs]: dL/d(theta obs), dL/d(NO3 obs)."""
        S = self._saved
        ks, alpha, n = S["ks"], S["alpha"], S["n"]
        dldy = dldy.to(device=self.dev, dtype=torch.float64).contiguous()
        row = {k: j for j, k in enumerate(self.obs_steps)}
        g_ks = torch.zeros(self.E, dtype=torch.float64, device=self.dev)
        g_an = torch.zeros(2, dtype=torch.float64, device=self.dev)
        lam = [torch.zeros(self.n, dtype=torch.float64, device=self.dev) for _ in range(2)]
        mu = [torch.zeros(3 * self.n, dtype=torch.float64, device=self.dev) for _ in range(2)]
        dldh = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        lam_next = mu_next = h_next = None
        prm = _lib.AdjParams()
        tp = _lib.AdjTrParams()
        prm.dt = tp.dt = self.dt
        for k2, v in enumerate(self.phi(alpha, n)):
            prm.phi[k2] = tp.phi[k2] = float(v)
        prm.tol, prm.maxiter = self.adj_tol, self.adj_maxiter
        tp.DL, tp.DT, tp.Dm = self.DL, self.DT, self.Dm
        for k2 in range(3):
            tp.mu[k2] = self.mu[k2]
        tp.rho_kd, tp.cauchy_flux = self.rho_kd, self.cauchy_flux
        tp.tol, tp.maxiter = self.adj_tol, self.adj_maxiter
        lib = self.ctx.lib
        stats = _lib.KrylovStats()
        self.adj_iters, self.adj_tr_iters = [], []
        for s in reversed(sorted(S["ckpt"])):
            e = min(s + self.checkpoint, self.nsteps)
            h_s, c_s = S["ckpt"][s]
            self._init(h_s, ks, alpha, n)
            hs, cs = [h_s], [c_s]
            for _ in range(s + 1, e + 1):
                self.ctx.richards_step(self.dt)
                hs.append(self.ctx.richards_state()[0])
                cs.append(self._species_step(hs[-1], cs[-1], ks, alpha, n)[0])
            for k in range(e, s, -1):
                j = row.get(k)
                dy_t = dldy[0, j].contiguous() if j is not None else None
                dy_c = dldy[1, j].contiguous() if j is not None else None
                tp.n_obs = prm.n_obs = int(self.obs.numel()) if j is not None else 0
                obs = ptr(self.obs) if j is not None else None
                m_out = mu[k % 2]
                rc = lib.wspm_adj_transport_step(self.ctx.h, C.byref(tp), ptr(ks), ptr(hs[k - s]), ptr(h_next),
                                                 ptr(cs[k - s]), ptr(cs[k - s - 1]), obs, ptr(dy_c), ptr(mu_next),
                                                 ptr(m_out), ptr(dldh), ptr(g_ks), ptr(g_an), C.byref(stats))
                self.ctx._ck(rc)
                self.adj_tr_iters.append(int(stats.iters))
                out = lam[k % 2]
                rc = lib.wspm_adj_step(self.ctx.h, C.byref(prm), ptr(ks), ptr(hs[k - s]), ptr(hs[k - s - 1]), obs,
                                       ptr(dy_t), ptr(lam_next), ptr(dldh), ptr(out), ptr(g_ks), ptr(g_an),
                                       C.byref(stats))
                self.ctx._ck(rc)
                self.adj_iters.append(int(stats.iters))
                lam_next, mu_next, h_next = out, m_out, hs[k - s]
        return g_ks * 1.0, g_an[0].clone(), g_an[1].clone()


class RichardsObservations(torch.autograd.Function):
    """y = theta(h_k)[obs] for the model's observation steps, differentiable
    in (log_ks [E], alpha [], n [])."""

    @staticmethod
    def forward(fctx, log_ks, alpha, n, model: RichardsAdjointModel):
        fctx.model = model
        fctx.shapes = (log_ks.device, alpha.device, n.device)
        return model.forward(log_ks.detach(), float(alpha), float(n))

    @staticmethod
    def backward(fctx, gy):
        model = fctx.model
        g_ks, g_a, g_n = model.backward(gy)
        dk, da, dn = fctx.shapes
        return g_ks.to(dk), g_a.to(da), g_n.to(dn), None


class NitrogenObservations(torch.autograd.Function):
    """y [2, n_obs_steps, n_obs] = (theta, NO3) at the model's observation
    nodes and steps, differentiable in (log_ks [E], alpha [], n [])."""

    @staticmethod
    def forward(fctx, log_ks, alpha, n, model: NitrogenAdjointModel):
        fctx.model = model
        fctx.shapes = (log_ks.device, alpha.device, n.device)
        return model.forward(log_ks.detach(), float(alpha), float(n))

    @staticmethod
    def backward(fctx, gy):
        g_ks, g_a, g_n = fctx.model.backward(gy)
        dk, da, dn = fctx.shapes
        return g_ks.to(dk), g_a.to(da), g_n.to(dn), None
