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
                                       ptr(self.obs) if j is not None else None, ptr(dy), ptr(lam_next), ptr(out),
                                       ptr(g_ks), ptr(g_an), C.byref(stats))
                self.ctx._ck(rc)
                self.adj_iters.append(int(stats.iters))
                lam_next = out
        return g_ks * 1.0, g_an[0].clone(), g_an[1].clone()

    def close(self):
        self.ctx.close()


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
