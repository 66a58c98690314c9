"""Discrete adjoint: gradients of theta observations w.r.t. per-element log Ks,
alpha and n, against finite differences of the oracle's forward (an
independent implementation of the same discrete model) and of the GPU forward
itself (16 random directions), through the torch.autograd.Function.

Both forwards run with tight Picard / Krylov tolerances so that the finite
differences resolve the converged-step model the adjoint differentiates.
Bar: relative error <= 1e-6.
"""
import numpy as np
import pytest
import torch

from oracle import mesh as omesh
from oracle import picard as opicard

pytestmark = pytest.mark.gpu

TS, TR = 0.40, 0.06
TIGHT = dict(tol=1e-14, maxiter=400, picard_tol=1e-13, picard_maxit=300)


def _problem(cells=8, seed=0):
    from jax_wspm_amd import mesh as M
    pts, tri = omesh.structured_tri(cells, 1.0)
    tri = np.asarray(tri)
    top = np.nonzero(pts[:, 1] == 1.0)[0]
    ks = M.correlated_log_ks(pts, tri, cells, mean_ks=1e-2, sigma=0.5, length=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    upper = np.nonzero((pts[:, 1] >= 0.5) & (pts[:, 1] < 1.0))[0]
    obs = np.sort(rng.choice(upper, size=10, replace=False))
    # wetting front from the top (h = -5) into h0 = -100: Picard needs 15-20
    # iterations per step at the tight tolerance, theta moves by ~0.1
    return pts, tri, [(top, np.full(top.shape, -5.0))], np.full(len(pts), -100.0), np.log(ks), obs


def _oracle_loss(pts, tri, bcs, h0, log_ks, alpha, n, obs, steps, obs_steps, w, dt):
    pb = opicard.Problem(points=pts, elems=tri, soil_model="vg", phi=(alpha, TS, TR, n, 1 - 1 / n, 1.0),
                         bc_lists=bcs, h0=h0, dt0=dt, nsteps=steps, adaptive=False, picard_tol=1e-13,
                         picard_maxit=300, picard_mode="rel", solver="bicgstab", precond="jacobi", tol=1e-14,
                         maxiter=400, Ks_elem=np.exp(log_ks))
    r = opicard.run(pb)
    return sum(float(w[j] @ r["theta"][k - 1][obs]) for j, k in enumerate(obs_steps))


def _model(pts, tri, bcs, h0, obs, steps, obs_steps, dt, checkpoint=2):
    from jax_wspm_amd.adjoint import RichardsAdjointModel
    return RichardsAdjointModel(pts, tri, bcs, h0, dt, steps, obs, obs_steps, TS, TR, checkpoint=checkpoint,
                                **TIGHT)


def test_adjoint_gradient_matches_oracle_finite_differences():
    pts, tri, bcs, h0, log_ks, obs = _problem()
    steps, obs_steps, dt = 4, [2, 4], 0.05
    alpha, n = 0.1, 2.0
    w = np.random.default_rng(1).normal(size=(len(obs_steps), len(obs)))
    print("[case] adjoint vs oracle FD 8x8, 4 steps", flush=True)
    model = _model(pts, tri, bcs, h0, obs, steps, obs_steps, dt)
    lk = torch.tensor(log_ks, dtype=torch.float64, device="cuda")
    y = model.forward(lk, alpha, n)
    g_ks, g_a, g_n = model.backward(torch.tensor(w, device="cuda"))
    g_ks = g_ks.cpu().numpy()
    model.close()
    # the oracle forward reproduces the GPU observations
    L0 = _oracle_loss(pts, tri, bcs, h0, log_ks, alpha, n, obs, steps, obs_steps, w, dt)
    assert abs(float((y.cpu().numpy() * w).sum()) - L0) <= 1e-11 * abs(L0)

    def fd(f, x, eps):
        return (f(x + eps) - f(x - eps)) / (2 * eps)

    ga = fd(lambda a: _oracle_loss(pts, tri, bcs, h0, log_ks, a, n, obs, steps, obs_steps, w, dt), alpha, 1e-6 * alpha)
    gn = fd(lambda m: _oracle_loss(pts, tri, bcs, h0, log_ks, alpha, m, obs, steps, obs_steps, w, dt), n, 1e-6 * n)
    assert abs(float(g_a) - ga) <= 1e-6 * abs(ga), (float(g_a), ga)
    assert abs(float(g_n) - gn) <= 1e-6 * abs(gn), (float(g_n), gn)
    idx = np.argsort(-np.abs(g_ks))[:6]           # the elements that matter most
    for e in idx:
        def f(v, e=e):
            lk2 = log_ks.copy()
            lk2[e] = v
            return _oracle_loss(pts, tri, bcs, h0, lk2, alpha, n, obs, steps, obs_steps, w, dt)
        ge = fd(f, log_ks[e], 1e-5)
        assert abs(g_ks[e] - ge) <= 1e-6 * np.abs(g_ks).max(), (e, g_ks[e], ge)


def test_adjoint_directional_derivatives_and_autograd():
    """16 random directions in (log Ks, alpha, n) against central differences
    of the GPU forward; gradients arrive through torch.autograd."""
    from jax_wspm_amd.adjoint import RichardsObservations
    pts, tri, bcs, h0, log_ks, obs = _problem(10, seed=3)
    steps, obs_steps, dt = 6, [3, 6], 0.05
    print("[case] adjoint directional FD 10x10, 6 steps, checkpoint 4", flush=True)
    model = _model(pts, tri, bcs, h0, obs, steps, obs_steps, dt, checkpoint=4)
    w = torch.tensor(np.random.default_rng(2).normal(size=(len(obs_steps), len(obs))), device="cuda")
    lk = torch.tensor(log_ks, dtype=torch.float64, device="cuda", requires_grad=True)
    a = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    nn = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    loss = (RichardsObservations.apply(lk, a, nn, model) * w).sum()
    loss.backward()
    g = np.concatenate([lk.grad.cpu().numpy(), [float(a.grad)], [float(nn.grad)]])
    assert max(model.adj_iters) < 1000

    def L(p):
        with torch.no_grad():
            y = model.forward(torch.tensor(p[:-2], device="cuda"), float(p[-2]), float(p[-1]))
            return float((y * w).sum())

    p0 = np.concatenate([log_ks, [0.1, 2.0]])
    rng = np.random.default_rng(4)
    scale = np.concatenate([np.ones(len(log_ks)), [0.1, 2.0]])
    for _ in range(16):
        d = rng.normal(size=p0.size) * scale
        eps = 1e-6
        fdv = (L(p0 + eps * d) - L(p0 - eps * d)) / (2 * eps)
        assert abs(g @ d - fdv) <= 1e-6 * abs(fdv), (g @ d, fdv)
    model.close()
