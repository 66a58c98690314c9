"""Adjoint of Richards flow + the nitrogen species chain (NH4 -> NO2 -> NO3):
gradients of theta and NO3 observations w.r.t. per-element log Ks, alpha and
n, against central differences of the oracle's coupled forward (an
independent restatement: oracle Picard flow, oracle flux recovery with the
per-element Ks, oracle species systems) and of the GPU forward itself.

Both forwards run with tight Picard / Krylov tolerances so that the finite
differences resolve the converged-step model the adjoint differentiates.
Bar: relative error <= 1e-6.
"""
import numpy as np
import pytest
import torch

from oracle import mesh as omesh
from oracle import picard as opicard
from oracle import transport as otr

pytestmark = pytest.mark.gpu

TS, TR = 0.40, 0.06
TIGHT = dict(tol=1e-14, maxiter=400, picard_tol=1e-13, picard_maxit=300)
SPECIES = dict(DL=0.05, DT=0.01, Dm=1e-4, mu=(0.2, 0.5, 0.05), rho_kd=0.5, cauchy_flux=-0.05,
               c_inlet=(1.0, 0.0, 0.4), tr_tol=1e-14, tr_maxiter=400)


def _problem(cells=8, seed=0):
    from jax_wspm_amd import mesh as M
    pts, tri = omesh.structured_tri(cells, 1.0)
    tri = np.asarray(tri)
    top = np.nonzero(pts[:, 1] == 1.0)[0]
    ks = M.correlated_log_ks(pts, tri, cells, mean_ks=1e-2, sigma=0.5, length=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    upper = np.nonzero((pts[:, 1] >= 0.5) & (pts[:, 1] < 1.0))[0]
    obs = np.sort(rng.choice(upper, size=10, replace=False))
    return pts, tri, top, [(top, np.full(top.shape, -5.0))], np.full(len(pts), -100.0), np.log(ks), obs


def _oracle_obs(pts, tri, top, bcs, h0, log_ks, alpha, n, obs, steps, obs_steps, dt):
    """theta and NO3 at the observation nodes and steps, oracle forward."""
    ks = np.exp(log_ks)
    phi = (alpha, TS, TR, n, 1 - 1 / n, 1.0)
    pb = opicard.Problem(points=pts, elems=tri, soil_model="vg", phi=phi, bc_lists=bcs, h0=h0, dt0=dt,
                         nsteps=steps, adaptive=False, picard_tol=1e-13, picard_maxit=300, picard_mode="rel",
                         solver="bicgstab", precond="jacobi", tol=1e-14, maxiter=400, Ks_elem=ks)
    r = opicard.run(pb)
    S = SPECIES
    prm = dict(DL=S["DL"], DT=S["DT"], Dm=S["Dm"], rho=1.0, Kd=S["rho_kd"], mu1=S["mu"][0], mu2=S["mu"][1],
               mu3=S["mu"][2])
    c = [np.zeros(len(pts)) for _ in range(3)]
    y = []
    for k in range(1, steps + 1):
        h = r["pressure_head"][k - 1]
        f = otr.fluxes_dispersion(pts, tri, h, np.array(phi), S["DL"], S["DT"], S["Dm"], ks_elem=ks)
        new = []
        for s in (1, 2, 3):
            A, b = otr.nitrogen_system(pts, tri, s, f, prm, c[s - 1], new[-1] if new else None, dt, None, top,
                                       np.zeros(0, dtype=int), S["cauchy_flux"], S["c_inlet"][s - 1])
            new.append(otr._solve(A, b, c[s - 1], "bicgstab", "jacobi", ktol=1e-14, maxiter=400))
        c = new
        if k in obs_steps:
            y.append(np.stack([f[6][obs], c[2][obs]]))
    return np.stack(y, axis=1)


def _model(pts, tri, top, bcs, h0, obs, steps, obs_steps, dt, checkpoint=2):
    from jax_wspm_amd.adjoint import NitrogenAdjointModel
    return NitrogenAdjointModel(pts, tri, bcs, h0, dt, steps, obs, obs_steps, TS, TR, inlet_nodes=top,
                                checkpoint=checkpoint, **SPECIES, **TIGHT)


def test_nitrogen_adjoint_matches_oracle_finite_differences():
    pts, tri, top, bcs, h0, log_ks, obs = _problem()
    steps, obs_steps, dt = 4, [2, 4], 0.05
    alpha, n = 0.1, 2.0
    rng = np.random.default_rng(1)
    w = rng.normal(size=(2, len(obs_steps), len(obs)))
    w[1] *= 5.0                           # NO3 values are below theta's: weight them up
    print("[case] nitrogen adjoint vs oracle FD 8x8, 4 steps", flush=True)
    model = _model(pts, tri, top, bcs, h0, obs, steps, obs_steps, dt)
    y = model.forward(torch.tensor(log_ks, device="cuda"), alpha, n).cpu().numpy()
    g_ks, g_a, g_n = model.backward(torch.tensor(w, device="cuda"))
    g_ks = g_ks.cpu().numpy()
    model.close()
    yo = _oracle_obs(pts, tri, top, bcs, h0, log_ks, alpha, n, obs, steps, obs_steps, dt)
    assert np.abs(y - yo).max() <= 1e-10 * np.abs(yo).max(), np.abs(y - yo).max()
    assert np.abs(y[1]).max() > 1e-4            # NO3 reaches the observation nodes

    def L(log_ks_, a_, n_):
        return float((w * _oracle_obs(pts, tri, top, bcs, h0, log_ks_, a_, n_, obs, steps, obs_steps, dt)).sum())

    ga = (L(log_ks, alpha * (1 + 1e-6), n) - L(log_ks, alpha * (1 - 1e-6), n)) / (2e-6 * alpha)
    gn = (L(log_ks, alpha, n * (1 + 1e-6)) - L(log_ks, alpha, n * (1 - 1e-6))) / (2e-6 * n)
    assert abs(float(g_a) - ga) <= 1e-6 * abs(ga), (float(g_a), ga)
    assert abs(float(g_n) - gn) <= 1e-6 * abs(gn), (float(g_n), gn)
    for e in np.argsort(-np.abs(g_ks))[:5]:
        lp, lm = log_ks.copy(), log_ks.copy()
        lp[e] += 1e-5
        lm[e] -= 1e-5
        ge = (L(lp, alpha, n) - L(lm, alpha, n)) / 2e-5
        assert abs(g_ks[e] - ge) <= 1e-6 * np.abs(g_ks).max(), (e, g_ks[e], ge)


@pytest.mark.parametrize("which", ["no3", "both"])
def test_nitrogen_adjoint_directional_derivatives(which):
    """16 random directions against central differences of the GPU forward,
    through torch.autograd; "no3" weighs the NO3 observations only, so the
    species adjoint alone carries the gradient."""
    from jax_wspm_amd.adjoint import NitrogenObservations
    pts, tri, top, bcs, h0, log_ks, obs = _problem(10, seed=3)
    steps, obs_steps, dt = 6, [3, 6], 0.05
    print(f"[case] nitrogen adjoint directional FD 10x10, 6 steps, checkpoint 4, {which}", flush=True)
    model = _model(pts, tri, top, bcs, h0, obs, steps, obs_steps, dt, checkpoint=4)
    w = np.random.default_rng(2).normal(size=(2, len(obs_steps), len(obs)))
    if which == "no3":
        w[0] = 0.0
    w = torch.tensor(w, device="cuda")
    lk = torch.tensor(log_ks, dtype=torch.float64, device="cuda", requires_grad=True)
    a = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    nn = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    loss = (NitrogenObservations.apply(lk, a, nn, model) * w).sum()
    loss.backward()
    g = np.concatenate([lk.grad.cpu().numpy(), [float(a.grad)], [float(nn.grad)]])
    assert max(model.adj_iters) < 1000 and max(model.adj_tr_iters) < 3000

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
