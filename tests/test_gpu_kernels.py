"""GPU parity of the HIP kernels against the oracle, called through the C ABI.

Tolerances: assembly entries agree to 1e-13 of their row's largest magnitude
(the same sums in another order); b and the Jacobi diagonal to 1e-13
relative; SpMV to 1e-14 relative per row scale; Krylov solutions to 1e-10
relative L2 with identical iteration counts (the same recurrences, FP64).
"""
import numpy as np
import pytest

from oracle import assembly as oasm
from oracle import krylov as okr
from oracle import mesh as omesh
from oracle import soil as osoil

from tests._helpers import rel_l2, row_scaled_err

pytestmark = pytest.mark.gpu

VG = (0.1, 0.40, 0.06, 2.0, 0.5, 1e-3)
GARDNER = (0.164, 0.45, 0.15, 0.1)


def _ctx(points, elems):
    from jax_wspm_amd.device import Context
    return Context(points, elems)


def _vg_k_exact(h, phi):
    """Mualem-van Genuchten K in 40-digit arithmetic (h >= 0: Ks)."""
    import mpmath
    mpmath.mp.dps = 40
    alpha, _, _, n, m, ks = [mpmath.mpf(float(v)) for v in phi]
    out = np.empty(len(h))
    for j, hv in enumerate(h):
        if hv >= 0:
            out[j] = float(ks)
            continue
        ahn = (alpha * -mpmath.mpf(float(hv))) ** n
        base = 1 + ahn
        out[j] = float(ks * mpmath.sqrt(base ** -m) * (1 - (ahn / base) ** m) ** 2)
    return out


def _state(n, seed, lo=-1200.0, hi=-5.0):
    rng = np.random.default_rng(seed)
    h_m = rng.uniform(lo, hi, n)
    h_n = h_m + rng.uniform(-3.0, 3.0, n)
    return h_m, h_n


CASES = [
    ("tri8_vg", lambda: omesh.structured_tri(8, 1.0), "vg", VG, False),
    ("tri25_gardner", lambda: omesh.structured_tri(25, 15.24), "gardner", GARDNER, False),
    ("tri32_vg_ks", lambda: omesh.structured_tri(32, 1.0), "vg", VG, True),
    ("tet6_vg", lambda: omesh.structured_tet(6, 1.0), "vg", VG, False),
    ("tet20_gardner", lambda: omesh.structured_tet(20, 15.24), "gardner", GARDNER, False),
    ("tet5_vg_ks", lambda: omesh.structured_tet(5, 1.0), "vg", VG, True),
]


@pytest.mark.parametrize("kernel", ["auto", "inc", "lds"])
@pytest.mark.parametrize("name,mk,model,phi,use_ks", CASES, ids=[c[0] for c in CASES])
def test_pattern_soil_assembly(name, mk, model, phi, use_ks, kernel, monkeypatch):
    """Every assembly kernel: the default (2D: register kernel gathering the
    row's neighbourhood once), the register kernel with per-incidence gathers
    (WSPM_ASM_KERNEL=inc) and the LDS-staged one (WSPM_ASM_KERNEL=lds, used
    for long rows)."""
    import torch
    pts, el = mk()
    n = len(pts)
    print(f"[case] {name} n={n} E={len(el)} kernel={kernel}", flush=True)
    if kernel != "auto":
        monkeypatch.setenv("WSPM_ASM_KERNEL", kernel)
    ctx = _ctx(pts, el)
    pat = oasm.Pattern(el, n)
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, pat.indptr) and np.array_equal(ci, pat.indices)

    h_m, h_n = _state(n, 1)
    h_m[:: 7] = np.abs(h_m[:: 7]) * 0.01   # some saturated nodes (h >= 0 branch)
    phi_nodal = np.array(phi, dtype=np.float64)
    ks = None
    if use_ks:
        ks = np.exp(np.log(1e-3) + 0.5 * np.random.default_rng(3).standard_normal(len(el)))
        phi_nodal = phi_nodal.copy()
        phi_nodal[-1] = 1.0
    Cm, Km, thm = osoil.soil(model, h_m, phi_nodal)
    _, _, thn = osoil.soil(model, h_n, phi_nodal)

    gC, gK, gth = ctx.soil(model, phi_nodal, torch.as_tensor(h_m, device="cuda"))
    for g, o in ((gC, Cm), (gth, thm)):
        g = g.cpu().numpy()
        # device pow/exp vs NumPy differ by an ulp
        assert np.allclose(g, o, rtol=1e-12, atol=1e-300), np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300))
    gK = gK.cpu().numpy()
    if model == "vg":
        # the reference's Mualem expression 1-(1-Se^(1/m))^m cancels in dry
        # soil (about 4e-12 relative error here); the device evaluates the
        # same function without cancellation: within 1e-13 of the exact value
        # (40-digit arithmetic) and within the expression's own error of the
        # oracle's
        exact = _vg_k_exact(h_m, phi_nodal)
        assert np.max(np.abs(gK - exact) / exact) <= 1e-13
        assert np.allclose(gK, Km, rtol=1e-10, atol=1e-300)
    else:
        assert np.allclose(gK, Km, rtol=1e-12, atol=1e-300), np.max(np.abs(gK - Km) / Km)

    # Dirichlet: two overlapping ordered lists (later wins)
    top = np.nonzero(pts[:, -1] == pts[:, -1].max())[0]
    side = np.nonzero(pts[:, 0] == pts[:, 0].min())[0]
    bcl = [(top, np.linspace(-10.0, -20.0, top.size)), (side, np.full(side.size, -15.0))]
    ctx.set_dirichlet(bcl)
    dt = 1e-4
    vals, b = oasm.assemble(pat, pts, thm, thn, h_m, Km, Cm, dt, bcl, ks)
    dinv = oasm.jacobi_inverse(pat, vals)
    A, gb, gd = ctx.assemble_richards(h_m, Cm, Km, thm, thn, dt, Ks_elem=ks)
    A, gb, gd = A.cpu().numpy(), gb.cpu().numpy(), gd.cpu().numpy()
    assert row_scaled_err(A, vals, pat.indptr) <= 1e-13
    assert rel_l2(gb, b) <= 1e-13 and np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(gd - dinv) / np.abs(dinv)) <= 1e-13

    # SpMV against scipy on the oracle's values
    x = np.random.default_rng(5).standard_normal(n)
    y = ctx.spmv(torch.as_tensor(vals, device="cuda"), torch.as_tensor(x, device="cuda")).cpu().numpy()
    yr = pat.to_scipy(vals) @ x
    scale = np.abs(pat.to_scipy(np.abs(vals)) @ np.abs(x))
    assert np.max(np.abs(y - yr) / np.maximum(scale, 1e-300)) <= 1e-14
    ctx.close()


def test_register_kernels_agree_bitwise(monkeypatch):
    """The neighbourhood kernels (records staged in LDS, the default; records
    selected from registers, WSPM_ASM_KERNEL=blend) do the incidence kernel's
    arithmetic in the same order: A, b and the Jacobi diagonal are bitwise
    equal (jittered mesh, per-element Ks, Dirichlet rows)."""
    import torch
    pts, el = omesh.structured_tri(17, 1.0)
    rng = np.random.default_rng(8)
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts = pts.copy()
    pts[inner] += 0.3 / 17 * rng.uniform(-1, 1, size=(inner.sum(), 2))
    n = len(pts)
    h_m, h_n = _state(n, 4)
    Cm, Km, thm = osoil.van_genuchten(h_m, VG)
    _, _, thn = osoil.van_genuchten(h_n, VG)
    ks = np.exp(np.log(1e-3) + 0.5 * rng.standard_normal(len(el)))
    top = np.nonzero(pts[:, 1] == 1.0)[0]
    out = []
    for kernel in ("auto", "inc", "blend"):
        print(f"[case] bitwise register kernels {kernel}", flush=True)
        monkeypatch.setenv("WSPM_ASM_KERNEL", kernel)
        ctx = _ctx(pts, el)
        ctx.set_dirichlet([(top, -75.0)])
        A, b, d = ctx.assemble_richards(h_m, Cm, Km, thm, thn, 1e-4, Ks_elem=ks)
        out.append([t.cpu().numpy() for t in (A, b, d)])
        ctx.close()
    for kernel, other in zip(("inc", "blend"), out[1:]):
        for name, x, y in zip("Abd", out[0], other):
            assert np.array_equal(x, y), (kernel, name, int(np.sum(x != y)),
                                          float(np.max(np.abs(x - y) / np.maximum(np.abs(x), 1e-300))))


def _system(cells=12, dim=2, seed=0, dt=1e-4):
    pts, el = (omesh.structured_tri if dim == 2 else omesh.structured_tet)(cells, 1.0)
    n = len(pts)
    pat = oasm.Pattern(el, n)
    rng = np.random.default_rng(seed)
    h_m = rng.uniform(-1000.0, -50.0, n)
    Cm, Km, thm = osoil.van_genuchten(h_m, VG)
    _, _, thn = osoil.van_genuchten(h_m - 1.0, VG)
    top = np.nonzero(pts[:, -1] == 1.0)[0]
    bcl = [(top, np.full(top.size, -75.0))]
    vals, b = oasm.assemble(pat, pts, thm, thn, h_m, Km, Cm, dt, bcl)
    return pts, el, pat, vals, b, h_m, bcl


def _gardner_system(cells, dim, dt, seed=0):
    """Gardner soil on the 15.24 domain, top/bottom Dirichlet: all six
    solver/preconditioner pairs converge here in 1-80 iterations (the
    unpreconditioned solvers stall on the unit-domain van Genuchten rows)."""
    pts, el = (omesh.structured_tri if dim == 2 else omesh.structured_tet)(cells, 15.24)
    n = len(pts)
    pat = oasm.Pattern(el, n)
    rng = np.random.default_rng(seed)
    h_m = rng.uniform(-15.24, -1.0, n)
    Cm, Km, thm = osoil.exponential(h_m, GARDNER)
    _, _, thn = osoil.exponential(h_m - 0.5, GARDNER)
    top = np.nonzero(pts[:, -1] == 15.24)[0]
    bot = np.nonzero(pts[:, -1] == 0.0)[0]
    bcl = [(top, np.full(top.size, -5.0)), (bot, np.full(bot.size, -15.0))]
    vals, b = oasm.assemble(pat, pts, thm, thn, h_m, Km, Cm, dt, bcl)
    return pts, el, pat, vals, b, h_m, bcl


@pytest.mark.parametrize("method", ["cg", "bicgstab", "gmres"])
@pytest.mark.parametrize("precond", ["none", "jacobi"])
@pytest.mark.parametrize("dim", [2, 3])
def test_krylov_matches_oracle(method, precond, dim):
    import torch
    print(f"[case] krylov {method} {precond} dim={dim}", flush=True)
    # dt large enough that the solvers take several (up to ~80) iterations
    pts, el, pat, vals, b, h_m, _ = _gardner_system(14 if dim == 2 else 5, dim, 0.1 if dim == 2 else 1.0)
    A = pat.to_scipy(vals)
    dinv = oasm.jacobi_inverse(pat, vals) if precond == "jacobi" else None
    x_o, info = okr.solve(method, A.dot, b, h_m, dinv, 1e-10, 0.0, 100, 30)
    assert np.linalg.norm(b - A @ x_o) <= 1e-9 * np.linalg.norm(b)   # the oracle converged
    ctx = _ctx(pts, el)
    dv = torch.as_tensor(vals, device="cuda")
    x_g, st = ctx.krylov(method, dv, torch.as_tensor(b, device="cuda"), torch.as_tensor(h_m, device="cuda"),
                         dinv=None if dinv is None else torch.as_tensor(dinv, device="cuda"),
                         tol=1e-10, maxiter=100, restart=30)
    x_g = x_g.cpu().numpy()
    assert st["iters"] == info["iters"], (st, info)
    assert rel_l2(x_g, x_o) <= 1e-10
    ctx.close()


def test_krylov_small_restart_and_truncation():
    """GMRES with restart 5 needs several cycles; CG truncated at maxiter=3."""
    import torch
    pts, el, pat, vals, b, h_m, _ = _gardner_system(10, 2, 10.0)
    A = pat.to_scipy(vals)
    dinv = oasm.jacobi_inverse(pat, vals)
    ctx = _ctx(pts, el)
    dv, db, dx0, dd = (torch.as_tensor(a, device="cuda") for a in (vals, b, h_m, dinv))
    x_o, info = okr.gmres(A.dot, b, h_m, 1e-10, 0.0, 5, 100, dinv)
    x_g, st = ctx.krylov("gmres", dv, db, dx0, dinv=dd, restart=5)
    assert st["iters"] == info["iters"] and st["arnoldi"] == info["arnoldi"]
    assert rel_l2(x_g.cpu().numpy(), x_o) <= 1e-10
    x_o, info = okr.cg(A.dot, b, h_m, 1e-10, 0.0, 3, dinv)
    x_g, st = ctx.krylov("cg", dv, db, dx0, dinv=dd, maxiter=3)
    assert st["iters"] == 3 == info["iters"]
    assert rel_l2(x_g.cpu().numpy(), x_o) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("method", ["cg", "bicgstab", "gmres"])
def test_mixed_precision_refinement(method):
    import torch
    print(f"[case] mixed {method}", flush=True)
    pts, el, pat, vals, b, h_m, _ = _gardner_system(16, 2, 1.0)
    A = pat.to_scipy(vals)
    dinv = oasm.jacobi_inverse(pat, vals)
    x_ref = np.linalg.solve(A.toarray(), b)
    ctx = _ctx(pts, el)
    dv, db, dx0, dd = (torch.as_tensor(a, device="cuda") for a in (vals, b, h_m, dinv))
    x_g, st = ctx.krylov(method, dv, db, dx0, dinv=dd, precision="mixed", tol=1e-12, inner_tol=1e-5)
    assert st["refinements"] >= 1
    assert rel_l2(x_g.cpu().numpy(), x_ref) <= 1e-10
    ctx.close()


def test_deterministic_repeat():
    """Same inputs -> bitwise identical assembly and solve (no float atomics)."""
    import torch
    pts, el, pat, vals, b, h_m, bcl = _system(cells=40, dim=2, dt=1.0)
    ctx = _ctx(pts, el)
    ctx.set_dirichlet(bcl)
    Cm, Km, thm = osoil.van_genuchten(h_m, VG)
    outs = []
    for _ in range(2):
        A, gb, gd = ctx.assemble_richards(h_m, Cm, Km, thm, thm, 1.0)
        x, _ = ctx.krylov("bicgstab", A, gb, torch.as_tensor(h_m, device="cuda"), dinv=gd)
        outs.append((A.cpu().numpy(), gb.cpu().numpy(), x.cpu().numpy()))
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(a, c)
    ctx.close()


@pytest.mark.parametrize("method", ["cg", "bicgstab", "gmres"])
@pytest.mark.parametrize("precond", ["block_jacobi", "ssor", "ilu"])
def test_krylov_general_preconditioners_match_oracle(method, precond):
    """Block Jacobi, SSOR and ILU applied between the Krylov kernels (level-
    scheduled triangular sweeps) against the oracle's factorisations."""
    import torch
    from oracle import precond as opc
    print(f"[case] krylov {method} {precond}", flush=True)
    pts, el, pat, vals, b, h_m, _ = _gardner_system(14, 2, 0.1)
    A = pat.to_scipy(vals)
    x_o, info = okr.solve(method, A.dot, b, h_m, opc.make(precond, A), 1e-10, 0.0, 100, 30)
    ctx = _ctx(pts, el)
    x_g, st = ctx.krylov_pc(method, precond, torch.as_tensor(vals, device="cuda"), torch.as_tensor(b, device="cuda"),
                            torch.as_tensor(h_m, device="cuda"), tol=1e-10, maxiter=100, restart=30)
    ctx.close()
    assert st["iters"] == info["iters"], (st, info)
    if method == "gmres":
        assert st["arnoldi"] == info["arnoldi"]
    assert rel_l2(x_g.cpu().numpy(), x_o) <= 1e-10


def test_ell_and_csr_time_steps_agree(monkeypatch):
    """The time-step solves read the operator in ELL order where the rows
    allow it; WSPM_SPMV=csr keeps CSR.  Both formats give the same
    trajectory (different summation order only: <= 1e-13) with the same
    Picard and Krylov iteration counts."""
    import numpy as np
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    from jax_wspm_amd.richards import RichardsSolver
    print("[case] ELL vs CSR time steps 48x48", flush=True)
    out = {}
    for fmt in ("ell", "csr"):
        if fmt == "csr":
            monkeypatch.setenv("WSPM_SPMV", "csr")
        else:
            monkeypatch.delenv("WSPM_SPMV", raising=False)
        pb = from_config(baseline_vg_config(48, 2, solver="bicgstab", precond="jacobi",
                                            ks_field={"sigma": 0.5, "length": 0.1, "seed": 1}, nsteps=6, dt=1e-4))
        s = RichardsSolver(pb)
        res = s.solve()
        assert s.ctx.spmv_format == fmt
        s.ctx.close()
        out[fmt] = res
    a, b = out["ell"], out["csr"]
    assert list(a["iterations"]) == list(b["iterations"])
    assert list(a["krylov_iterations"]) == list(b["krylov_iterations"])
    for k in range(len(a["times"])):
        assert np.linalg.norm(a["pressure_head"][k] - b["pressure_head"][k]) <= 1e-13 * np.linalg.norm(b["pressure_head"][k])


def _jittered_problem(cells, solver, seed=3, nsteps=5):
    from dataclasses import replace
    from jax_wspm_amd.problems import baseline_vg_config, from_config
    pb = from_config(baseline_vg_config(cells, 2, solver=solver, precond="jacobi", nsteps=nsteps, dt=1e-4,
                                        ks_field={"sigma": 0.5, "length": 0.1, "seed": seed}))
    pts = pb.points.copy()
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts[inner] += 0.3 / cells * np.random.default_rng(seed).uniform(-1, 1, size=(inner.sum(), 2))
    return replace(pb, points=pts)


@pytest.mark.parametrize("solver", ["bicgstab", "cg"])
def test_fan_order_time_steps(monkeypatch, solver):
    """The time-step operator in fan order (each row: the node, then its
    neighbours around it, so the assembly adds every triangle into slots
    fixed at compile time) against the sorted ELL order of the neighbourhood
    kernel (WSPM_ASM_FAN=0) and against the oracle, on a jittered 33x33 mesh
    with per-element Ks: the same Picard and Krylov counts, trajectories
    within 1e-12 of each other (summation order only) and 1e-10 of the
    oracle."""
    from jax_wspm_amd.richards import RichardsSolver
    from oracle import picard as opicard
    from tests._helpers import to_oracle
    pb = _jittered_problem(33, solver)
    out = {}
    for fan in ("1", "0"):
        print(f"[case] fan order {solver} WSPM_ASM_FAN={fan}", flush=True)
        monkeypatch.setenv("WSPM_ASM_FAN", fan)
        s = RichardsSolver(pb)
        out[fan] = s.solve()
        assert s.ctx.operator_order == ("fan" if fan == "1" else "sorted")
        s.ctx.close()
    a, b = out["1"], out["0"]
    assert list(a["iterations"]) == list(b["iterations"])
    assert list(a["krylov_iterations"]) == list(b["krylov_iterations"])
    o = opicard.run(to_oracle(pb))
    assert list(a["iterations"]) == list(o["iterations"])
    for k in range(len(a["times"])):
        assert rel_l2(a["pressure_head"][k], b["pressure_head"][k]) <= 1e-12, k
        assert rel_l2(a["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
        assert rel_l2(a["theta"][k], o["theta"][k]) <= 1e-10, k


def test_fan_order_needs_consistent_orientation():
    """One triangle listed in the opposite orientation: its vertices' rows are
    no longer fans of consistently oriented triangles, so the operator keeps
    the sorted ELL order (and the results their parity)."""
    from dataclasses import replace
    from jax_wspm_amd.richards import RichardsSolver
    from oracle import picard as opicard
    from tests._helpers import to_oracle
    pb = _jittered_problem(17, "bicgstab", nsteps=3)
    el = pb.elements.copy()
    el[40] = el[40][::-1]
    pb = replace(pb, elements=el)
    print("[case] fan order, one flipped triangle", flush=True)
    s = RichardsSolver(pb)
    g = s.solve()
    assert s.ctx.spmv_format == "ell" and s.ctx.operator_order == "sorted"
    s.ctx.close()
    o = opicard.run(to_oracle(pb))
    assert list(g["iterations"]) == list(o["iterations"])
    for k in range(len(g["times"])):
        assert rel_l2(g["pressure_head"][k], o["pressure_head"][k]) <= 1e-10, k
