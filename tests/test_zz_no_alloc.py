"""wspm_krylov_noalloc, the Krylov solve that allocates nothing, against
wspm_krylov (the same CG / BiCGSTAB kernels on a context's workspace).

The no-allocation entry point takes the CSR matrix from the caller and keeps
its Krylov vectors, partial sums and solver states in a static device array
of the library.  Its launch shapes are those a context derives for the same
pattern, so the results must be bitwise those of wspm_krylov, with the same
iteration counts: on two calls in a row, the second on a smaller system (the
array keeps the first call's contents), for both methods with Jacobi and
without.  Sizes beyond the array are refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import assembly as oasm
from oracle import mesh as omesh
from oracle import soil as osoil

pytestmark = pytest.mark.gpu

VG = (0.1, 0.40, 0.06, 2.0, 0.5, 1e-3)


def _system(cells, seed):
    pts, el = omesh.structured_tri(cells, 1.0)
    n = len(pts)
    pat = oasm.Pattern(el, n)
    rng = np.random.default_rng(seed)
    h_m = rng.uniform(-1000.0, -50.0, n)
    Cm, Km, thm = osoil.van_genuchten(h_m, VG)
    _, _, thn = osoil.van_genuchten(h_m - 1.0, VG)
    top = np.nonzero(pts[:, -1] == 1.0)[0]
    vals, b = oasm.assemble(pat, pts, thm, thn, h_m, Km, Cm, 1e-4, [(top, np.full(top.size, -75.0))])
    return pts, el, pat, vals, b, h_m


def _noalloc(lib, method, pat, A, dinv, b, x0, tol, maxiter):
    import torch
    from jax_wspm_amd import _lib
    rp = torch.as_tensor(pat.indptr.astype(np.int32), device="cuda")
    ci = torch.as_tensor(pat.indices.astype(np.int32), device="cuda")
    x = torch.empty_like(x0)
    st = _lib.KrylovStats()
    rc = lib.wspm_krylov_noalloc(_lib.METHOD[method], int(pat.n), int(pat.indices.size), _lib.ptr(rp), _lib.ptr(ci),
                                 _lib.ptr(A), _lib.ptr(dinv), _lib.ptr(b), _lib.ptr(x0), _lib.ptr(x), float(tol), 0.0,
                                 int(maxiter), None, C.byref(st))
    torch.cuda.synchronize()
    return rc, x, st


@pytest.mark.parametrize("method,precond", [("bicgstab", "jacobi"), ("cg", "jacobi"), ("bicgstab", "none")])
def test_noalloc_matches_context_solve(method, precond):
    import torch
    from jax_wspm_amd import _lib
    from jax_wspm_amd.device import Context
    lib = _lib.load()
    for call, (cells, seed) in enumerate([(48, 1), (20, 2)]):     # the second call on a smaller system
        print(f"[case] noalloc {method}+{precond} call {call} cells {cells}", flush=True)
        pts, el, pat, vals, b, h_m = _system(cells, seed)
        A = torch.as_tensor(vals, device="cuda")
        bt = torch.as_tensor(b, device="cuda")
        x0 = torch.as_tensor(h_m, device="cuda")
        dinv = torch.as_tensor(oasm.jacobi_inverse(pat, vals), device="cuda") if precond == "jacobi" else None
        ctx = Context(pts, el)
        x_ref, st_ref = ctx.krylov(method, A, bt, x0, dinv=dinv, tol=1e-12, maxiter=400)
        ctx.close()
        rc, x, st = _noalloc(lib, method, pat, A, dinv, bt, x0, 1e-12, 400)
        assert rc == 0, lib.wspm_last_error(None)
        assert st.iters == st_ref["iters"] and st.code == st_ref["code"]
        assert st.res_norm == st_ref["res_norm"] and st.b_norm == st_ref["b_norm"]
        assert torch.equal(x, x_ref), (call, float((x - x_ref).abs().max()))


def test_noalloc_refuses_sizes_beyond_its_array():
    import torch
    from jax_wspm_amd import _lib
    lib = _lib.load()
    n = (1 << 19) + 1
    big = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    v = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = _lib.KrylovStats()
    rc = lib.wspm_krylov_noalloc(_lib.METHOD["cg"], n, n, _lib.ptr(big), _lib.ptr(big), _lib.ptr(v), None,
                                 _lib.ptr(v), _lib.ptr(v), _lib.ptr(v), 1e-10, 0.0, 10, None, C.byref(st))
    assert rc == _lib.WSPM_EINVAL and b"outside" in lib.wspm_last_error(None)
    rc = lib.wspm_krylov_noalloc(_lib.METHOD["gmres"], 10, 10, _lib.ptr(big), _lib.ptr(big), _lib.ptr(v), None,
                                 _lib.ptr(v), _lib.ptr(v), _lib.ptr(v), 1e-10, 0.0, 10, None, C.byref(st))
    assert rc == _lib.WSPM_EINVAL
