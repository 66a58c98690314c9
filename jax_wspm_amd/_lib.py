"""ctypes binding of the C ABI declared in include/wspm.h.

The shared library is built in-tree (``jax_wspm_amd/_lib/libwspm.so``) by
``jax_wspm_amd.build``.  There is no fallback: if the library is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libwspm.so"

WSPM_OK, WSPM_EINVAL, WSPM_EHIP, WSPM_ENOMEM, WSPM_ESTATE = 0, -1, -2, -3, -4
SOIL = {"gardner": 0, "exponential": 0, "vg": 1, "van_genuchten": 1}
METHOD = {"cg": 0, "bicgstab": 1, "gmres": 2}
STEP_METHOD = dict(METHOD, direct=3)       # Picard steps also take the banded-LU direct solve
PRECOND = {"none": 0, None: 0, "jacobi": 1, "block_jacobi": 2, "ssor": 3, "ilu": 4}
PRECISION = {"fp64": 0, "mixed": 1}
INNER = {"krylov": 0, "jacobi": 1}
PICARD = {"rel": 0, "abs": 1}


class WspmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wspm error {code}: {msg}")
        self.code = code


class KrylovStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("arnoldi", C.c_int32), ("code", C.c_int32),
                ("refinements", C.c_int32), ("res_norm", C.c_double), ("b_norm", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("picard_iters", C.c_int32), ("krylov_iters", C.c_int32), ("converged", C.c_int32),
                ("reserved", C.c_int32), ("err", C.c_double)]


class RichardsParams(C.Structure):
    _fields_ = [("soil_model", C.c_int32), ("method", C.c_int32), ("precond", C.c_int32),
                ("precision", C.c_int32), ("phi", C.c_double * 6), ("tol", C.c_double),
                ("atol", C.c_double), ("maxiter", C.c_int32), ("restart", C.c_int32),
                ("picard_tol", C.c_double), ("picard_maxit", C.c_int32), ("picard_mode", C.c_int32),
                ("inner_tol", C.c_double), ("inner_method", C.c_int32), ("inner_iters", C.c_int32)]


class TransportParams(C.Structure):
    _fields_ = [("lumped", C.c_int32), ("inlet_cauchy", C.c_int32), ("outlet_freedrain", C.c_int32),
                ("reserved", C.c_int32), ("dt", C.c_double), ("mu", C.c_double), ("rho_kd", C.c_double),
                ("mu_src", C.c_double), ("cauchy_flux", C.c_double), ("c_inlet", C.c_double)]


class AdjParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("phi", C.c_double * 6), ("tol", C.c_double), ("maxiter", C.c_int32),
                ("n_obs", C.c_int32)]


class AdjTrParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("phi", C.c_double * 6), ("DL", C.c_double), ("DT", C.c_double),
                ("Dm", C.c_double), ("mu", C.c_double * 3), ("rho_kd", C.c_double), ("cauchy_flux", C.c_double),
                ("tol", C.c_double), ("maxiter", C.c_int32), ("n_obs", C.c_int32)]


EDGE = {"inlet": 0, "outlet": 1}

_P = C.c_void_p
_I = C.c_int32
_D = C.c_double

# name -> (restype, argtypes); must match include/wspm.h
PROTOTYPES = {
    "wspm_version": (_I, []),
    "wspm_last_error": (C.c_char_p, [_P]),
    "wspm_ctx_create": (_I, [_I, _I, _I, _P, _P, _P, C.POINTER(_P)]),
    "wspm_ctx_destroy": (None, [_P]),
    "wspm_mesh_rcm": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "wspm_set_stream": (_I, [_P, _P]),
    "wspm_ctx_info": (_I, [_P, C.POINTER(C.c_int64)]),
    "wspm_get_pattern": (_I, [_P, _P, _P]),
    "wspm_bc_set": (_I, [_P, _I, _P, _P]),
    "wspm_soil": (_I, [_P, _I, _P, _I, _P, _P, _P, _P]),
    "wspm_assemble_richards": (_I, [_P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
    "wspm_spmv": (_I, [_P, _P, _P, _P]),
    "wspm_krylov": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _D, _I, _I, C.POINTER(KrylovStats)]),
    "wspm_krylov_noalloc": (_I, [_I, _I, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _D, _D, _I, _P,
                                 C.POINTER(KrylovStats)]),
    "wspm_krylov_pc": (_I, [_P, _I, _I, _P, _P, _P, _P, _D, _D, _I, _I, C.POINTER(KrylovStats)]),
    "wspm_krylov_mixed": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _D, _I, _I, _D, _I,
                               C.POINTER(KrylovStats)]),
    "wspm_richards_init": (_I, [_P, C.POINTER(RichardsParams), _P, _P]),
    "wspm_richards_step": (_I, [_P, _D, C.POINTER(StepStats)]),
    "wspm_richards_state": (_I, [_P, _P, _P, _P]),
    "wspm_timing_enable": (_I, [_P, _I]),
    "wspm_timing_read": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "wspm_timing_read_asm": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "wspm_launch_count": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "wspm_timing_read_over": (_I, [_P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int64)]),
    "wspm_edge_set": (_I, [_P, _I, _I, _P]),
    "wspm_richards_edges": (_I, [_P, _I, _D, _I]),
    "wspm_transport_fields": (_I, [_P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P]),
    "wspm_assemble_transport": (_I, [_P, C.POINTER(TransportParams), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _P]),
    "wspm_adj_step": (_I, [_P, C.POINTER(AdjParams), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                           C.POINTER(KrylovStats)]),
    "wspm_adj_transport_step": (_I, [_P, C.POINTER(AdjTrParams), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     C.POINTER(KrylovStats)]),
    "wspm_ens_create": (_I, [_I, _I, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "wspm_ens_richards_init": (_I, [_P, C.POINTER(RichardsParams), _P, _P, _P]),
    "wspm_ens_step_stats": (_I, [_P, _P, _P, _P, _P]),
    "wspm_ens_shape": (_I, [_P, C.POINTER(_I)]),
    "wspm_domain_integral": (_I, [_P, _P, C.POINTER(_D)]),
    "wspm_direct": (_I, [_P, _P, _P, _P, C.POINTER(_I)]),
    "wspm_direct_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "wspm_element_size": (_I, [_P, _P]),
    "wspm_peclet_courant": (_I, [_P, _P, _P, _P, _P, _P, _D, _I, _P, _P, C.POINTER(_D)]),
}

_lib = None


def load(path: os.PathLike | None = None):
    """Load libwspm.so (once).  Raises if it is missing: there is no fallback.

    PyTorch is imported first: its wheel bundles a HIP runtime whose SONAME
    (libamdhip64.so.7) is the one libwspm.so links against, so the library
    binds to the runtime that owns the tensors and streams handed to it.
    Loaded the other way round, the process would carry two HIP runtimes."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - plain C use without torch
        pass
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python -m jax_wspm_amd.build`")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, ctx=None):
    if rc != WSPM_OK:
        lib = load()
        msg = lib.wspm_last_error(ctx)
        raise WspmError(rc, msg.decode() if msg else "")


def ptr(t):
    """Device (or host) address of a tensor/ndarray, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    return t.ctypes.data_as(C.c_void_p)
