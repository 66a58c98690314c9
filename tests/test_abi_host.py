"""CPU checks of the boundary and the host layer (no GPU compute calls)."""
import ctypes
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import mesh as omesh

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "wspm.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wspm_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from jax_wspm_amd import _lib
    assert header_functions() == sorted(_lib.PROTOTYPES)


def test_library_exports_every_declared_symbol():
    from jax_wspm_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libwspm.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in header_functions():
        assert hasattr(lib, name), name
    loaded = _lib.load()
    assert loaded.wspm_version() >= 10000
    # argument validation needs no GPU: null output pointer is rejected
    rc = loaded.wspm_ctx_create(2, 3, 1, None, None, None, None)
    assert rc == _lib.WSPM_EINVAL
    assert b"null" in loaded.wspm_last_error(None)


def test_struct_layouts_match_header(tmp_path):
    """Every parameter struct of the binding has the size and field offsets
    the C compiler gives the header's struct (a C program built with gcc
    against include/wspm.h prints them)."""
    import shutil
    import subprocess
    from jax_wspm_amd import _lib
    pairs = {"wspm_krylov_stats": _lib.KrylovStats, "wspm_step_stats": _lib.StepStats,
             "wspm_richards_params": _lib.RichardsParams, "wspm_transport_params": _lib.TransportParams,
             "wspm_adj_params": _lib.AdjParams, "wspm_adj_tr_params": _lib.AdjTrParams}
    assert ctypes.sizeof(_lib.RichardsParams) == 16 + 48 + 16 + 8 + 8 + 8 + 8 + 8
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "wspm.h"', "int main(void) {"]
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = ln.split()
        got[(cname, field)] = int(val)
    for cname, cls in pairs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


@pytest.mark.parametrize("N,L", [(25, 15.24), (50, 1.0), (7, 2.0)])
def test_product_square_generator_matches_oracle(N, L):
    from jax_wspm_amd import mesh as M
    p, t = M.unit_square(N, L)
    po, to = omesh.structured_tri(N, L)
    assert np.array_equal(p, po) and np.array_equal(t, to)


@pytest.mark.parametrize("N", [3, 20])
def test_product_cube_generator_matches_oracle(N):
    from jax_wspm_amd import mesh as M
    p, t = M.unit_cube(N, 15.24)
    po, to = omesh.structured_tet(N, 15.24)
    assert np.array_equal(p, po) and np.array_equal(t, to)


def test_product_generators_match_fixture_checksums(golden):
    import hashlib
    from jax_wspm_amd import mesh as M
    ck = json.loads((golden / "mesh_checksums.json").read_text())
    for name, d in ck.items():
        gen = M.unit_square if d["kind"] == "tri" else M.unit_cube
        p, t = gen(d["cells"], d["length"])
        assert hashlib.sha256(p.tobytes()).hexdigest() == d["points_sha256"], name
        assert hashlib.sha256(t.tobytes()).hexdigest() == d["elems_sha256"], name


def test_boundary_sets_and_problem_definitions():
    from jax_wspm_amd import mesh as M
    from jax_wspm_amd.config import SimulationConfig
    from jax_wspm_amd.problems import TRACY_BOTTOM, baseline_vg_config, from_config
    assert TRACY_BOTTOM == -15.0
    for tc in ("Test1", "Test2", "Test3"):
        cfg = SimulationConfig(test_case=tc)
        pb = from_config(cfg)
        top, bot = omesh.boundary_nodes_2d(pb.points, tc)
        assert np.array_equal(pb.bc_lists[0][0], top) and np.array_equal(pb.bc_lists[1][0], bot)
        assert np.all(pb.bc_lists[1][1] == -15.0)
        assert pb.picard_mode == "rel" and pb.picard_tol == 1e-6 and pb.picard_maxit == 100
        # Tracy top value equals -15.24 at the corners (sin = 0)
        x = pb.points[top, 0]
        hup = pb.bc_lists[0][1]
        assert np.allclose(hup[(x == 0) | (x == 15.24)], -15.24, atol=1e-9)
    cfg = SimulationConfig(test_case="Test3D")
    pb = from_config(cfg)
    assert pb.n == 9261 and pb.picard_mode == "abs" and pb.picard_maxit == 30 and pb.theta_at_start
    t3, b3 = omesh.boundary_nodes_3d(pb.points)
    assert np.array_equal(pb.bc_lists[0][0], t3) and np.array_equal(pb.bc_lists[1][0], b3)
    pb = from_config(baseline_vg_config(25, 2))
    assert pb.n == 676 and pb.bc_lists[0][0].size == 26 and np.all(pb.bc_lists[0][1] == -75.0)
    assert pb.phi == (0.1, 0.40, 0.06, 2.0, 0.5, 1e-3)
    pb = from_config(baseline_vg_config(16, 3))
    assert pb.n == 17 ** 3 and np.all(pb.points[pb.bc_lists[0][0], 2] == 1.0)


def test_correlated_ks_field_statistics():
    from jax_wspm_amd import mesh as M
    p, t = M.unit_square(128, 1.0)
    ks = M.correlated_log_ks(p, t, 128, mean_ks=1e-3, sigma=0.5, length=0.05, seed=0)
    z = (np.log(ks) - np.log(1e-3)) / 0.5
    assert ks.shape == (t.shape[0],) and np.all(ks > 0)
    assert abs(z.mean()) < 1e-12 and abs(z.std() - 1) < 1e-12
    # spatial correlation: neighbouring cells are strongly correlated
    zc = z[0::2].reshape(128, 128)
    c1 = np.corrcoef(zc[:, :-1].ravel(), zc[:, 1:].ravel())[0, 1]
    assert c1 > 0.9
    ks2 = M.correlated_log_ks(p, t, 128, mean_ks=1e-3, sigma=0.5, length=0.05, seed=0)
    assert np.array_equal(ks, ks2)


def test_config_roundtrip_and_reference_defaults():
    from jax_wspm_amd.config import SimulationConfig, create_config_from_dict
    cfg = SimulationConfig()
    d = cfg.to_dict()
    assert d["solver"]["tol"] == 1e-10 and d["solver"]["maxiter"] == 100 and d["solver"]["restart"] == 30
    assert abs(d["van_genuchten"]["m"] - 0.358974358974359) < 1e-15   # frozen 1 - 1/1.56
    assert d["time"]["m_it"] == 5 and d["time"]["M_it"] == 7
    cfg2 = create_config_from_dict(json.loads(json.dumps(d)))
    assert cfg2.to_dict() == d


def test_adapt_time_step_rule():
    from jax_wspm_amd.richards import adapt_time_step
    assert adapt_time_step(1.0, 5) == 2.0 and adapt_time_step(1.0, 6) == 1.0
    assert adapt_time_step(1.0, 7) == 1.0 and adapt_time_step(1.0, 8) == 0.5


def test_loader_validation(tmp_path):
    from jax_wspm_amd import mesh as M
    with pytest.raises(FileNotFoundError):
        M.load_and_validate_mesh("25", str(tmp_path), "Test1")
    (tmp_path / "richards").mkdir()
    np.savetxt(tmp_path / "richards" / "p3.csv", np.array([[0, 0], [1, 0], [0, 1.0]]), delimiter=",")
    np.savetxt(tmp_path / "richards" / "t3.csv", np.array([[1, 2, 4]]), delimiter=",", fmt="%d")
    with pytest.raises(ValueError):
        M.load_and_validate_mesh("3", str(tmp_path), "Test1")
    np.savetxt(tmp_path / "richards" / "t3.csv", np.array([[1, 2, 3]]), delimiter=",", fmt="%d")
    p, t, info = M.load_and_validate_mesh("3", str(tmp_path), "Test1")
    assert t.tolist() == [[0, 1, 2]] and info["num_points"] == 3


def test_product_tracy_matches_oracle():
    from jax_wspm_amd import tracy as pt
    from oracle import tracy as ot
    p, _ = omesh.structured_tri(25, 15.24)
    eps0, d = np.exp(0.164 * -15.24), 0.164 * 0.3 / 0.1
    for tn in (1, 2, 3):
        for t in (0.02, 0.19, 0.7):
            a = ot.exact_solution(p[:, 0], p[:, 1], t, tn, 15.24, 0.164, eps0, d)
            b = pt.exact_solution(p[:, 0], p[:, 1], t, tn, 15.24, 0.164, eps0, d)
            assert np.allclose(a[0], b[0], rtol=1e-13, atol=1e-12) and np.allclose(a[1], b[1], rtol=1e-13)


def test_cli_flags_and_defaults_match_reference():
    from jax_wspm_amd.cli import build_config, parse_args
    a = parse_args([])
    # src/cli.py:59-82 defaults
    assert (a.mesh_size, a.test_case, a.solver, a.preconditioner, a.dt, a.tmax, a.output_dir,
            a.save_frequency) == ("25", "Test1", "gmres", "none", 1e-6, 1.0, "results", 5)
    cfg = build_config(parse_args(["--test-case", "Test3D", "--solver", "bicgstab", "--dt", "1e-4",
                                   "--tmax", "0.1"]))
    d = cfg.to_dict()
    assert d["solver"]["solver_type"] == "bicgstab" and d["time"]["dt_init"] == 1e-4 and d["test_case"] == "Test3D"
    with pytest.raises(SystemExit):
        parse_args(["--solver", "lu"])
