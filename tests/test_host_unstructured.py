"""CPU checks of the unstructured solute meshes and their renumbering.

The fixture tests/golden/pinns_meshes.npz holds the reference's
data/mesh/solute/{p,t}_Pinns_{25,1024}.csv (elements 0-based).  These tests
need no GPU: the mesh loader, the boundary lists of SoluteTest, and
wspm_mesh_rcm (a host-only entry point of libwspm) with the maps of
jax_wspm_amd.reorder, checked with the oracle run on both numberings.
"""
import numpy as np
import pytest

from oracle import picard as opicard
from oracle.assembly import Pattern


def _mesh(golden, tag):
    z = np.load(golden / "pinns_meshes.npz")
    return z[f"points_{tag}"], z[f"elements_{tag}"].astype(np.int64)


def _lib_or_skip():
    from jax_wspm_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libwspm.so not built (run __graft_entry__.build())")


def test_loader_reads_pinns_csv(golden, tmp_path):
    """src/mesh/loader.py:33-36 + :72-73: SoluteTest reads solute/p_Pinns_{size}.csv
    and t_Pinns_{size}.csv and shifts the elements to 0-based."""
    from jax_wspm_amd.mesh import load_and_validate_mesh
    pts, tri = _mesh(golden, "25")
    d = tmp_path / "solute"
    d.mkdir()
    np.savetxt(d / "p_Pinns_25.csv", pts, delimiter=",", fmt="%.17g")
    np.savetxt(d / "t_Pinns_25.csv", tri + 1, delimiter=",", fmt="%d")
    p, t, info = load_and_validate_mesh("25", str(tmp_path), "SoluteTest")
    assert np.array_equal(np.asarray(p), pts) and np.array_equal(np.asarray(t), tri)
    assert info["num_points"] == 3828 and info["num_elements"] == 7426 and info["mesh_size"] == "Pinns"


def test_pinns_rows_exceed_register_kernels(golden):
    """Rows of up to 9 (Pinns_25) and 10 (Pinns_1024) entries: the time step
    takes the LDS / CSR kernels there."""
    for tag, longest in (("25", 9), ("1024", 10)):
        pts, tri = _mesh(golden, tag)
        pat = Pattern(tri, len(pts))
        assert int(np.diff(pat.indptr).max()) == longest, tag


def test_solute_boundary_lists(golden):
    """boundary.py:499-504: the inlet |x| <= 0.05 on y = 2 (11 nodes, joined
    by mesh edges in index order) and the bottom y = 0 (41 nodes, not joined
    in index order: right-hand-side terms only)."""
    from jax_wspm_amd.transport import transport_boundary_nodes
    pts, tri = _mesh(golden, "25")
    top, bottom = transport_boundary_nodes(pts, "SoluteTest")
    edges = {tuple(sorted(e)) for t in tri for e in ((t[0], t[1]), (t[1], t[2]), (t[0], t[2]))}
    assert len(top) == 11 and len(bottom) == 41
    assert all(tuple(sorted(p)) in edges for p in zip(top[:-1], top[1:]))
    assert not all(tuple(sorted(p)) in edges for p in zip(bottom[:-1], bottom[1:]))


def _band(tri, n, perm=None):
    pat = Pattern(tri, n)
    rows = np.repeat(np.arange(n), np.diff(pat.indptr))
    cols = pat.indices
    if perm is not None:
        rows, cols = perm[rows], perm[cols]
    return int((rows - cols).max()), int((cols - rows).max())


@pytest.mark.parametrize("tag,shuffle", [("25", False), ("25", True), ("1024", True)])
def test_rcm_renumbering(golden, tag, shuffle):
    """The reference's solute meshes are already banded in file order
    (Pinns_25: bandwidth 100, RCM 113), so renumbering is opt-in; on the same
    mesh with its nodes numbered at random RCM brings the band back down."""
    _lib_or_skip()
    from jax_wspm_amd.reorder import rcm
    pts, tri = _mesh(golden, tag)
    n = len(pts)
    if shuffle:
        q = np.random.default_rng(5).permutation(n)
        pts, tri = pts[np.argsort(q)], q[tri]
    rn = rcm(tri, n)
    assert np.array_equal(np.sort(rn.node_new), np.arange(n))
    assert np.array_equal(rn.node_old[rn.node_new], np.arange(n))
    assert np.array_equal(np.sort(rn.elem_old), np.arange(len(tri)))
    # the reported bandwidths are those of the node graph before and after
    assert rn.band[:2] == _band(tri, n)
    assert rn.band[2:] == _band(tri, n, rn.node_new)
    if shuffle:
        assert max(rn.band[2:]) * 10 < max(rn.band[:2])
    else:
        assert rn.band == (100, 100, 113, 113)
    # the same triangles, renumbered; elements ordered by their first new vertex
    t2 = rn.elements(tri)
    assert np.array_equal(np.sort(rn.node_old[t2], axis=1), np.sort(tri[rn.elem_old], axis=1))
    assert np.all(np.diff(t2.min(axis=1)) >= 0)
    assert np.array_equal(rn.points(pts)[rn.node_new], pts)
    x = np.random.default_rng(0).normal(size=n)
    assert np.array_equal(rn.nodal_back(rn.nodal(x)), x)


def test_rcm_oracle_trajectory_maps_back(golden):
    """The oracle's time step on the renumbered Pinns_1024 mesh (boundary
    list mapped in its order), mapped back, equals the run in file order to
    rounding: the maps of jax_wspm_amd.reorder carry every input and output."""
    _lib_or_skip()
    from jax_wspm_amd.problems import Problem
    from jax_wspm_amd.richards import _renumbered
    from jax_wspm_amd.reorder import rcm
    from tests._helpers import rel_l2, to_oracle
    pts, tri = _mesh(golden, "1024")
    top = np.nonzero(pts[:, 1] == pts[:, 1].max())[0]
    ks = np.exp(np.random.default_rng(2).normal(np.log(1e-3), 0.5, len(tri)))
    pb = Problem(name="VGPinns", points=pts, elements=tri, soil_model="vg", phi=(0.1, 0.40, 0.06, 2.0, 0.5, 1.0),
                 bc_lists=[(top, np.full(top.shape, -75.0))], h0=np.full(len(pts), -1000.0), dt0=1e-4, nsteps=2,
                 solver="bicgstab", precond="jacobi", Ks_elem=ks)
    rn = rcm(tri, len(pts))
    o1 = opicard.run(to_oracle(pb))
    o2 = opicard.run(to_oracle(_renumbered(pb, rn)))
    assert list(o1["iterations"]) == list(o2["iterations"])
    for k in range(2):
        assert rel_l2(rn.nodal_back(o2["pressure_head"][k]), o1["pressure_head"][k]) <= 1e-12
