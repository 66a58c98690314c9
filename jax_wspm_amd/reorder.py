"""Node and element renumbering of unstructured meshes (``wspm_mesh_rcm``).

The reference runs its unstructured solute meshes (data/mesh/solute, loaded
by src/mesh/loader.py:33-36) in file order.  On the GPU the order decides how
far apart a row's neighbours sit in memory: the SpMV gathers x and the
assembly gathers node records through each row's columns, so a narrow band
keeps those gathers inside few cache lines.  A :class:`Renumbering` holds the
reverse Cuthill-McKee order of the nodes (and the elements sorted by their
first new vertex) and maps arrays between the two numberings.  Boundary node
lists are mapped entry by entry, so they keep the reference's list order (the
edge terms of src/models/boundary.py:108-421 depend on it).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Tuple

import numpy as np

from . import _lib
from ._lib import check, ptr


@dataclass
class Renumbering:
    node_new: np.ndarray     # new index of every original node
    node_old: np.ndarray     # original node at every new index
    elem_old: np.ndarray     # original element at every new position
    band: Tuple[int, int, int, int]   # (kl, ku) of the node graph before, then after

    def points(self, points: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(points)[self.node_old])

    def elements(self, elements: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(self.node_new[np.asarray(elements)[self.elem_old]])

    def nodal(self, a):
        """An array over the original nodes, in the new order."""
        return None if a is None else np.asarray(a)[..., self.node_old]

    def nodal_back(self, a):
        """An array over the new nodes, in the original order."""
        return np.asarray(a)[..., self.node_new]

    def elemental(self, a):
        return None if a is None else np.asarray(a)[..., self.elem_old]

    def node_ids(self, ids) -> np.ndarray:
        """Original node ids -> new ids, list order kept."""
        return self.node_new[np.asarray(ids, dtype=np.int64)]


def rcm(elements: np.ndarray, n_nodes: int) -> Renumbering:
    """Reverse Cuthill-McKee renumbering of the mesh (libwspm, host only)."""
    el = np.ascontiguousarray(np.asarray(elements), dtype=np.int32)
    E, K = el.shape
    lib = _lib.load()
    node_new = np.zeros(n_nodes, dtype=np.int32)
    elem_old = np.zeros(E, dtype=np.int32)
    band = np.zeros(4, dtype=np.int32)
    check(lib.wspm_mesh_rcm(K - 1, int(n_nodes), E, ptr(el), ptr(node_new), ptr(elem_old), ptr(band)), None)
    node_new = node_new.astype(np.int64)
    node_old = np.empty_like(node_new)
    node_old[node_new] = np.arange(n_nodes)
    return Renumbering(node_new=node_new, node_old=node_old, elem_old=elem_old.astype(np.int64),
                       band=tuple(int(b) for b in band))


def renumbering(kind, elements: np.ndarray, n_nodes: int):
    """None (keep the file order) or "rcm"."""
    if kind is None or kind == "none":
        return None
    if kind == "rcm":
        return rcm(elements, n_nodes)
    raise ValueError(f"unknown reordering {kind!r} (none, rcm)")
