"""Network indexing: the reference's ordering pipeline and the device layout.

Host wrappers over the integer routines of libcmfhip (no GPU is needed for
these).  ``finalize_connectivity`` reproduces MERITMap._finalize_connectivity
(cmfgpu/params/merit_map.py:416-456): outlets, Kahn topological order,
basin-size reorder, downstream indices with self-pointing mouths.
``level_order`` is the breadth-first renumbering the device path runs in.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def trace_outlets(catchment_id, downstream_id) -> np.ndarray:
    """cmfgpu/params/utils.py:20-80."""
    cid, did = _i64(catchment_id), _i64(downstream_id)
    out = np.empty(cid.shape[0], dtype=np.int64)
    _lib.check(_lib.lib().cmf_trace_outlets(cid.shape[0], _p(cid, ctypes.c_int64),
                                            _p(did, ctypes.c_int64), _p(out, ctypes.c_int64)),
               "trace_outlets")
    return out


def topological_sort(catchment_id, downstream_id) -> np.ndarray:
    """cmfgpu/params/utils.py:83-138 (ValueError on rings, as :135-136)."""
    cid, did = _i64(catchment_id), _i64(downstream_id)
    out = np.empty(cid.shape[0], dtype=np.int32)
    rc = _lib.lib().cmf_topological_sort(cid.shape[0], _p(cid, ctypes.c_int64),
                                         _p(did, ctypes.c_int64), _p(out, ctypes.c_int32))
    if rc == _lib.CMF_ERR_INVALID:
        raise ValueError(_lib.lib().cmf_last_error().decode())
    _lib.check(rc, "topological_sort")
    return out


def reorder_by_basin_size(topo_idx, basin_id):
    """cmfgpu/params/utils.py:483-503 -> (order int64, basin_sizes int64)."""
    t, b = _i64(topo_idx), _i64(basin_id)
    n = t.shape[0]
    order = np.empty(n, dtype=np.int64)
    sizes = np.empty(max(n, 1), dtype=np.int64)
    nb = ctypes.c_int64(0)
    _lib.check(_lib.lib().cmf_reorder_by_basin_size(n, _p(t, ctypes.c_int64), _p(b, ctypes.c_int64),
                                                    _p(order, ctypes.c_int64), _p(sizes, ctypes.c_int64),
                                                    ctypes.byref(nb)),
               "reorder_by_basin_size")
    return order, sizes[: nb.value].copy()


def find_indices_in(query, ref) -> np.ndarray:
    """Index of each query value in ``ref`` (-1 when absent)."""
    q, r = _i64(query), _i64(ref)
    out = np.empty(q.shape[0], dtype=np.int64)
    _lib.check(_lib.lib().cmf_find_indices_in(q.shape[0], _p(q, ctypes.c_int64), r.shape[0],
                                              _p(r, ctypes.c_int64), _p(out, ctypes.c_int64)),
               "find_indices_in")
    return out


def check_flow_direction(downstream_idx) -> None:
    """cmfgpu/params/merit_map.py:1063-1070 (ValueError when violated)."""
    d = _i32(downstream_idx)
    rc = _lib.lib().cmf_check_flow_direction(d.shape[0], _p(d, ctypes.c_int32))
    if rc == _lib.CMF_ERR_INVALID:
        raise ValueError(_lib.lib().cmf_last_error().decode())
    _lib.check(rc, "check_flow_direction")


def accumulate_downstream(downstream_idx, local) -> np.ndarray:
    """cmfgpu/params/estimate_river_geometry.py:76-104 (upstream-first order)."""
    d = _i32(downstream_idx)
    x = np.ascontiguousarray(local, dtype=np.float64)
    out = np.empty_like(x)
    _lib.check(_lib.lib().cmf_accumulate_downstream(d.shape[0], _p(d, ctypes.c_int32),
                                                    _p(x, ctypes.c_double), _p(out, ctypes.c_double)),
               "accumulate_downstream")
    return out


@dataclass
class LevelOrder:
    """Breadth-first renumbering (see cmf_level_order in cmf_network.cpp).

    perm[k]: network index of internal catchment k; inv = perm^-1;
    up_ptr[k]..up_ptr[k+1]: the internal upstream range of k; dn[k]: internal
    downstream index; level_ptr: level boundaries, deepest level first.
    """

    perm: np.ndarray
    inv: np.ndarray
    up_ptr: np.ndarray
    dn: np.ndarray
    level_ptr: np.ndarray

    @property
    def num_levels(self) -> int:
        return self.level_ptr.shape[0] - 1


def level_order(downstream_idx) -> LevelOrder:
    d = _i32(downstream_idx)
    n = d.shape[0]
    perm = np.empty(n, np.int32)
    inv = np.empty(n, np.int32)
    up_ptr = np.empty(n + 1, np.int32)
    dn = np.empty(n, np.int32)
    nl = ctypes.c_int64(0)
    cap = n + 2
    level_ptr = np.empty(cap, np.int64)
    _lib.check(_lib.lib().cmf_level_order(n, _p(d, ctypes.c_int32), _p(perm, ctypes.c_int32),
                                          _p(inv, ctypes.c_int32), _p(up_ptr, ctypes.c_int32),
                                          _p(dn, ctypes.c_int32), ctypes.byref(nl), cap,
                                          _p(level_ptr, ctypes.c_int64)),
               "level_order")
    return LevelOrder(perm, inv, up_ptr, dn, level_ptr[: nl.value + 1].copy())


@dataclass
class BasinLevelOrder:
    """Breadth-first levels laid out basin by basin (cmf_basin_level_order).

    Same arrays as LevelOrder; basin_ptr[b]..basin_ptr[b+1] is basin b (the tree
    of one mouth, mouths in ascending network index) in internal order.
    """

    perm: np.ndarray
    inv: np.ndarray
    up_ptr: np.ndarray
    dn: np.ndarray
    basin_ptr: np.ndarray

    @property
    def num_basins(self) -> int:
        return self.basin_ptr.shape[0] - 1


def basin_level_order(downstream_idx) -> BasinLevelOrder:
    d = _i32(downstream_idx)
    n = d.shape[0]
    perm = np.empty(n, np.int32)
    inv = np.empty(n, np.int32)
    up_ptr = np.empty(n + 1, np.int32)
    dn = np.empty(n, np.int32)
    nb = ctypes.c_int64(0)
    cap = n + 1
    basin_ptr = np.empty(cap, np.int64)
    _lib.check(_lib.lib().cmf_basin_level_order(n, _p(d, ctypes.c_int32), _p(perm, ctypes.c_int32),
                                                _p(inv, ctypes.c_int32), _p(up_ptr, ctypes.c_int32),
                                                _p(dn, ctypes.c_int32), ctypes.byref(nb), cap,
                                                _p(basin_ptr, ctypes.c_int64)),
               "basin_level_order")
    return BasinLevelOrder(perm, inv, up_ptr, dn, basin_ptr[: nb.value + 1].copy())


def keyed_level_order(downstream_idx, key=None) -> LevelOrder:
    """Breadth-first levels of the whole forest with each level's sibling
    groups arranged by work class (cmf_keyed_level_order): `key` holds one
    small class code per catchment (network order), a group's class is the OR
    of its members'.  `up_ptr` is None: the upstream ranges stay contiguous
    but do not follow their owners' order."""
    d = _i32(downstream_idx)
    n = d.shape[0]
    k = None
    if key is not None:
        k = np.ascontiguousarray(key, dtype=np.uint8)
        if k.shape != (n,):
            raise ValueError(f"key must have one entry per catchment ({n}), got shape {k.shape}")
    perm = np.empty(n, np.int32)
    inv = np.empty(n, np.int32)
    dn = np.empty(n, np.int32)
    nl = ctypes.c_int64(0)
    cap = n + 2
    level_ptr = np.empty(cap, np.int64)
    _lib.check(_lib.lib().cmf_keyed_level_order(n, _p(d, ctypes.c_int32),
                                                _p(k, ctypes.c_uint8) if k is not None else None,
                                                _p(perm, ctypes.c_int32), _p(inv, ctypes.c_int32),
                                                _p(dn, ctypes.c_int32), ctypes.byref(nl), cap,
                                                _p(level_ptr, ctypes.c_int64)),
               "keyed_level_order")
    return LevelOrder(perm, inv, None, dn, level_ptr[: nl.value + 1].copy())


@dataclass
class Connectivity:
    """Result of finalize_connectivity: the reference order of the catchments."""

    order: np.ndarray            # generation index of each ordered catchment
    catchment_id: np.ndarray
    downstream_id: np.ndarray    # mouths point to themselves
    downstream_idx: np.ndarray   # int32, == own index at mouths
    catchment_basin_id: np.ndarray
    basin_sizes: np.ndarray
    root_mouth: np.ndarray


def finalize_connectivity(catchment_id, downstream_id) -> Connectivity:
    """MERITMap._finalize_connectivity, cmfgpu/params/merit_map.py:416-456.

    ``downstream_id`` < 0 marks a mouth on input.
    """
    cid, did = _i64(catchment_id), _i64(downstream_id)
    root_mouth = trace_outlets(cid, did)
    topo = topological_sort(cid, did)
    order, sizes = reorder_by_basin_size(topo, root_mouth)
    cid_o = cid[order]
    did_o = did[order].copy()
    rm_o = root_mouth[order]
    changes = np.concatenate(([True], rm_o[1:] != rm_o[:-1])) if rm_o.size else np.zeros(0, bool)
    basin_id = (np.cumsum(changes) - 1).astype(np.int64)
    didx = find_indices_in(did_o, cid_o)
    mouth = didx < 0
    did_o[mouth] = cid_o[mouth]
    didx[mouth] = np.flatnonzero(mouth)
    didx32 = didx.astype(np.int32)
    check_flow_direction(didx32)
    return Connectivity(order, cid_o, did_o, didx32, basin_id, sizes, rm_o)
