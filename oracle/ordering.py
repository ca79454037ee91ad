"""Pure-Python restatement of the reference's network-ordering routines.

TEST INFRASTRUCTURE ONLY (the checker for cmfgpu_amd.network).  Small inputs.

  trace_outlets          cmfgpu/params/utils.py:20-80
  topological_sort       cmfgpu/params/utils.py:83-138
  reorder_by_basin_size  cmfgpu/params/utils.py:483-503 (equal-size basins in
                         ascending label order: the reference's np.argsort
                         default kind leaves that tie order unspecified)
  level_order            breadth-first levels from the mouths (the device layout)
  basin_level_order      the same levels laid out basin by basin
  keyed_level_order      the forest-wide levels, sibling groups by work class
"""
from __future__ import annotations

from collections import deque

import numpy as np


def trace_outlets(catchment_id, downstream_id):
    down = {int(c): int(d) for c, d in zip(catchment_id, downstream_id)}
    out = []
    for c in catchment_id:
        cur = int(c)
        seen = 0
        while cur in down and down[cur] >= 0:
            nxt = down[cur]
            if nxt not in down:      # points at a non-cell: stop there
                cur = nxt
                break
            cur = nxt
            seen += 1
            if seen > len(down) + 1:
                raise ValueError("ring")
        out.append(cur)
    return np.array(out, dtype=np.int64)


def topological_sort(catchment_id, downstream_id):
    idx = {int(c): i for i, c in enumerate(catchment_id)}
    n = len(catchment_id)
    indeg = [0] * n
    for d in downstream_id:
        j = idx.get(int(d), -1)
        if j >= 0:
            indeg[j] += 1
    q = deque(i for i in range(n) if indeg[i] == 0)
    out = []
    while q:
        u = q.popleft()
        out.append(u)
        j = idx.get(int(downstream_id[u]), -1)
        if j >= 0:
            indeg[j] -= 1
            if indeg[j] == 0:
                q.append(j)
    if len(out) != n:
        raise ValueError("Rings exist and topological sorting is not possible")
    return np.array(out, dtype=np.int32)


def reorder_by_basin_size(topo_idx, basin_id):
    topo_basin = [int(basin_id[t]) for t in topo_idx]
    labels = sorted(set(topo_basin))
    counts = {b: topo_basin.count(b) for b in labels}
    ranked = sorted(labels, key=lambda b: (-counts[b], b))
    rank = {b: r for r, b in enumerate(ranked)}
    order = sorted(range(len(topo_idx)), key=lambda k: (rank[topo_basin[k]], k))
    return (np.array([int(topo_idx[k]) for k in order], dtype=np.int64),
            np.array([counts[b] for b in ranked], dtype=np.int64))


def level_order(downstream_idx):
    n = len(downstream_idx)
    children = [[] for _ in range(n)]
    for i, d in enumerate(downstream_idx):
        if d != i:
            children[d].append(i)
    levels = [[i for i in range(n) if downstream_idx[i] == i]]
    while True:
        nxt = [c for u in levels[-1] for c in children[u]]
        if not nxt:
            break
        levels.append(nxt)
    perm = [i for lev in reversed(levels) for i in lev]
    inv = [0] * n
    for k, i in enumerate(perm):
        inv[i] = k
    up_ptr = [0]
    for k in range(n):
        up_ptr.append(up_ptr[-1] + len(children[perm[k]]))
    dn = [inv[downstream_idx[perm[k]]] for k in range(n)]
    return (np.array(perm, np.int32), np.array(inv, np.int32), np.array(up_ptr, np.int32),
            np.array(dn, np.int32), len(levels))


def basin_level_order(downstream_idx):
    """Breadth-first levels inside each basin, basins by ascending mouth index,
    deepest level first inside a basin (the layout of cmf_basin_level_order)."""
    n = len(downstream_idx)
    children = [[] for _ in range(n)]
    for i, d in enumerate(downstream_idx):
        if d != i:
            children[d].append(i)
    perm, basin_ptr = [], [0]
    for m in range(n):
        if downstream_idx[m] != m:
            continue
        levels = [[m]]
        while True:
            nxt = [c for u in levels[-1] for c in children[u]]
            if not nxt:
                break
            levels.append(nxt)
        perm += [i for lev in reversed(levels) for i in lev]
        basin_ptr.append(len(perm))
    inv = [0] * n
    for k, i in enumerate(perm):
        inv[i] = k
    dn = [inv[downstream_idx[perm[k]]] for k in range(n)]
    # a mouth's range ends with the mouth itself when the mouth comes next
    up_ptr = [0]
    for k in range(n):
        end = up_ptr[-1] + len(children[perm[k]])
        if dn[k] == k and end == k:
            end += 1
        up_ptr.append(end)
    return (np.array(perm, np.int32), np.array(inv, np.int32), np.array(up_ptr, np.int32),
            np.array(dn, np.int32), np.array(basin_ptr, np.int64))


_CLASS_LEAD = (1, 3, 2, 0)


def _class_rank(c):
    return _CLASS_LEAD.index(c) if c in _CLASS_LEAD else c


def keyed_level_order(downstream_idx, key=None):
    """Forest-wide breadth-first levels (the layout of cmf_keyed_level_order):
    level 0 lists the mouths (each its own group) in ascending index, level
    d + 1 the children of level d's catchments in that level's order, one group
    per parent (children ascending).  A group's class is the OR of its
    members' keys; each level lists its groups by class rank (1, 3, 2, 0, then
    the other values ascending), ties in formation order, reversed on odd
    depths; levels are emitted deepest first."""
    n = len(downstream_idx)
    key = [0] * n if key is None else [int(k) for k in key]
    children = [[] for _ in range(n)]
    for i, d in enumerate(downstream_idx):
        if d != i:
            children[d].append(i)
    groups = [[m] for m in range(n) if downstream_idx[m] == m]
    levels = []
    depth = 0
    while groups:
        def cls(g):
            c = 0
            for i in g:
                c |= key[i]
            return c
        ranked = sorted(range(len(groups)), key=lambda q: (_class_rank(cls(groups[q])), q))
        if depth % 2 == 1:
            # reverse class order, formation order kept inside a class
            ranked = sorted(range(len(groups)), key=lambda q: (-_class_rank(cls(groups[q])), q))
        level = [i for q in ranked for i in groups[q]]
        levels.append(level)
        groups = [children[u] for u in level if children[u]]
        depth += 1
    perm = [i for lev in reversed(levels) for i in lev]
    inv = [0] * n
    for k, i in enumerate(perm):
        inv[i] = k
    dn = [inv[downstream_idx[perm[k]]] for k in range(n)]
    return np.array(perm, np.int32), np.array(inv, np.int32), np.array(dn, np.int32), len(levels)
