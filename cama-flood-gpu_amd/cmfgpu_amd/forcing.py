"""Runoff forcing mapped from a grid onto catchments on the device.

The reference turns gridded runoff into catchment runoff with a sparse
catchment x grid-cell matrix of area weights (the mapping .npz written by
hydroforge's ``generate_mapping_table``: ``catchment_ids``, ``sparse_data``,
``sparse_indices``, ``sparse_indptr``, ``matrix_shape``, ``coord_lon``,
``coord_lat``; read at cmfgpu/params/export_bin.py:963-975) and applies it in
hydroforge's ``shard_forcing`` before ``set_inputs``
(scripts/run_daily_bin.py:98-119).  Here the product runs in one HIP pass
(cmf_hip_map_runoff) that writes the model's runoff field directly in the
device layout (rows in the mapping's order, each sum scattered to its
catchment's place).  The unit factor is the caller's: ``scale`` multiplies every
row sum (hydroforge's conversion is not available to pin, so it is a
parameter here).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import torch

from . import _lib


class RunoffMapping:
    """CSR matrix, rows = catchments in network order, columns = grid cells."""

    def __init__(self, indptr, indices, weights, num_cells: int, scale: float = 1.0):
        self.indptr = np.asarray(indptr, dtype=np.int64)
        self.indices = np.asarray(indices, dtype=np.int32)
        self.weights = np.asarray(weights, dtype=np.float32)
        self.num_cells = int(num_cells)
        self.scale = float(scale)
        if self.indptr.ndim != 1 or self.indptr[0] != 0 or np.any(np.diff(self.indptr) < 0):
            raise ValueError("indptr must start at 0 and be non-decreasing")
        if self.indptr[-1] != self.indices.shape[0] or self.indices.shape != self.weights.shape:
            raise ValueError("indices and weights must have indptr[-1] entries")
        if self.indices.size and (self.indices.min() < 0 or self.indices.max() >= self.num_cells):
            raise ValueError("grid index out of range")
        self.model = None

    @property
    def rows(self) -> int:
        return self.indptr.shape[0] - 1

    @classmethod
    def from_npz(cls, path, catchment_ids, scale: float = 1.0) -> "RunoffMapping":
        """Rows of the reference's mapping file for `catchment_ids` (model order);
        catchments the file does not list get no runoff."""
        with np.load(Path(path), allow_pickle=False) as z:
            ids = z["catchment_ids"].astype(np.int64)
            data, idx, ptr = z["sparse_data"], z["sparse_indices"], z["sparse_indptr"]
            shape = tuple(int(x) for x in z["matrix_shape"])
        row_of = {int(c): r for r, c in enumerate(ids)}
        want = np.asarray(catchment_ids, dtype=np.int64)
        counts = np.zeros(want.shape[0], np.int64)
        rows = []
        for k, c in enumerate(want):
            r = row_of.get(int(c))
            rows.append(r)
            if r is not None:
                counts[k] = ptr[r + 1] - ptr[r]
        indptr = np.zeros(want.shape[0] + 1, np.int64)
        indptr[1:] = np.cumsum(counts)
        indices = np.empty(indptr[-1], np.int32)
        weights = np.empty(indptr[-1], np.float32)
        for k, r in enumerate(rows):
            if r is not None:
                indices[indptr[k]:indptr[k + 1]] = idx[ptr[r]:ptr[r + 1]]
                weights[indptr[k]:indptr[k + 1]] = data[ptr[r]:ptr[r + 1]]
        return cls(indptr, indices, weights, shape[1], scale)

    def bind(self, model) -> "RunoffMapping":
        """Upload the matrix in its own (network) row order with each row's
        place in the model's device layout: consecutive rows gather nearby
        grid cells whatever the device layout, and the kernel scatters the
        row sums to their catchments."""
        if model.N != self.rows:
            raise ValueError(f"mapping has {self.rows} rows, model {model.N} catchments")
        if model.T > 1 and "runoff" in model.batched_names:
            raise ValueError("the mapped runoff is shared by all trials")
        dev = model.device
        self.d_indptr = torch.from_numpy(self.indptr).to(dev)
        self.d_indices = torch.from_numpy(self.indices).to(dev)
        self.d_weights = torch.from_numpy(self.weights).to(dev)
        self.d_place = torch.from_numpy(np.asarray(model.inv, dtype=np.int32)).to(dev)
        self.model = model
        return self

    def apply(self, grid) -> None:
        """Map one runoff field (float32, num_cells values, any shape) into the
        bound model's runoff."""
        if self.model is None:
            raise RuntimeError("bind(model) first")
        m = self.model
        if not (isinstance(grid, torch.Tensor) and grid.device == m.device and grid.dtype == torch.float32
                and grid.is_contiguous()):
            g = grid if isinstance(grid, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(grid))
            grid = g.to(device=m.device, dtype=torch.float32).contiguous()
        g = grid.reshape(-1)
        if g.shape[0] != self.num_cells:
            raise ValueError(f"grid has {g.shape[0]} cells, mapping {self.num_cells}")
        out = m.t["runoff"]
        prec = _lib.PRECISION_F32_F64 if out.dtype == torch.float32 else _lib.PRECISION_F64_F64
        _lib.check(_lib.lib().cmf_hip_map_runoff(self.rows, _lib.ptr(self.d_indptr), _lib.ptr(self.d_indices),
                                                 _lib.ptr(self.d_weights), _lib.ptr(g), self.scale,
                                                 _lib.ptr(self.d_place), prec, _lib.ptr(out),
                                                 _lib.current_stream()), "cmf_hip_map_runoff")
        self._keep = g
