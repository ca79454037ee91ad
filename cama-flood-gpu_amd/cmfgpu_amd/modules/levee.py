"""LeveeModule (cmfgpu/modules/levee.py:73-244)."""
from __future__ import annotations

import numpy as np


def interp_lookup(table: np.ndarray, rows_idx: np.ndarray, position: np.ndarray) -> np.ndarray:
    """LeveeModule._interp_lookup (levee.py:140-207) for a shared table.

    Linear interpolation at `position` in [0, L] where virtual level 0 is
    depth 0 and level k > 0 is table[:, k - 1]; arithmetic in the table dtype.
    """
    L = table.shape[-1]
    rows = table[rows_idx]
    dt = table.dtype.type
    position = np.clip(position.astype(table.dtype), dt(0.0), dt(L))
    lower = np.floor(position).astype(np.int32)
    upper = lower + 1
    frac = (position - lower.astype(table.dtype)).astype(table.dtype)
    lower_val = np.take_along_axis(rows, np.clip(lower - 1, 0, None)[:, None], axis=1)[:, 0]
    lower_val = np.where(lower == 0, dt(0.0), lower_val).astype(table.dtype)
    upper_val = np.take_along_axis(rows, np.minimum(upper - 1, L - 1)[:, None], axis=1)[:, 0]
    return (lower_val + frac * (upper_val - lower_val)).astype(table.dtype)


class LeveeModule:
    module_name = "levee"

    def __init__(self, src: dict, base_fields: dict, real=np.float32):
        n = base_fields["downstream_idx"].shape[0]
        lidx = np.ascontiguousarray(src["levee_catchment_idx"], dtype=np.int32)
        if lidx.size and (lidx.min() < 0 or lidx.max() >= n):
            raise ValueError("levee catchment index out of range")
        if np.unique(lidx).size != lidx.size:
            raise ValueError("levee_catchment_idx must be unique (one levee per catchment)")
        frac = np.ascontiguousarray(src["levee_fraction"], dtype=real)
        crown = np.ascontiguousarray(src["levee_crown_height"], dtype=real)
        self.validate_levee_fraction(frac)
        table = base_fields["flood_depth_table"]
        L = table.shape[-1]
        # levee.py:209-217: position = levee_fraction * num_flood_levels
        base = interp_lookup(table, lidx, (frac * real(L)).astype(real))
        is_levee = np.zeros(n, np.bool_)
        is_levee[lidx] = True
        self.fields = {
            "levee_catchment_idx": lidx,
            "levee_fraction": frac,
            "levee_crown_height": crown,
            "levee_base_height": np.ascontiguousarray(base, dtype=real),
            "is_levee": is_levee,
        }
        if "levee_catchment_id" in src:
            self.fields["levee_catchment_id"] = np.asarray(src["levee_catchment_id"], dtype=np.int64)
        self.num_levees = int(lidx.shape[0])
        self.validate_linked_state()

    @staticmethod
    def validate_levee_fraction(frac: np.ndarray) -> None:
        """levee.py:222-231."""
        invalid = ~np.isfinite(frac) | (frac < 0) | (frac >= 1)
        if invalid.any():
            raise ValueError("levee_fraction must be finite and lie within [0, 1)")

    def validate_linked_state(self) -> None:
        """levee.py:233-244."""
        crown = self.fields["levee_crown_height"]
        base = self.fields["levee_base_height"]
        invalid = ~np.isfinite(crown) | ~np.isfinite(base) | (base >= crown)
        num = int(invalid.sum())
        if num > 0:
            raise ValueError(
                f"{num} levees have non-finite heights or a base height greater than or equal to the crown height")
