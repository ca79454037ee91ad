"""BifurcationModule (cmfgpu/modules/bifurcation.py:64-214)."""
from __future__ import annotations

import numpy as np

PATH_LEVEL_PARAMS = ("bifurcation_manning", "bifurcation_width", "bifurcation_elevation")


class BifurcationModule:
    module_name = "bifurcation"

    def __init__(self, src: dict, num_catchments: int, real=np.float32):
        ci = np.ascontiguousarray(src["bifurcation_catchment_idx"], dtype=np.int32)
        di = np.ascontiguousarray(src["bifurcation_downstream_idx"], dtype=np.int32)
        if ci.shape != di.shape or ci.ndim != 1:
            raise ValueError("bifurcation_catchment_idx and bifurcation_downstream_idx must be 1-D of equal length")
        if ci.size and (ci.min() < 0 or di.min() < 0 or ci.max() >= num_catchments or di.max() >= num_catchments):
            raise ValueError("bifurcation catchment index out of range")
        P = ci.shape[0]
        width = np.ascontiguousarray(src["bifurcation_width"], dtype=real)
        if width.ndim != 2 or width.shape[0] != P:
            raise ValueError("bifurcation_width must be [num_bifurcation_paths, num_bifurcation_levels]")
        Lb = width.shape[1]
        f = {
            "bifurcation_catchment_idx": ci,
            "bifurcation_downstream_idx": di,
            "bifurcation_width": width,
            "bifurcation_manning": np.ascontiguousarray(src.get("bifurcation_manning", np.full((P, Lb), 0.03)), dtype=real),
            "bifurcation_length": np.ascontiguousarray(src["bifurcation_length"], dtype=real),
            "bifurcation_elevation": np.ascontiguousarray(src["bifurcation_elevation"], dtype=real).copy(),
            "bifurcation_outflow": np.ascontiguousarray(src.get("bifurcation_outflow", np.zeros((P, Lb))), dtype=real).copy(),
            "bifurcation_cross_section_depth": np.ascontiguousarray(
                src.get("bifurcation_cross_section_depth", np.zeros((P, Lb))), dtype=real).copy(),
            "global_bifurcation_outflow": np.zeros(num_catchments, np.float64),
            "limit_rate": np.zeros(num_catchments, real),
        }
        for k in ("bifurcation_path_id", "bifurcation_catchment_id", "bifurcation_downstream_id"):
            if k in src:
                f[k] = np.asarray(src[k], dtype=np.int64)
        self.num_bifurcation_paths = P
        self.num_bifurcation_levels = Lb
        self.fields = f
