"""BaseModule: core catchment fields (cmfgpu/modules/base.py:65-591)."""
from __future__ import annotations

import numpy as np

CATCHMENT_PARAMS = (
    "river_width", "river_length", "river_height", "catchment_elevation",
    "catchment_area", "downstream_distance",
)
# (name, dtype class, default) of init_state fields, base.py:197-271
INIT_STATE = (
    ("river_storage", "hpfloat", 0.0),
    ("flood_storage", "hpfloat", 0.0),
    ("protected_storage", "hpfloat", 0.0),
    ("protected_depth", "float", 0.0),
    ("river_depth", "float", 0.0),
    ("flood_depth", "float", 0.0),
    ("river_outflow", "float", 0.0),
    ("flood_outflow", "float", 0.0),
    ("river_cross_section_depth", "float", 0.0),
    ("flood_cross_section_depth", "float", 0.0),
    ("flood_cross_section_area", "float", 0.0),
    ("flood_fraction", "float", 0.0),
)
# per-trial state: replicated over trials
TRIAL_STATE = tuple(name for name, _, _ in INIT_STATE) + (
    "river_inflow", "flood_inflow", "outgoing_storage", "total_storage",
    "water_surface_elevation", "protected_water_surface_elevation", "runoff",
)
HPFLOAT = {"river_storage", "flood_storage", "protected_storage", "river_inflow",
           "flood_inflow", "outgoing_storage", "total_storage"}


def _count(name, arr, bad) -> int:
    return int(np.count_nonzero(bad(arr)))


class BaseModule:
    module_name = "base"

    def __init__(self, src: dict, opened_modules, real=np.float32, gravity: float = 9.8,
                 min_kinematic_slope: float = 1e-5):
        if not gravity > 0 or not np.isfinite(gravity):
            raise ValueError("gravity must be finite and greater than zero")
        if not min_kinematic_slope > 0 or not np.isfinite(min_kinematic_slope):
            raise ValueError("min_kinematic_slope must be finite and greater than zero")
        self.gravity = float(gravity)
        self.min_kinematic_slope = float(min_kinematic_slope)
        self.opened = set(opened_modules)
        self.real = real
        n = int(np.asarray(src["downstream_idx"]).shape[0])
        self.num_catchments = n
        f = {}
        f["downstream_idx"] = np.ascontiguousarray(src["downstream_idx"], dtype=np.int32)
        for k in ("catchment_id", "downstream_id", "catchment_basin_id"):
            if k in src:
                f[k] = np.asarray(src[k], dtype=np.int64)
        for k in CATCHMENT_PARAMS:
            f[k] = np.ascontiguousarray(src[k], dtype=real)
        f["river_manning"] = np.ascontiguousarray(src.get("river_manning", np.full(n, 0.03)), dtype=real)
        f["flood_manning"] = np.ascontiguousarray(src.get("flood_manning", np.full(n, 0.1)), dtype=real)
        f["flood_depth_table"] = np.ascontiguousarray(src["flood_depth_table"], dtype=real)
        f["runoff"] = np.ascontiguousarray(src.get("runoff", np.zeros(n)), dtype=real)
        for name, kind, default in INIT_STATE:
            dt = np.float64 if kind == "hpfloat" else real
            f[name] = np.ascontiguousarray(src.get(name, np.full(n, default)), dtype=dt).copy()
        # computed state (base.py:325-419)
        f["outgoing_storage"] = np.zeros(n, np.float64)
        f["river_inflow"] = np.zeros(n, np.float64)
        f["flood_inflow"] = np.zeros(n, np.float64)
        if self.has_total_storage:
            f["total_storage"] = (f["river_storage"] + f["flood_storage"] + f["protected_storage"])
        if self.has_water_surface:
            f["water_surface_elevation"] = np.zeros(n, real)
        if self.has_protected_water_surface:
            f["protected_water_surface_elevation"] = np.zeros(n, real)
        self.fields = f
        self.validate_physical_parameters()
        self.validate_flood_depth_table_monotonicity()

    # base.py:297-318
    @property
    def has_total_storage(self) -> bool:
        return bool({"bifurcation", "reservoir"} & self.opened)

    @property
    def has_water_surface(self) -> bool:
        return "bifurcation" in self.opened

    @property
    def has_protected_water_surface(self) -> bool:
        return {"bifurcation", "levee"} <= self.opened

    @property
    def num_flood_levels(self) -> int:
        return int(self.fields["flood_depth_table"].shape[-1])

    def validate_physical_parameters(self) -> None:
        """base.py:462-575."""
        f = self.fields
        positive = {k: f[k] for k in ("river_width", "river_length", "river_height", "catchment_area",
                                      "downstream_distance", "river_manning", "flood_manning")}
        bad = {k: _count(k, v, lambda a: ~np.isfinite(a) | (a <= 0)) for k, v in positive.items()}
        bad = {k: c for k, c in bad.items() if c}
        if bad:
            details = ", ".join(f"{k}={c}" for k, c in bad.items())
            raise ValueError(
                "River geometry, catchment area, downstream distance, and Manning coefficients "
                f"must contain only finite values greater than zero; invalid element counts: {details}")
        finite = {k: f[k] for k in ("catchment_elevation", "flood_depth_table")}
        bad = {k: _count(k, v, lambda a: ~np.isfinite(a)) for k, v in finite.items()}
        bad = {k: c for k, c in bad.items() if c}
        if bad:
            details = ", ".join(f"{k}={c}" for k, c in bad.items())
            raise ValueError(f"Elevation and flood-depth lookup values must be finite; invalid element counts: {details}")
        neg = int(np.count_nonzero(f["flood_depth_table"] < 0))
        if neg:
            raise ValueError(f"flood_depth_table must be non-negative; invalid element count: {neg}")
        state = ("river_storage", "flood_storage", "protected_storage", "protected_depth", "river_depth",
                 "flood_depth", "river_cross_section_depth", "flood_cross_section_depth",
                 "flood_cross_section_area", "flood_fraction")
        bad = {k: _count(k, f[k], lambda a: ~np.isfinite(a) | (a < 0)) for k in state}
        bad = {k: c for k, c in bad.items() if c}
        if bad:
            details = ", ".join(f"{k}={c}" for k, c in bad.items())
            raise ValueError(
                "Initial storage, depth, and cross-section state must contain only finite non-negative "
                f"values; invalid element counts: {details}")
        above = int(np.count_nonzero(f["flood_fraction"] > 1))
        if above:
            raise ValueError(f"Initial flood_fraction must not exceed one; invalid element count: {above}")
        bad = {k: _count(k, f[k], lambda a: ~np.isfinite(a)) for k in ("river_outflow", "flood_outflow")}
        bad = {k: c for k, c in bad.items() if c}
        if bad:
            details = ", ".join(f"{k}={c}" for k, c in bad.items())
            raise ValueError(f"Initial river and flood outflow must contain only finite values; invalid element counts: {details}")

    def validate_flood_depth_table_monotonicity(self) -> None:
        """base.py:577-591."""
        t = self.fields["flood_depth_table"]
        if t.shape[-1] > 1:
            d = np.diff(t, axis=-1)
            bad = int(np.count_nonzero(d < 0))
            if bad:
                raise ValueError(
                    f"flood_depth_table must be monotonically non-decreasing; found {bad} decreasing "
                    f"intervals (minimum diff={float(d.min()):.6f}). Regenerate the parameter file with MERITMap.")
