"""ReservoirModule (cmfgpu/modules/reservoir.py:64-265)."""
from __future__ import annotations

import numpy as np

RESERVOIR_PARAMS = ("reservoir_capacity", "conservation_volume", "emergency_volume",
                    "normal_outflow", "flood_control_outflow", "reservoir_area")


class ReservoirModule:
    module_name = "reservoir"

    def __init__(self, src: dict, base_fields: dict, real=np.float32):
        n = base_fields["downstream_idx"].shape[0]
        ridx = np.ascontiguousarray(src["reservoir_catchment_idx"], dtype=np.int32)
        if ridx.size and (ridx.min() < 0 or ridx.max() >= n):
            raise ValueError("reservoir catchment index out of range")
        if np.unique(ridx).size != ridx.size:
            raise ValueError("reservoir_catchment_idx must be unique (one reservoir per catchment)")
        f = {"reservoir_catchment_idx": ridx}
        for k in RESERVOIR_PARAMS:
            f[k] = np.ascontiguousarray(src[k], dtype=real)
        if "reservoir_catchment_id" in src:
            f["reservoir_catchment_id"] = np.asarray(src["reservoir_catchment_id"], dtype=np.int64)
        self.real = real
        self.fields = f
        self.num_reservoirs = int(ridx.shape[0])
        self.validate_reservoir_volumes()
        # derived parameters, reservoir.py:174-206, in the parameter dtype
        cons, emer = f["conservation_volume"], f["emergency_volume"]
        qn, qf = f["normal_outflow"], f["flood_control_outflow"]
        r = real
        flood_volume = ((emer - cons) / r(0.95)).astype(real)
        f["flood_volume"] = flood_volume
        f["adjustment_volume"] = (cons + flood_volume * r(0.1)).astype(real)
        seconds_per_year = 365.0 * 24 * 60 * 60
        seconds_per_180_days = 180.0 * 24 * 60 * 60
        vyr = (qn * r(seconds_per_year)).astype(real)
        qsto = ((cons * r(0.7) + vyr / r(4.0)) / r(seconds_per_180_days)).astype(real)
        f["effective_normal_outflow"] = (np.minimum(qn, qsto) * r(1.5)).astype(real)
        f["adjustment_outflow"] = ((f["effective_normal_outflow"] + qf) * r(0.5)).astype(real)
        # masks, reservoir.py:129-163
        dn = base_fields["downstream_idx"]
        is_res = np.zeros(n, np.bool_)
        is_res[ridx] = True
        idx = np.arange(n)
        not_mouth = dn != idx
        upstream = (~is_res) & not_mouth & is_res[dn]
        f["is_reservoir"] = is_res
        f["is_dam_related"] = is_res | upstream
        f["is_dam_upstream"] = f["is_dam_related"] & ~is_res
        f["reservoir_total_inflow"] = np.zeros(n, np.float64)

    def validate_reservoir_volumes(self) -> None:
        """reservoir.py:259-265."""
        f = self.fields
        if np.any(f["emergency_volume"] > f["reservoir_capacity"]):
            raise ValueError("Emergency volume exceeds reservoir capacity")
        if np.any(f["conservation_volume"] > f["emergency_volume"]):
            raise ValueError("Conservation volume exceeds emergency volume")

    def initialize_state(self, base_fields: dict, bif_fields: dict | None) -> tuple[int, int]:
        """Cold start (reservoir.py:208-254): dam storage and dam-related paths."""
        idx = self.fields["reservoir_catchment_idx"]
        river = base_fields["river_storage"][..., idx]
        flood = base_fields["flood_storage"][..., idx]
        cons = self.fields["conservation_volume"].astype(river.dtype)
        update = (river + flood) < cons
        fixed = int(update.sum())
        if fixed:
            base_fields["river_storage"][..., idx] = np.where(update, cons, river)
            base_fields["flood_storage"][..., idx] = np.where(update, np.zeros_like(flood), flood)
        masked = 0
        if bif_fields is not None:
            rel = self.fields["is_dam_related"]
            m = rel[bif_fields["bifurcation_catchment_idx"]] | rel[bif_fields["bifurcation_downstream_idx"]]
            bif_fields["bifurcation_elevation"][..., m, :] = 1.0e20
            masked = int(m.sum())
        return fixed, masked
