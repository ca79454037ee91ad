"""Backend-neutral kernel ABIs, restated from cmfgpu/phys/specs.py:1-803.

The reference builds these with hydroforge.contracts.KernelSpec (absent
here); this module carries its own minimal KernelSpec with the same names,
parameter order, access modes and optional/feature rules, and the same
construction-time checks (specs.py:57-103): every pointer has exactly one
access mode and every remaining parameter is a known constant or runtime
scalar.
"""
from __future__ import annotations

from dataclasses import dataclass, field

RUNTIME_SCALARS = {
    "gravity": "precision",
    "time_step": "precision",
    "adaptive_time_factor": "precision",
    "num_catchments": "index",
    "num_trials": "index",
    "num_sea_level_boundaries": "index",
    "num_bifurcation_paths": "index",
    "num_reservoirs": "index",
    "num_inflow_gauges": "int32",
    "num_levees": "index",
}
STRUCTURAL_CONSTANTS = {"num_bifurcation_levels": "int32", "num_flood_levels": "int32"}
FEATURE_SOURCES = {
    "HAS_BIFURCATION": ("module_enabled", "bifurcation"),
    "HAS_INFLOW": ("module_enabled", "inflow"),
    "HAS_LEVEE": ("module_enabled", "levee"),
    "HAS_RESERVOIR": ("module_enabled", "reservoir"),
    "HAS_SEA_LEVEL": ("module_enabled", "sea_level"),
    "HAS_TOTAL_STORAGE": ("module_flag", "has_total_storage"),
    "HAS_WATER_SURFACE": ("module_flag", "has_water_surface"),
    "HAS_PROTECTED_WATER_SURFACE": ("module_flag", "has_protected_water_surface"),
}
ACCESS_MODES = ("read", "write", "read_write", "atomic_add", "atomic_max")


@dataclass(frozen=True)
class KernelSpec:
    name: str
    parameters: tuple
    size_key: str
    buffers: dict
    optional_buffers: dict
    compile_time: dict
    feature_sources: dict
    runtime_scalars: dict
    optional_values: dict = field(default_factory=dict)

    def pointer_fields(self):
        """Module field names of the pointer parameters (``*_ptr`` minus the suffix)."""
        return tuple(p[:-4] for p in self.parameters if p.endswith("_ptr"))

    def required_fields(self, enabled_features: set) -> tuple:
        out = []
        for p in self.parameters:
            if not p.endswith("_ptr"):
                continue
            gate = self.optional_buffers.get(p, "required")
            if p in self.optional_buffers and (gate is None or gate not in enabled_features):
                continue
            out.append(p[:-4])
        return tuple(out)


def _spec(name, parameters, size_key, *, optional=None, constants=None, optional_values=None,
          read=(), write=(), read_write=(), atomic_add=(), atomic_max=()):
    pointer_names = {p for p in parameters if p.endswith("_ptr")}
    groups = {"read": set(read), "write": set(write), "read_write": set(read_write),
              "atomic_add": set(atomic_add), "atomic_max": set(atomic_max)}
    unknown = {a: n - pointer_names for a, n in groups.items() if n - pointer_names}
    if unknown:
        raise ValueError(f"{name}: buffer access names are not pointer arguments: {unknown}")
    claimed = {}
    for access, names in groups.items():
        for p in names:
            prev = claimed.setdefault(p, access)
            if prev != access:
                raise ValueError(f"{name}: pointer {p!r} has conflicting access {prev!r} and {access!r}")
    missing = pointer_names - set(claimed)
    if missing:
        raise ValueError(f"{name}: every pointer requires one explicit access mode; missing={sorted(missing)}")
    buffers = {p: claimed[p] for p in pointer_names}
    constants = dict(constants or {})
    constants.update({p: STRUCTURAL_CONSTANTS[p] for p in parameters if p in STRUCTURAL_CONSTANTS})
    scalars = set(parameters) - set(buffers) - set(constants)
    bad = scalars - set(RUNTIME_SCALARS)
    if bad:
        raise ValueError(f"{name}: undefined runtime scalar kind(s): {bad}")
    return KernelSpec(
        name=name, parameters=tuple(parameters), size_key=size_key, buffers=buffers,
        optional_buffers=dict(optional or {}), compile_time=constants,
        feature_sources={p: FEATURE_SOURCES[p] for p in constants if p in FEATURE_SOURCES},
        runtime_scalars={p: RUNTIME_SCALARS[p] for p in scalars},
        optional_values=dict(optional_values or {}),
    )


_BATCH_BASE = ("batched_river_manning", "batched_flood_manning", "batched_river_width",
               "batched_river_length", "batched_river_height", "batched_catchment_elevation",
               "batched_downstream_distance", "batched_sea_surface_elevation")

OUTFLOW = _spec(
    "compute_outflow",
    ("downstream_idx_ptr", "river_inflow_ptr", "river_outflow_ptr", "river_manning_ptr",
     "river_depth_ptr", "river_width_ptr", "river_length_ptr", "river_height_ptr",
     "river_storage_ptr", "flood_inflow_ptr", "flood_outflow_ptr", "flood_manning_ptr",
     "flood_depth_ptr", "protected_depth_ptr", "is_levee_ptr", "catchment_elevation_ptr",
     "downstream_distance_ptr", "flood_storage_ptr", "protected_storage_ptr",
     "river_cross_section_depth_ptr", "flood_cross_section_depth_ptr",
     "flood_cross_section_area_ptr", "global_bifurcation_outflow_ptr", "total_storage_ptr",
     "outgoing_storage_ptr", "water_surface_elevation_ptr",
     "protected_water_surface_elevation_ptr", "gravity", "time_step_ptr", "num_catchments",
     "HAS_BIFURCATION", "HAS_TOTAL_STORAGE", "HAS_WATER_SURFACE", "HAS_PROTECTED_WATER_SURFACE",
     "HAS_LEVEE", "is_dam_upstream_ptr", "HAS_RESERVOIR", "MIN_KINEMATIC_SLOPE",
     "sea_surface_elevation_ptr", "catchment_sea_level_idx_ptr", "HAS_SEA_LEVEL", "num_trials")
    + _BATCH_BASE + ("num_sea_level_boundaries",),
    "num_catchments",
    optional={"global_bifurcation_outflow_ptr": "HAS_BIFURCATION", "total_storage_ptr": None,
              "water_surface_elevation_ptr": None, "protected_water_surface_elevation_ptr": None,
              "is_levee_ptr": "HAS_LEVEE", "is_dam_upstream_ptr": "HAS_RESERVOIR",
              "sea_surface_elevation_ptr": "HAS_SEA_LEVEL", "catchment_sea_level_idx_ptr": "HAS_SEA_LEVEL"},
    constants={**{k: "bool" for k in ("HAS_BIFURCATION", "HAS_TOTAL_STORAGE", "HAS_WATER_SURFACE",
                                      "HAS_PROTECTED_WATER_SURFACE", "HAS_LEVEE", "HAS_RESERVOIR",
                                      "HAS_SEA_LEVEL")},
               "MIN_KINEMATIC_SLOPE": "precision", **{k: "bool" for k in _BATCH_BASE}},
    optional_values={"num_sea_level_boundaries": ("HAS_SEA_LEVEL", 0)},
    read=("downstream_idx_ptr", "river_manning_ptr", "river_depth_ptr", "river_width_ptr",
          "river_length_ptr", "river_height_ptr", "river_storage_ptr", "flood_manning_ptr",
          "flood_depth_ptr", "protected_depth_ptr", "is_levee_ptr", "catchment_elevation_ptr",
          "downstream_distance_ptr", "flood_storage_ptr", "protected_storage_ptr", "time_step_ptr",
          "is_dam_upstream_ptr", "sea_surface_elevation_ptr", "catchment_sea_level_idx_ptr"),
    write=("river_inflow_ptr", "flood_inflow_ptr", "global_bifurcation_outflow_ptr", "total_storage_ptr",
           "water_surface_elevation_ptr", "protected_water_surface_elevation_ptr"),
    read_write=("river_outflow_ptr", "flood_outflow_ptr", "river_cross_section_depth_ptr",
                "flood_cross_section_depth_ptr", "flood_cross_section_area_ptr"),
    atomic_add=("outgoing_storage_ptr",),
)

INFLOW = _spec(
    "compute_inflow",
    ("downstream_idx_ptr", "river_outflow_ptr", "flood_outflow_ptr", "river_storage_ptr",
     "flood_storage_ptr", "outgoing_storage_ptr", "river_inflow_ptr", "flood_inflow_ptr",
     "limit_rate_ptr", "reservoir_total_inflow_ptr", "is_reservoir_ptr", "num_catchments",
     "HAS_BIFURCATION", "HAS_RESERVOIR", "num_trials"),
    "num_catchments",
    optional={"limit_rate_ptr": "HAS_BIFURCATION", "reservoir_total_inflow_ptr": "HAS_RESERVOIR",
              "is_reservoir_ptr": "HAS_RESERVOIR"},
    constants={"HAS_BIFURCATION": "bool", "HAS_RESERVOIR": "bool"},
    read=("downstream_idx_ptr", "river_storage_ptr", "flood_storage_ptr", "outgoing_storage_ptr",
          "is_reservoir_ptr"),
    write=("limit_rate_ptr",),
    read_write=("river_outflow_ptr", "flood_outflow_ptr"),
    atomic_add=("river_inflow_ptr", "flood_inflow_ptr", "reservoir_total_inflow_ptr"),
)

ADAPTIVE_TIME = _spec(
    "compute_adaptive_time_step",
    ("river_depth_ptr", "downstream_distance_ptr", "is_dam_related_ptr", "max_sub_steps_ptr",
     "outer_time_step_ptr", "adaptive_time_factor", "gravity", "num_catchments", "HAS_RESERVOIR",
     "num_trials", "batched_downstream_distance"),
    "num_catchments",
    optional={"is_dam_related_ptr": "HAS_RESERVOIR"},
    constants={"HAS_RESERVOIR": "bool", "batched_downstream_distance": "bool"},
    read=("river_depth_ptr", "downstream_distance_ptr", "is_dam_related_ptr", "outer_time_step_ptr"),
    atomic_max=("max_sub_steps_ptr",),
)

_BIF_PARAMS = ("bifurcation_catchment_idx_ptr", "bifurcation_downstream_idx_ptr", "bifurcation_manning_ptr",
               "bifurcation_outflow_ptr", "bifurcation_width_ptr", "bifurcation_length_ptr",
               "bifurcation_elevation_ptr", "bifurcation_cross_section_depth_ptr",
               "water_surface_elevation_ptr")
_BIF_BATCH = ("batched_bifurcation_manning", "batched_bifurcation_width", "batched_bifurcation_length",
              "batched_bifurcation_elevation")
_BIF_READ = ("bifurcation_catchment_idx_ptr", "bifurcation_downstream_idx_ptr", "bifurcation_manning_ptr",
             "bifurcation_width_ptr", "bifurcation_length_ptr", "bifurcation_elevation_ptr",
             "water_surface_elevation_ptr", "total_storage_ptr", "time_step_ptr")

BIFURCATION_OUTFLOW = _spec(
    "compute_bifurcation_outflow",
    _BIF_PARAMS + ("total_storage_ptr", "outgoing_storage_ptr", "gravity", "time_step_ptr",
                   "num_bifurcation_paths", "num_bifurcation_levels", "num_trials", "num_catchments")
    + _BIF_BATCH,
    "num_bifurcation_paths",
    constants={k: "bool" for k in _BIF_BATCH},
    read=_BIF_READ,
    read_write=("bifurcation_outflow_ptr", "bifurcation_cross_section_depth_ptr"),
    atomic_add=("outgoing_storage_ptr",),
)

BIFURCATION_INFLOW = _spec(
    "compute_bifurcation_inflow",
    ("bifurcation_catchment_idx_ptr", "bifurcation_downstream_idx_ptr", "limit_rate_ptr",
     "bifurcation_outflow_ptr", "global_bifurcation_outflow_ptr", "num_bifurcation_paths",
     "num_bifurcation_levels", "num_trials", "num_catchments"),
    "num_bifurcation_paths",
    read=("bifurcation_catchment_idx_ptr", "bifurcation_downstream_idx_ptr", "limit_rate_ptr"),
    read_write=("bifurcation_outflow_ptr",),
    atomic_add=("global_bifurcation_outflow_ptr",),
)

RESERVOIR_OUTFLOW = _spec(
    "compute_reservoir_outflow",
    ("reservoir_catchment_idx_ptr", "downstream_idx_ptr", "reservoir_total_inflow_ptr",
     "river_outflow_ptr", "flood_outflow_ptr", "river_storage_ptr", "flood_storage_ptr",
     "conservation_volume_ptr", "emergency_volume_ptr", "adjustment_volume_ptr",
     "effective_normal_outflow_ptr", "adjustment_outflow_ptr", "flood_control_outflow_ptr",
     "runoff_ptr", "total_storage_ptr", "outgoing_storage_ptr", "time_step_ptr", "num_reservoirs",
     "num_catchments", "num_trials", "batched_runoff"),
    "num_reservoirs",
    constants={"batched_runoff": "bool"},
    read=("reservoir_catchment_idx_ptr", "downstream_idx_ptr", "river_storage_ptr", "flood_storage_ptr",
          "conservation_volume_ptr", "emergency_volume_ptr", "adjustment_volume_ptr",
          "effective_normal_outflow_ptr", "adjustment_outflow_ptr", "flood_control_outflow_ptr",
          "runoff_ptr", "time_step_ptr"),
    write=("total_storage_ptr",),
    read_write=("river_outflow_ptr", "flood_outflow_ptr", "reservoir_total_inflow_ptr"),
    atomic_add=("outgoing_storage_ptr",),
)

_STAGE_PARAMS = ("river_inflow_ptr", "flood_inflow_ptr", "river_outflow_ptr", "flood_outflow_ptr",
                 "global_bifurcation_outflow_ptr", "runoff_ptr", "inflow_ptr", "catchment_inflow_idx_ptr",
                 "time_step_ptr", "outgoing_storage_ptr", "river_storage_ptr", "flood_storage_ptr",
                 "protected_storage_ptr", "river_depth_ptr", "flood_depth_ptr", "protected_depth_ptr",
                 "flood_fraction_ptr", "river_height_ptr", "flood_depth_table_ptr", "catchment_area_ptr",
                 "river_width_ptr", "river_length_ptr")
_STAGE_READ = ("river_inflow_ptr", "flood_inflow_ptr", "river_outflow_ptr", "flood_outflow_ptr",
               "global_bifurcation_outflow_ptr", "runoff_ptr", "inflow_ptr", "catchment_inflow_idx_ptr",
               "time_step_ptr", "river_height_ptr", "flood_depth_table_ptr", "catchment_area_ptr",
               "river_width_ptr", "river_length_ptr")
_STAGE_WRITE = ("outgoing_storage_ptr", "river_depth_ptr", "flood_depth_ptr", "protected_depth_ptr",
                "flood_fraction_ptr")
_STAGE_RW = ("river_storage_ptr", "flood_storage_ptr", "protected_storage_ptr")
_STAGE_BATCH = ("batched_runoff", "batched_inflow", "batched_river_height", "batched_flood_depth_table",
                "batched_catchment_area", "batched_river_width", "batched_river_length")

FLOOD_STAGE = _spec(
    "compute_flood_stage",
    _STAGE_PARAMS + ("num_catchments", "num_flood_levels", "HAS_BIFURCATION", "HAS_INFLOW", "num_trials")
    + _STAGE_BATCH + ("num_inflow_gauges",),
    "num_catchments",
    optional={"global_bifurcation_outflow_ptr": "HAS_BIFURCATION", "inflow_ptr": "HAS_INFLOW",
              "catchment_inflow_idx_ptr": "HAS_INFLOW"},
    constants={"HAS_BIFURCATION": "bool", "HAS_INFLOW": "bool", **{k: "bool" for k in _STAGE_BATCH}},
    optional_values={"num_inflow_gauges": ("HAS_INFLOW", 0)},
    read=_STAGE_READ, write=_STAGE_WRITE, read_write=_STAGE_RW,
)

LOG_SUM_PTRS = ("total_storage_pre_sum_ptr", "total_storage_next_sum_ptr", "total_storage_new_sum_ptr",
                "total_inflow_sum_ptr", "total_outflow_sum_ptr", "total_storage_stage_sum_ptr",
                "river_storage_sum_ptr", "flood_storage_sum_ptr", "flood_area_sum_ptr",
                "total_inflow_error_sum_ptr", "total_stage_error_sum_ptr")

FLOOD_STAGE_LOG = _spec(
    "compute_flood_stage_log",
    _STAGE_PARAMS + ("is_levee_ptr",) + LOG_SUM_PTRS
    + ("current_step_ptr", "num_catchments", "num_flood_levels", "HAS_BIFURCATION", "HAS_INFLOW", "HAS_LEVEE"),
    "num_catchments",
    optional={"global_bifurcation_outflow_ptr": "HAS_BIFURCATION", "inflow_ptr": "HAS_INFLOW",
              "catchment_inflow_idx_ptr": "HAS_INFLOW", "is_levee_ptr": "HAS_LEVEE"},
    constants={"HAS_BIFURCATION": "bool", "HAS_INFLOW": "bool", "HAS_LEVEE": "bool"},
    read=_STAGE_READ + ("is_levee_ptr", "current_step_ptr"), write=_STAGE_WRITE, read_write=_STAGE_RW,
    atomic_add=LOG_SUM_PTRS,
)

_LEVEE_PARAMS = ("levee_catchment_idx_ptr", "river_storage_ptr", "flood_storage_ptr", "protected_storage_ptr",
                 "river_depth_ptr", "flood_depth_ptr", "protected_depth_ptr", "river_height_ptr",
                 "flood_depth_table_ptr", "catchment_area_ptr", "river_width_ptr", "river_length_ptr",
                 "levee_base_height_ptr", "levee_crown_height_ptr", "levee_fraction_ptr", "flood_fraction_ptr")
_LEVEE_READ = ("levee_catchment_idx_ptr", "river_height_ptr", "flood_depth_table_ptr", "catchment_area_ptr",
               "river_width_ptr", "river_length_ptr", "levee_base_height_ptr", "levee_crown_height_ptr",
               "levee_fraction_ptr")
_LEVEE_RW = ("river_storage_ptr", "flood_storage_ptr", "river_depth_ptr", "flood_depth_ptr", "flood_fraction_ptr")
_LEVEE_BATCH = ("batched_river_length", "batched_river_width", "batched_river_height", "batched_catchment_area",
                "batched_levee_crown_height", "batched_levee_fraction", "batched_levee_base_height",
                "batched_flood_depth_table")

LEVEE_STAGE = _spec(
    "compute_levee_stage",
    _LEVEE_PARAMS + ("num_levees", "num_flood_levels", "num_trials", "num_catchments") + _LEVEE_BATCH,
    "num_levees",
    constants={k: "bool" for k in _LEVEE_BATCH},
    read=_LEVEE_READ, write=("protected_storage_ptr", "protected_depth_ptr"), read_write=_LEVEE_RW,
)

_LEVEE_LOG_PTRS = ("total_storage_stage_sum_ptr", "river_storage_sum_ptr", "flood_storage_sum_ptr",
                   "flood_area_sum_ptr", "total_stage_error_sum_ptr")
LEVEE_STAGE_LOG = _spec(
    "compute_levee_stage_log",
    _LEVEE_PARAMS + _LEVEE_LOG_PTRS + ("current_step_ptr", "num_levees", "num_flood_levels"),
    "num_levees",
    read=_LEVEE_READ + ("current_step_ptr",), write=("protected_storage_ptr", "protected_depth_ptr"),
    read_write=_LEVEE_RW, atomic_add=_LEVEE_LOG_PTRS,
)

LEVEE_BIFURCATION_OUTFLOW = _spec(
    "compute_levee_bifurcation_outflow",
    _BIF_PARAMS + ("protected_water_surface_elevation_ptr", "total_storage_ptr", "outgoing_storage_ptr",
                   "gravity", "time_step_ptr", "num_bifurcation_paths", "num_bifurcation_levels",
                   "num_trials", "num_catchments") + _BIF_BATCH,
    "num_bifurcation_paths",
    constants={k: "bool" for k in _BIF_BATCH},
    read=_BIF_READ + ("protected_water_surface_elevation_ptr",),
    read_write=("bifurcation_outflow_ptr", "bifurcation_cross_section_depth_ptr"),
    atomic_add=("outgoing_storage_ptr",),
)

ALL_SPECS = {s.name: s for s in (OUTFLOW, INFLOW, ADAPTIVE_TIME, BIFURCATION_OUTFLOW, BIFURCATION_INFLOW,
                                 RESERVOIR_OUTFLOW, FLOOD_STAGE, FLOOD_STAGE_LOG, LEVEE_STAGE,
                                 LEVEE_STAGE_LOG, LEVEE_BIFURCATION_OUTFLOW)}
