"""Assembly of the complete host-side field set of a model configuration.

``assemble`` runs the module classes over a parameter source and returns the
fields every kernel reads or writes, in the reference's network order, with
the feature flags of cmf_fields (the reference's HAS_* constants,
cmfgpu/phys/specs.py:28-39, resolved from the opened modules).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .modules import (AdaptiveTimeModule, BaseModule, BifurcationModule, LeveeModule, LogModule,
                      ReservoirModule)

KNOWN_MODULES = ("base", "adaptive_time", "bifurcation", "levee", "reservoir", "log")
# the hpfloat fields (modules/base.py:197-216,325-410, bifurcation.py, reservoir.py)
STORAGE_FIELDS = ("river_storage", "flood_storage", "protected_storage", "outgoing_storage", "total_storage",
                  "river_inflow", "flood_inflow", "global_bifurcation_outflow", "reservoir_total_inflow")


@dataclass
class HostState:
    fields: dict
    flags: int
    num_flood_levels: int
    num_bifurcation_levels: int = 0
    log_buffer_size: int = 0
    gravity: float = 9.8
    adaptive_time_factor: float = 0.7
    min_kinematic_slope: float = 1e-5
    real: type = np.float32
    modules: dict = field(default_factory=dict)
    sto: type = np.float64

    @property
    def num_catchments(self) -> int:
        return int(self.fields["downstream_idx"].shape[0])


def real_dtype(precision) -> type:
    if precision in ("float32", np.float32, "f32", 0):
        return np.float32
    if precision in ("float64", np.float64, "f64", 1):
        return np.float64
    raise ValueError(f"unknown precision {precision!r}")


def assemble(src: dict, opened_modules=("base", "adaptive_time"), *, precision="float32",
             log_buffer_size: int = 0, gravity: float = 9.8, adaptive_time_factor: float = 0.7,
             min_kinematic_slope: float = 1e-5, cold_start: bool = True,
             precise_surfaces: bool = False, mixed_precision: bool = True) -> HostState:
    """``precise_surfaces``: in the FP32 profile form water surfaces and their
    differences in double (the PRECISE_SURFACE flag); no effect in FP64.
    ``mixed_precision``: storages and accumulators (the hpfloat fields) in
    FP64 under ``precision="float32"``; with False they are FP32 as well, the
    reference's float32 profile without mixed precision
    (cmfgpu/phys/triton/utils.py:18-30).  The water-balance sums stay FP64."""
    opened = set(opened_modules) | {"base"}
    unknown = opened - set(KNOWN_MODULES)
    if unknown:
        raise ValueError(f"unsupported modules: {sorted(unknown)} (inflow and sea_level are not part of this path)")
    real = real_dtype(precision)
    mods = {}
    base = BaseModule(src, opened, real, gravity=gravity, min_kinematic_slope=min_kinematic_slope)
    mods["base"] = base
    fields = dict(base.fields)
    flags = _lib.PRECISE_SURFACE if precise_surfaces else 0
    nb_levels = 0
    if "adaptive_time" in opened:
        at = AdaptiveTimeModule(adaptive_time_factor)
        mods["adaptive_time"] = at
        flags |= _lib.HAS_ADAPTIVE_TIME
    bif = None
    if "bifurcation" in opened:
        bif = BifurcationModule(src, base.num_catchments, real)
        mods["bifurcation"] = bif
        fields.update(bif.fields)
        nb_levels = bif.num_bifurcation_levels
        flags |= _lib.HAS_BIFURCATION
    if "levee" in opened:
        lev = LeveeModule(src, fields, real)
        mods["levee"] = lev
        fields.update(lev.fields)
        flags |= _lib.HAS_LEVEE
    if "reservoir" in opened:
        res = ReservoirModule(src, fields, real)
        mods["reservoir"] = res
        fields.update(res.fields)
        flags |= _lib.HAS_RESERVOIR
        if cold_start:
            res.initialize_state(fields, fields if bif is not None else None)
    if "log" in opened:
        log = LogModule(log_buffer_size)
        mods["log"] = log
        fields.update(log.fields)
        flags |= _lib.HAS_LOG
    if base.has_total_storage:
        flags |= _lib.HAS_TOTAL_STORAGE
        fields["total_storage"] = fields["river_storage"] + fields["flood_storage"] + fields["protected_storage"]
    if base.has_water_surface:
        flags |= _lib.HAS_WATER_SURFACE
    if base.has_protected_water_surface:
        flags |= _lib.HAS_PROTECTED_WATER_SURFACE
    sto = np.float64
    if not mixed_precision and real == np.float32:
        sto = np.float32
        for name in STORAGE_FIELDS:
            if name in fields:
                fields[name] = np.ascontiguousarray(fields[name], dtype=np.float32)
    return HostState(fields=fields, flags=flags, num_flood_levels=base.num_flood_levels,
                     num_bifurcation_levels=nb_levels, log_buffer_size=log_buffer_size if "log" in opened else 0,
                     gravity=gravity, adaptive_time_factor=adaptive_time_factor,
                     min_kinematic_slope=min_kinematic_slope, real=real, modules=mods, sto=sto)
