"""ctypes binding of libcmfhip.so (the C ABI declared in include/cmf_hip.h).

torch is imported before the library is opened: torch ships its own HIP
runtime, and loading it first makes libcmfhip bind to that same runtime, so
device pointers and streams from torch are valid in the library.

There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime libcmfhip binds to)

PACKAGE_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PACKAGE_ROOT / "lib" / "libcmfhip.so"

# ----------------------------------------------------------------- enums --
CMF_OK = 0
CMF_ERR_INVALID = 1
CMF_ERR_HIP = 2
CMF_ERR_NOMEM = 3
CMF_ERR_UNSUPPORTED = 4
CMF_ERR_INTERNAL = 5

HAS_BIFURCATION = 1 << 0
HAS_LEVEE = 1 << 1
HAS_RESERVOIR = 1 << 2
HAS_LOG = 1 << 3
HAS_TOTAL_STORAGE = 1 << 4
HAS_WATER_SURFACE = 1 << 5
HAS_PROTECTED_WATER_SURFACE = 1 << 6
HAS_ADAPTIVE_TIME = 1 << 7
PER_TRIAL_DT = 1 << 8
PRECISE_SURFACE = 1 << 9

BATCHED_BITS = {
    "river_manning": 1 << 0,
    "flood_manning": 1 << 1,
    "river_width": 1 << 2,
    "river_length": 1 << 3,
    "river_height": 1 << 4,
    "catchment_elevation": 1 << 5,
    "downstream_distance": 1 << 6,
    "runoff": 1 << 7,
    "flood_depth_table": 1 << 8,
    "catchment_area": 1 << 9,
    "levee_crown_height": 1 << 10,
    "levee_fraction": 1 << 11,
    "levee_base_height": 1 << 12,
    "bifurcation_manning": 1 << 13,
    "bifurcation_width": 1 << 14,
    "bifurcation_length": 1 << 15,
    "bifurcation_elevation": 1 << 16,
}

PRECISION_F32_F64 = 0
PRECISION_F64_F64 = 1
PRECISION_F32_F32 = 2

SCHEDULES = {"three_pass": 0, "block_local": 1, "limiter_pass": 2}

LOG_FIELDS = (
    "total_storage_pre_sum",
    "total_storage_next_sum",
    "total_storage_new_sum",
    "total_inflow_error_sum",
    "total_inflow_sum",
    "total_outflow_sum",
    "total_storage_stage_sum",
    "total_stage_error_sum",
    "river_storage_sum",
    "flood_storage_sum",
    "flood_area_sum",
)

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64

# Pointer members of cmf_fields in declaration order (after the scalar
# shared-state block).  Names are the reference's module field names.
POINTER_FIELDS = (
    "time_step", "outer_time_step", "current_step", "max_sub_steps",
    "downstream_idx",
    "river_inflow", "flood_inflow", "river_outflow", "flood_outflow",
    "river_manning", "flood_manning", "river_depth", "flood_depth", "protected_depth",
    "river_width", "river_length", "river_height", "catchment_elevation",
    "catchment_area", "downstream_distance", "river_storage", "flood_storage",
    "protected_storage", "river_cross_section_depth", "flood_cross_section_depth",
    "flood_cross_section_area", "flood_fraction", "runoff", "flood_depth_table",
    "outgoing_storage", "total_storage", "water_surface_elevation",
    "protected_water_surface_elevation", "is_levee",
    "bifurcation_catchment_idx", "bifurcation_downstream_idx", "bifurcation_manning",
    "bifurcation_outflow", "bifurcation_width", "bifurcation_length",
    "bifurcation_elevation", "bifurcation_cross_section_depth",
    "global_bifurcation_outflow", "limit_rate",
    "levee_catchment_idx", "levee_base_height", "levee_crown_height", "levee_fraction",
    "reservoir_catchment_idx", "conservation_volume", "emergency_volume",
    "adjustment_volume", "effective_normal_outflow", "adjustment_outflow",
    "flood_control_outflow", "reservoir_total_inflow", "is_reservoir",
    "is_dam_upstream", "is_dam_related", "levee_index",
)


class CmfFields(ctypes.Structure):
    """Mirror of struct cmf_fields."""

    _fields_ = [
        ("num_catchments", _i64),
        ("num_trials", _i64),
        ("num_flood_levels", _i64),
        ("num_bifurcation_paths", _i64),
        ("num_bifurcation_levels", _i64),
        ("num_levees", _i64),
        ("num_reservoirs", _i64),
        ("log_buffer_size", _i64),
        ("trial_stride", _i64),
        ("catchment_stride", _i64),
        ("path_trial_stride", _i64),
        ("path_stride", _i64),
        ("precision", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("batched", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("gravity", ctypes.c_double),
        ("adaptive_time_factor", ctypes.c_double),
        ("min_kinematic_slope", ctypes.c_double),
    ] + [(name, _vp) for name in POINTER_FIELDS] + [
        ("log_sums", _vp * len(LOG_FIELDS)),
    ]


class CmfError(RuntimeError):
    """Raised when a libcmfhip entry point returns a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


_LIB = None


def _declare(lib):
    p = ctypes.POINTER
    i32p = p(ctypes.c_int32)
    i64p = p(ctypes.c_int64)
    f64p = p(ctypes.c_double)
    sig = {
        "cmf_last_error": (ctypes.c_char_p, []),
        "cmf_version": (ctypes.c_int, []),
        "cmf_fields_size": (_i64, []),
        "cmf_trace_outlets": (ctypes.c_int, [_i64, i64p, i64p, i64p]),
        "cmf_topological_sort": (ctypes.c_int, [_i64, i64p, i64p, i32p]),
        "cmf_reorder_by_basin_size": (ctypes.c_int, [_i64, i64p, i64p, i64p, i64p, i64p]),
        "cmf_find_indices_in": (ctypes.c_int, [_i64, i64p, _i64, i64p, i64p]),
        "cmf_check_flow_direction": (ctypes.c_int, [_i64, i32p]),
        "cmf_accumulate_downstream": (ctypes.c_int, [_i64, i32p, f64p, f64p]),
        "cmf_level_order": (ctypes.c_int, [_i64, i32p, i32p, i32p, i32p, i32p, i64p, _i64, i64p]),
        "cmf_basin_level_order": (ctypes.c_int, [_i64, i32p, i32p, i32p, i32p, i32p, i64p, _i64, i64p]),
        "cmf_keyed_level_order": (ctypes.c_int, [_i64, i32p, p(ctypes.c_uint8), i32p, i32p, i32p, i64p, _i64, i64p]),
        "cmf_network_create": (ctypes.c_int, [_i64, i32p, _i64, i32p, i32p, _i64, p(_vp)]),
        "cmf_network_create_grouped": (ctypes.c_int, [_i64, i32p, _i64, i32p, i32p, _i64, _i64, i64p, p(_vp)]),
        "cmf_network_num_groups": (_i64, [_vp]),
        "cmf_network_set_block_sizes": (ctypes.c_int, [_vp, i32p, ctypes.c_int32]),
        "cmf_network_set_outflow_accumulator": (ctypes.c_int, [_vp, _vp]),
        "cmf_hip_map_runoff": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, ctypes.c_int32, _vp,
                                              _vp]),
        "cmf_network_get_block_sizes": (ctypes.c_int, [_vp, i32p, ctypes.c_int32]),
        "cmf_network_destroy": (ctypes.c_int, [_vp]),
        "cmf_network_is_contiguous": (ctypes.c_int, [_vp]),
        "cmf_hip_substep": (ctypes.c_int, [_vp, p(CmfFields), ctypes.c_int32, ctypes.c_int32, _vp]),
        "cmf_hip_advance": (ctypes.c_int, [_vp, p(CmfFields), ctypes.c_double, ctypes.c_int32,
                                           i32p, _vp]),
        "cmf_hip_advance_graph": (ctypes.c_int, [_vp, p(CmfFields), ctypes.c_double, ctypes.c_int32,
                                                 i32p, _vp]),
        "cmf_hip_substep_timed": (ctypes.c_int, [_vp, p(CmfFields), ctypes.c_int32, ctypes.c_int32, _vp,
                                                 f64p, p(ctypes.c_char_p), ctypes.c_int32, i32p]),
    }
    for op in OP_SYMBOLS.values():
        sig[op] = (ctypes.c_int, [_vp, p(CmfFields), _vp])
    sig["cmf_hip_flood_stage_log_noalloc"] = (ctypes.c_int, [p(CmfFields), _vp])
    sig["cmf_network_set_schedule"] = (ctypes.c_int, [_vp, ctypes.c_int32])
    i64p = ctypes.POINTER(ctypes.c_int64)
    sig["cmf_network_local_counts"] = (ctypes.c_int, [_vp, i64p, i64p, i64p])
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


# reference op name -> exported symbol (cmfgpu/phys/*.py registries)
OP_SYMBOLS = {
    "compute_outflow": "cmf_hip_outflow",
    "compute_inflow": "cmf_hip_inflow",
    "compute_flood_stage": "cmf_hip_flood_stage",
    "compute_flood_stage_log": "cmf_hip_flood_stage_log",
    "compute_adaptive_time_step": "cmf_hip_adaptive_time_step",
    "compute_bifurcation_outflow": "cmf_hip_bifurcation_outflow",
    "compute_bifurcation_inflow": "cmf_hip_bifurcation_inflow",
    "compute_levee_stage": "cmf_hip_levee_stage",
    "compute_levee_stage_log": "cmf_hip_levee_stage_log",
    "compute_levee_bifurcation_outflow": "cmf_hip_levee_bifurcation_outflow",
    "compute_reservoir_outflow": "cmf_hip_reservoir_outflow",
}


def lib():
    """Return the loaded library, loading it on first use (raises if absent)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("CMF_HIP_LIBRARY", str(LIB_PATH))
        if not Path(path).exists():
            raise FileNotFoundError(
                f"libcmfhip.so not found at {path}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`"
            )
        handle = _declare(ctypes.CDLL(path))
        size = handle.cmf_fields_size()
        if size != ctypes.sizeof(CmfFields):
            raise RuntimeError(
                f"cmf_fields ABI mismatch: library {size} bytes, binding "
                f"{ctypes.sizeof(CmfFields)} bytes"
            )
        _LIB = handle
    return _LIB


def check(status: int, what: str = "") -> None:
    if status != CMF_OK:
        msg = lib().cmf_last_error().decode(errors="replace")
        raise CmfError(status, f"{what}: {msg}" if what else msg)


def ptr(a) -> ctypes.c_void_p:
    """Device (or host) address of a torch tensor or numpy array; None -> NULL."""
    if a is None:
        return ctypes.c_void_p(0)
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return ctypes.c_void_p(a.ctypes.data)


def current_stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
