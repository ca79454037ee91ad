"""ctypes wrapper of the CPU oracle (libcmf_oracle.so, built from cmf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.

State is a dict of numpy arrays keyed by the reference's module field names
(the arrays are updated in place, as the reference's kernels update module
fields).  Catchment-indexed arrays are in the network order the caller uses.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcmf_oracle.so"

# Field table order of cmf_oracle.cpp (enum Field).
FIELDS = (
    "downstream_idx",
    "river_inflow", "flood_inflow", "river_outflow", "flood_outflow",
    "river_manning", "flood_manning", "river_depth", "flood_depth",
    "protected_depth", "river_width", "river_length", "river_height",
    "catchment_elevation", "catchment_area", "downstream_distance",
    "river_storage", "flood_storage", "protected_storage",
    "river_cross_section_depth", "flood_cross_section_depth", "flood_cross_section_area",
    "flood_fraction", "runoff", "flood_depth_table", "outgoing_storage", "total_storage",
    "water_surface_elevation", "protected_water_surface_elevation", "is_levee",
    "bifurcation_catchment_idx", "bifurcation_downstream_idx", "bifurcation_manning",
    "bifurcation_outflow", "bifurcation_width", "bifurcation_length", "bifurcation_elevation",
    "bifurcation_cross_section_depth", "global_bifurcation_outflow", "limit_rate",
    "levee_catchment_idx", "levee_base_height", "levee_crown_height", "levee_fraction",
    "reservoir_catchment_idx", "conservation_volume", "emergency_volume",
    "adjustment_volume", "effective_normal_outflow", "adjustment_outflow",
    "flood_control_outflow", "reservoir_total_inflow", "is_reservoir",
    "is_dam_upstream", "is_dam_related",
    "log_sums",
)

HAS_BIFURCATION, HAS_LEVEE, HAS_RESERVOIR, HAS_LOG = 1, 2, 4, 8
HAS_TOTAL_STORAGE, HAS_WATER_SURFACE, HAS_PROTECTED_WATER_SURFACE, HAS_ADAPTIVE = 16, 32, 64, 128
PRECISE_SURFACE = 512   # FP32 profile: water surfaces formed in double

KERNELS = {
    "outflow": 0, "inflow": 1, "flood_stage": 2, "bifurcation_outflow": 3,
    "bifurcation_inflow": 4, "levee_stage": 5, "reservoir": 6,
    "levee_bifurcation_outflow": 7,
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(str(LIB_PATH))
        vpp = ctypes.POINTER(ctypes.c_void_p)
        i64p = ctypes.POINTER(ctypes.c_int64)
        f64p = ctypes.POINTER(ctypes.c_double)
        h.cmf_oracle_advance.restype = ctypes.c_int
        h.cmf_oracle_advance.argtypes = [ctypes.c_int, vpp, i64p, f64p, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
        h.cmf_oracle_set_threads.restype = None
        h.cmf_oracle_set_threads.argtypes = [ctypes.c_int]
        h.cmf_oracle_max_threads.restype = ctypes.c_int
        h.cmf_oracle_max_threads.argtypes = []
        h.cmf_oracle_adaptive_steps.restype = ctypes.c_int
        h.cmf_oracle_adaptive_steps.argtypes = [ctypes.c_int, vpp, i64p, f64p, ctypes.c_int, ctypes.c_double]
        h.cmf_oracle_kernel.restype = ctypes.c_int
        h.cmf_oracle_kernel.argtypes = [ctypes.c_int, ctypes.c_int, vpp, i64p, f64p, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int]
        for name, rt, at in (("cmf_oracle_cbrtf", ctypes.c_float, ctypes.c_float),
                             ("cmf_oracle_cbrt", ctypes.c_double, ctypes.c_double),
                             ("cmf_oracle_logf", ctypes.c_float, ctypes.c_float),
                             ("cmf_oracle_expf", ctypes.c_float, ctypes.c_float)):
            fn = getattr(h, name)
            fn.restype = rt
            fn.argtypes = [at]
        h.cmf_oracle_num_fields.restype = ctypes.c_int
        if h.cmf_oracle_num_fields() != len(FIELDS):
            raise RuntimeError("oracle field table out of sync")
        _LIB = h
    return _LIB


def set_threads(threads: int) -> int:
    """Set the oracle's OpenMP thread count; returns the previous one."""
    h = lib()
    prev = int(h.cmf_oracle_max_threads())
    h.cmf_oracle_set_threads(int(threads))
    return prev


class OracleModel:
    """Sequential CPU model over a dict of numpy arrays.

    precision 0: REAL float32 / STO float64; 1: float64 / float64; 2: float32 /
    float32 (the water-balance sums stay float64).
    """

    def __init__(self, state: dict, *, flags: int, num_flood_levels: int,
                 num_bifurcation_levels: int = 0, log_buffer_size: int = 0,
                 gravity: float = 9.8, adaptive_time_factor: float = 0.7,
                 min_kinematic_slope: float = 1e-5, precision: int = 0):
        self.state = state
        self.flags = flags
        self.precision = precision
        n = state["downstream_idx"].shape[0]
        P = state["bifurcation_catchment_idx"].shape[0] if "bifurcation_catchment_idx" in state else 0
        V = state["levee_catchment_idx"].shape[0] if "levee_catchment_idx" in state else 0
        R = state["reservoir_catchment_idx"].shape[0] if "reservoir_catchment_idx" in state else 0
        self.dims = np.array([n, num_flood_levels, P, num_bifurcation_levels, V, R, log_buffer_size],
                             dtype=np.int64)
        self.scalars = np.array([gravity, adaptive_time_factor, min_kinematic_slope], dtype=np.float64)
        self._check_dtypes()

    def _check_dtypes(self):
        real = np.float64 if self.precision == 1 else np.float32
        sto = np.float32 if self.precision == 2 else np.float64
        for name, arr in self.state.items():
            if name not in FIELDS:
                continue
            if not arr.flags.c_contiguous:
                raise ValueError(f"{name} must be C-contiguous")
            if name.endswith("_idx"):
                want = np.int32
            elif name.startswith("is_"):
                want = np.bool_
            elif name in ("river_inflow", "flood_inflow", "river_storage", "flood_storage",
                          "protected_storage", "outgoing_storage", "total_storage",
                          "global_bifurcation_outflow", "reservoir_total_inflow"):
                want = sto
            elif name == "log_sums":
                want = np.float64
            else:
                want = real
            if arr.dtype != want:
                raise TypeError(f"{name}: dtype {arr.dtype}, expected {np.dtype(want)}")

    def _table(self):
        ptrs = (ctypes.c_void_p * len(FIELDS))()
        for k, name in enumerate(FIELDS):
            arr = self.state.get(name)
            ptrs[k] = None if arr is None else arr.ctypes.data
        return ptrs

    def _args(self):
        return (self._table(), self.dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                self.scalars.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def advance(self, outer_seconds: float, num_outer: int = 1, fixed_count: int = 1) -> np.ndarray:
        """Run `num_outer` outer steps (CaMaFlood.step_advance); returns the sub-step counts."""
        counts = np.zeros(num_outer, dtype=np.int32)
        t, d, s = self._args()
        lib().cmf_oracle_advance(self.precision, t, d, s, self.flags, float(outer_seconds),
                                 int(fixed_count), int(num_outer),
                                 counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return counts

    def adaptive_steps(self, outer_seconds: float) -> int:
        t, d, s = self._args()
        return int(lib().cmf_oracle_adaptive_steps(self.precision, t, d, s, self.flags, float(outer_seconds)))

    def kernel(self, name: str, time_step: float = 0.0, current_step: int = 0) -> None:
        t, d, s = self._args()
        rc = lib().cmf_oracle_kernel(self.precision, KERNELS[name], t, d, s, self.flags,
                                     float(time_step), int(current_step))
        if rc != 0:
            raise RuntimeError(f"oracle kernel {name} failed")


def cbrtf(x: float) -> float:
    return lib().cmf_oracle_cbrtf(x)
