"""Registered physics ops (cmfgpu/phys/{outflow,storage,adaptive_time,
bifurcation,levee,reservoir}.py), backed by the HIP library."""
from .outflow import compute_inflow, compute_outflow  # noqa: F401
from .storage import compute_flood_stage, compute_flood_stage_log  # noqa: F401
from .adaptive_time import compute_adaptive_time_step  # noqa: F401
from .bifurcation import compute_bifurcation_inflow, compute_bifurcation_outflow  # noqa: F401
from .levee import (  # noqa: F401
    compute_levee_bifurcation_outflow, compute_levee_stage, compute_levee_stage_log,
)
from .reservoir import compute_reservoir_outflow  # noqa: F401
