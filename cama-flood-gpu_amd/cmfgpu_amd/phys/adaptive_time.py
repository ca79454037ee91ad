"""Adaptive CFL sub-step count (cmfgpu/phys/adaptive_time.py:34-38)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import ADAPTIVE_TIME

compute_adaptive_time_step = BackendRegistry({"hip": hip_op("compute_adaptive_time_step")},
                                             name="compute_adaptive_time_step", spec=ADAPTIVE_TIME).selected
