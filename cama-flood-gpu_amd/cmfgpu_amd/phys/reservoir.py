"""Reservoir release (cmfgpu/phys/reservoir.py:33-37)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import RESERVOIR_OUTFLOW

compute_reservoir_outflow = BackendRegistry({"hip": hip_op("compute_reservoir_outflow")},
                                            name="compute_reservoir_outflow", spec=RESERVOIR_OUTFLOW).selected
