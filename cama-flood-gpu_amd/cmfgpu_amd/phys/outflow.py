"""Main-channel outflow and inflow (cmfgpu/phys/outflow.py:51-60)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import INFLOW, OUTFLOW

compute_outflow = BackendRegistry({"hip": hip_op("compute_outflow")}, name="compute_outflow", spec=OUTFLOW).selected
compute_inflow = BackendRegistry({"hip": hip_op("compute_inflow")}, name="compute_inflow", spec=INFLOW).selected
