"""Levee stage and levee-aware bifurcation (cmfgpu/phys/levee.py:71-94)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import LEVEE_BIFURCATION_OUTFLOW, LEVEE_STAGE, LEVEE_STAGE_LOG

compute_levee_stage = BackendRegistry({"hip": hip_op("compute_levee_stage")},
                                      name="compute_levee_stage", spec=LEVEE_STAGE).selected
compute_levee_stage_log = BackendRegistry({"hip": hip_op("compute_levee_stage_log")},
                                          name="compute_levee_stage_log", spec=LEVEE_STAGE_LOG).selected
compute_levee_bifurcation_outflow = BackendRegistry(
    {"hip": hip_op("compute_levee_bifurcation_outflow")},
    name="compute_levee_bifurcation_outflow", spec=LEVEE_BIFURCATION_OUTFLOW).selected
