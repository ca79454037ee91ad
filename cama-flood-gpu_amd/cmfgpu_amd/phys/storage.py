"""Flood stage (cmfgpu/phys/storage.py:48-57)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import FLOOD_STAGE, FLOOD_STAGE_LOG

compute_flood_stage = BackendRegistry({"hip": hip_op("compute_flood_stage")},
                                      name="compute_flood_stage", spec=FLOOD_STAGE).selected
compute_flood_stage_log = BackendRegistry({"hip": hip_op("compute_flood_stage_log")},
                                          name="compute_flood_stage_log", spec=FLOOD_STAGE_LOG).selected
