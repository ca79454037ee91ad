"""Bifurcation pathways (cmfgpu/phys/bifurcation.py:45-62)."""
from ._hip import hip_op
from .registry import BackendRegistry
from .specs import BIFURCATION_INFLOW, BIFURCATION_OUTFLOW

compute_bifurcation_outflow = BackendRegistry({"hip": hip_op("compute_bifurcation_outflow")},
                                              name="compute_bifurcation_outflow", spec=BIFURCATION_OUTFLOW).selected
compute_bifurcation_inflow = BackendRegistry({"hip": hip_op("compute_bifurcation_inflow")},
                                             name="compute_bifurcation_inflow", spec=BIFURCATION_INFLOW).selected
