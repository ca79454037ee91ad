"""Host-side module layer: field names, dtypes, defaults, derived fields and
validators of the reference's modules (cmfgpu/modules/*.py).

Each module class turns a parameter source (a dict of numpy arrays keyed by
field name, as a parameters.nc would hold them) into the complete set of
fields the kernels use, in the reference's order and dtypes:
  float   -> REAL (float32, or float64 for the FP64 profile)
  hpfloat -> STO (float64)
  idx     -> int32, int -> int64, bool -> bool
"""
from .base import BaseModule  # noqa: F401
from .adaptive_time import AdaptiveTimeModule  # noqa: F401
from .bifurcation import BifurcationModule  # noqa: F401
from .levee import LeveeModule  # noqa: F401
from .reservoir import ReservoirModule  # noqa: F401
from .log import LogModule  # noqa: F401

MODULE_CLASSES = {
    "base": BaseModule,
    "adaptive_time": AdaptiveTimeModule,
    "bifurcation": BifurcationModule,
    "levee": LeveeModule,
    "reservoir": ReservoirModule,
    "log": LogModule,
}
