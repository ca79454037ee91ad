"""Model driver (cmfgpu/models)."""
from .cama_flood_model import CaMaFlood  # noqa: F401
