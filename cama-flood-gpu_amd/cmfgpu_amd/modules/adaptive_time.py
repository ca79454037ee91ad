"""AdaptiveTimeModule (cmfgpu/modules/adaptive_time.py:21-45)."""
from __future__ import annotations

import numpy as np


class AdaptiveTimeModule:
    module_name = "adaptive_time"

    def __init__(self, adaptive_time_factor: float = 0.7, slots: int = 1):
        if not (adaptive_time_factor > 0) or not np.isfinite(adaptive_time_factor):
            raise ValueError("adaptive_time_factor must be finite and greater than zero")
        self.adaptive_time_factor = float(adaptive_time_factor)
        # max_sub_steps: one slot, or one per trial with per-trial time steps
        self.fields = {"max_sub_steps": np.zeros(slots, np.int32)}
