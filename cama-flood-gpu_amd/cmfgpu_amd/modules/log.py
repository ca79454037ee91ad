"""LogModule (cmfgpu/modules/log.py:51-326): water-balance sums and text log.

The 11 sums are FP64 here (the reference keeps them in model precision).
The text format is the reference's (log.py:80-202).
"""
from __future__ import annotations

import math
from datetime import datetime, timedelta
from pathlib import Path

import numpy as np

from .._lib import LOG_FIELDS

HEADERS = ["StepStartTime", "StoragePre", "StorageNext", "StorageNew", "InflowError", "Inflow",
           "Outflow", "StorageStage", "StageError", "RiverStorage", "FloodStorage", "FloodArea"]
FORMATS = ["%-18s"] + ["%16.6g"] * 3 + ["%16.3e"] + ["%16.6g"] * 3 + ["%16.3e"] + ["%16.6g"] * 3


class LogModule:
    module_name = "log"

    def __init__(self, log_buffer_size: int = 20000):
        if log_buffer_size < 0:
            raise ValueError("log_buffer_size must be >= 0")
        self.log_buffer_size = int(log_buffer_size)
        self.fields = {"log_sums": np.zeros((len(LOG_FIELDS), self.log_buffer_size), np.float64)}
        self._times = []
        self._time_step = 0.0
        self._num_steps = 0
        self._initialized = False

    def set_time(self, time_step: float, num_steps: int, current_time: datetime) -> None:
        """log.py:115-145."""
        if type(time_step) not in {int, float}:
            raise TypeError("time_step must be an exact int or float")
        time_step = float(time_step)
        if not math.isfinite(time_step) or time_step <= 0:
            raise ValueError("time_step must be finite and greater than zero")
        if type(num_steps) is not int:
            raise TypeError("num_steps must be an exact int")
        if num_steps < 1:
            raise ValueError("num_steps must be positive")
        if num_steps > self.log_buffer_size:
            raise ValueError(
                f"num_steps ({num_steps}) exceeds log_buffer_size ({self.log_buffer_size}). "
                "Increase log_buffer_size in the log module configuration.")
        times = [current_time + timedelta(seconds=time_step * i) for i in range(num_steps)]
        self._time_step = time_step
        self._num_steps = num_steps
        self._times = times

    def write_step(self, log_path: Path, sums: np.ndarray) -> None:
        """Append one outer step (log.py:166-202); `sums` is [11, log_buffer_size]."""
        log_path = Path(log_path)
        log_path.parent.mkdir(parents=True, exist_ok=True)
        if not self._initialized:
            widths = [18] + [16] * (len(HEADERS) - 1)
            with log_path.open("w") as fh:
                fh.write("".join(f"{h:<{w}}" if i == 0 else f"{h:>{w}}"
                                 for i, (h, w) in enumerate(zip(HEADERS, widths))) + "\n")
            self._initialized = True
        with log_path.open("a") as fh:
            fh.write(f"Time Step: {self._time_step:.4f} seconds    Number of Steps: {self._num_steps}\n")
            for i in range(self._num_steps):
                row = [self._times[i].strftime("%Y-%m-%d %H:%M")] + [sums[k, i] for k in range(len(LOG_FIELDS))]
                fh.write("".join(fmt % v for fmt, v in zip(FORMATS, row)) + "\n")
