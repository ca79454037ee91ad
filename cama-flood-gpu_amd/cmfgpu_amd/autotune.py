"""Block-size autotuner for the fused sub-step schedule.

Every tunable kernel of the fused schedule takes its block size as a launch
parameter (multiples of 64 up to 512; include/cmf_hip.h, CMF_TUNE_*).  For a
model, ``tune`` times one fused sub-step per candidate size with HIP events
around each kernel (cmf_hip_substep_timed), applying the candidate to every
kernel at once, and keeps for each kernel the size with the shortest mean
duration.  The model's state is saved before and restored after, so tuning
does not advance the simulation.  Choices are cached in a JSON file keyed by
the device name, the problem shape and the enabled features.

The reference passes a fixed BLOCK_SIZE to every launch
(cmfgpu/phys/specs.py:10-26); there is no counterpart to this module.
"""
from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

from . import _lib

CANDIDATES = (64, 128, 256, 512)
KERNELS = {
    "k_outflow": 0, "k_outflow_paths": 0, "k_outflow_prov": 0, "k_outflow_paths_prov": 0, "k_reservoir": 1, "k_limiter": 2, "k_limiter_pair": 2,
    "k_limiter_fix": 2, "k_stage_fused": 3,
    # the block-local schedule: its list passes take the limiter and stage
    # sizes (k_fused_local runs fixed 512-catchment blocks)
    "k_limiter_list": 2, "k_stage_list": 3,
}
NUM_TUNABLE = 4


def kernel_times(model, reps: int) -> dict:
    """Mean duration (ms) of each kernel of one fused sub-step (sub-step 0,
    inflow not stored), HIP events on the launch stream.  Advances the state."""
    cap = 64
    ms = (ctypes.c_double * cap)()
    names = (ctypes.c_char_p * cap)()
    count = ctypes.c_int32(0)
    acc = {}
    for _ in range(reps):
        _lib.check(_lib.lib().cmf_hip_substep_timed(model.network_handle, model.field_table_ref(), 0, 0,
                                                    _lib.current_stream(), ms, names, cap, ctypes.byref(count)),
                   "cmf_hip_substep_timed")
        for k in range(count.value):
            key = names[k].decode().strip("()")
            acc.setdefault(key, []).append(ms[k])
    return {k: sum(v) / len(v) for k, v in acc.items()}


def get_block_sizes(model) -> list:
    out = (ctypes.c_int32 * NUM_TUNABLE)()
    _lib.check(_lib.lib().cmf_network_get_block_sizes(model.network_handle, out, NUM_TUNABLE),
               "cmf_network_get_block_sizes")
    return list(out)


def set_block_sizes(model, sizes) -> None:
    arr = (ctypes.c_int32 * NUM_TUNABLE)(*[int(x) for x in sizes])
    _lib.check(_lib.lib().cmf_network_set_block_sizes(model.network_handle, arr, NUM_TUNABLE),
               "cmf_network_set_block_sizes")


def _key(model) -> str:
    import torch
    dev = torch.cuda.get_device_name(model.device) if model.device.type == "cuda" else str(model.device)
    return (f"{dev}|N={model.N}|T={model.T}|P={model.P}|V={model.V}|R={model.R}|G={model.G}|"
            f"flags={model.flags}|prec={model.real.__name__}|layout={model.layout}|"
            f"trials={model.trial_layout}|kernels={','.join(sorted(KERNELS, key=KERNELS.get))}")


def default_cache_path() -> Path:
    return Path(os.environ.get("CMF_TUNE_CACHE", Path.home() / ".cache" / "cmfgpu_amd" / "block_sizes.json"))


def tune(model, *, reps: int = 16, candidates=CANDIDATES, cache_path: str | Path | None = None,
         use_cache: bool = True) -> dict:
    """Pick and apply the block size of each tunable kernel; returns
    {kernel name: {"block": size, "ms": {size: mean ms}}}."""
    import torch
    path = Path(cache_path) if cache_path is not None else default_cache_path()
    key = _key(model)
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache and len(cache[key].get("sizes", ())) == NUM_TUNABLE:
            set_block_sizes(model, cache[key]["sizes"])
            return cache[key]["report"]
    saved = {name: t.clone() for name, t in model.t.items()}
    original = get_block_sizes(model)

    def restore():
        for name, t in model.t.items():
            t.copy_(saved[name])
        torch.cuda.synchronize()

    times = {}
    try:
        # every candidate runs the same sub-steps from the same state: the
        # stage pass's cost follows the wet fraction, which drifts as the
        # state advances
        for size in candidates:
            set_block_sizes(model, [size] * NUM_TUNABLE)
            restore()
            kernel_times(model, 2)               # warm-up
            restore()
            for name, ms in kernel_times(model, reps).items():
                base = name.split("<")[0]
                if base in KERNELS:
                    times.setdefault(base, {})[size] = ms
    finally:
        restore()
    sizes = list(original)
    report = {}
    for base, per in times.items():
        best = min(per, key=per.get)
        sizes[KERNELS[base]] = best
        report[base] = {"block": best, "ms": {str(k): round(v, 5) for k, v in per.items()}}
    set_block_sizes(model, sizes)
    if use_cache:
        cache[key] = {"sizes": sizes, "report": report}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1, sort_keys=True))
        except OSError:
            pass
    return report
