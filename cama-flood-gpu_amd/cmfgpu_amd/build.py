"""Build recipe of libcmfhip.so (hipcc, gfx950).

Compiled in-tree into cama-flood-gpu_amd/lib so the library travels with the
source tree.  Floating-point contraction is off and divide/sqrt are correctly
rounded, so each kernel evaluates exactly the operation sequence written;
packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "csrc"
LIB_DIR = ROOT / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CMF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-pass-failed",
]
SOURCES = ["cmf_capi.hip", "cmf_network.cpp"]


def build(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    out = LIB_DIR / "libcmfhip.so"
    tmp = LIB_DIR / "libcmfhip.so.tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-o", str(tmp), *[str(CSRC / s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
