"""FP32 (FP64 storage) against the FP64 profile over a simulated month.

The (FP64, FP64) profile runs the same kernels and is bit-identical to the
FP64 CPU oracle (tests/test_gpu_parity.py::test_fp64_profile_bitexact), so these are the
errors of the FP32 path against an FP64 oracle run, on the 60,000-catchment
plain network with 30 days of daily lognormal runoff (1440 outer steps).

Stated tolerances (relative L2 over all catchments; measured values on
MI355X are about a third of these):
  discharge (river + flood outflow)   5e-5
  river / flood outflow               1e-4
  river / flood storage               5e-5
  river / flood depth                 1e-4
"""
import numpy as np
import pytest

from cmfgpu_amd import synthetic

TOL = {"discharge": 5e-5, "river_outflow": 1e-4, "flood_outflow": 1e-4, "river_storage": 5e-5,
       "flood_storage": 5e-5, "river_depth": 1e-4, "flood_depth": 1e-4}


def _rel(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))


@pytest.mark.gpu
def test_fp32_against_fp64_over_30_days():
    from cmfgpu_amd.models import CaMaFlood
    net = synthetic.build_config(0)
    src = dict(net.fields)
    models = [CaMaFlood(src, opened_modules=("base", "adaptive_time"), device="cuda", step_duration=1800.0,
                        precision=p) for p in ("float32", "float64")]
    forcing = synthetic.Runoff(src["catchment_area"], 2000)
    for day in range(30):
        r = forcing.next_day()
        for m in models:
            m.set_inputs(r)
        for _ in range(48):
            a, b = (m.step_advance() for m in models)
            assert a == b, f"sub-step counts differ on day {day + 1}: {a} vs {b}"
        if day + 1 in (1, 15, 30):
            got = {name: _rel(*(m.get(name) for m in models)) for name in TOL if name != "discharge"}
            got["discharge"] = _rel(*((m.get("river_outflow").astype(np.float64) + m.get("flood_outflow"))
                                      for m in models))
            print(f"day {day + 1}: " + ", ".join(f"{k} {v:.2e}" for k, v in got.items()), flush=True)
            for name, tol in TOL.items():
                assert got[name] <= tol, f"day {day + 1} {name}: relative L2 error {got[name]:.3e} > {tol}"
    for m in models:
        m.close()
