"""The PRECISE_SURFACE flag of the CPU oracle on whole networks (CPU only).

With the flag the FP32 profile forms water surfaces, slopes and depths above
bed, bank and path sills in double from its FP32 depth, elevation and height.
Checked against the FP64 profile, which the flag leaves bit for bit as it is:
on 3,000 catchments with 100 paths and levees on 5 % (runoff x4, one day,
48 outer steps with the adaptive sub-step count) the discharge and the path
flows follow FP64 to 1e-6 (relative L2; measured 1.5e-7 and 1.2e-7), at least
ten times closer than the reference's FP32 surfaces (6.0e-6, 1.1e-5), with
the same sub-step counts in every outer step.
"""
import numpy as np

from cmfgpu_amd import synthetic
from helpers import FULL, oracle_for

MODS = ("base", "adaptive_time", "bifurcation", "levee")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _run(src, mods, precision, precise, days=1, scale=4.0):
    orc, _ = oracle_for(src, mods, precision=precision, precise_surfaces=precise)
    forcing = synthetic.Runoff(src["catchment_area"], 2000)
    counts = []
    for _ in range(days):
        orc.state["runoff"][:] = forcing.next_day() * np.float32(scale)
        counts.append(np.array(orc.advance(1800.0, 48), copy=True))
    keep = ("river_outflow", "flood_outflow", "bifurcation_outflow", "river_storage", "flood_storage",
            "protected_storage", "reservoir_total_inflow")
    return {k: np.array(orc.state[k], np.float64, copy=True) for k in keep if k in orc.state}, np.concatenate(counts)


def test_flag_changes_nothing_in_fp64():
    net = synthetic.generate_network(3000, 1004, levee=0.05, paths=100, reservoirs=30)
    src = dict(net.fields)
    a, ca = _run(src, FULL, "float64", False)
    b, cb = _run(src, FULL, "float64", True)
    assert np.array_equal(ca, cb)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_precise_fp32_follows_fp64():
    net = synthetic.generate_network(3000, 1003, levee=0.05, paths=100)
    src = dict(net.fields)
    ref, c64 = _run(src, MODS, "float64", False)
    err = {}
    for precise in (False, True):
        got, c32 = _run(src, MODS, "float32", precise)
        assert np.array_equal(c32, c64)
        err[precise] = (_rel(got["river_outflow"] + got["flood_outflow"], ref["river_outflow"] + ref["flood_outflow"]),
                        _rel(got["bifurcation_outflow"], ref["bifurcation_outflow"]))
    print(f"discharge / path flows vs FP64: reference surfaces {err[False]}, precise {err[True]}")
    assert err[True][0] <= 1e-6 and err[True][1] <= 1e-6, err
    assert err[False][0] >= 10 * err[True][0] and err[False][1] >= 10 * err[True][1], err
