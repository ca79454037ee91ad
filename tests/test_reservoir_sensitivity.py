"""Why no FP32 run of the reference scheme meets a 1e-5 discharge bound in the
reaches below reservoirs, shown with the FP64 oracle alone (CPU).

The release rule (triton/reservoir.py:114-163) raises the release above the
adjustment volume as ((S - S_adj) / (S_emg - S_adj))^0.1: its slope is
infinite at the threshold.  A storage 1e-7 above it (the size of one FP32
rounding of the storage) already releases a fifth of the way from the
adjustment to the flood-control outflow.  So a perturbation of the state the
size of FP32 rounding moves the discharge below a dam by per cents once a
reservoir sits near that threshold, in FP64 arithmetic, with no FP32 anywhere.

Stated bound, with this demonstration: discharge of the FP32 path against the
FP64 oracle <= 1e-5 outside the reaches below reservoirs (the dam and every
catchment downstream of it) with precise surfaces; inside those reaches the
FP64 scheme itself moves by more than 1e-4 under a 1e-7 perturbation.
"""
import numpy as np
import pytest

from cmfgpu_amd import synthetic
from helpers import FULL, oracle_for
from test_oracle_pins import reservoir_state
from oracle import oracle as O


def _release(total, inflow=10.0):
    s, p = reservoir_state(total, inflow)
    s = {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in s.items()}
    m = O.OracleModel(s, flags=O.HAS_RESERVOIR | O.HAS_TOTAL_STORAGE, num_flood_levels=10, precision=1)
    m.kernel("outflow", time_step=600.0)
    m.kernel("reservoir", time_step=600.0)
    return float(s["river_outflow"][1]), p


def test_release_has_an_infinite_slope_at_the_adjustment_volume():
    """FP64: S = S_adj (1 + 1e-7) releases Q_adj + ((S - S_adj)/(S_emg -
    S_adj))^0.1 (Q_fc - Q_adj), about a fifth of the way to Q_fc, while
    S = S_adj releases Q_adj."""
    _, p = _release(1e7)
    # the parameters as the state stores them (float32 values)
    av, ev, qa, qf = (float(np.float32(p[k])) for k in ("av", "ev", "qa", "qf"))
    at, _ = _release(av)
    above, _ = _release(av * (1 + 1e-7))
    expect = qa + ((av * 1e-7) / (ev - av)) ** 0.1 * (qf - qa)
    assert at == pytest.approx(qa, rel=1e-12)
    assert above == pytest.approx(expect, rel=1e-9)
    assert (above - at) / qa > 0.05


def _reaches_below(dn, start):
    below = np.zeros(dn.shape[0], bool)
    cur = np.asarray(start).copy()
    below[cur] = True
    while cur.size:
        nxt = dn[cur]
        cur = nxt[~below[nxt]]
        below[cur] = True
    return below


def _discharge(orc):
    return orc.state["river_outflow"].astype(np.float64) + orc.state["flood_outflow"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_fp64_scheme_amplifies_rounding_sized_perturbations_below_dams():
    """80,000 catchments (levees on 5 %, 2,000 paths, 200 reservoirs), one
    day: the FP64 oracle from the initial state and from that state with
    every river storage scaled by 1 +- 1e-7 agree to about 1e-7 away from
    the reservoirs and differ by more than 1e-4 over the whole network, all
    of it in the reaches below dams; the FP32 path with precise surfaces is
    within 1e-5 of FP64 outside those reaches."""
    n = 80_000
    net = synthetic.generate_network(n, 1004, levee=0.05, paths=n // 40, reservoirs=n // 400)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 4000).next_day()
    def run(scale=None, **kw):
        s = dict(src)
        if scale is not None:
            s["river_storage"] = np.asarray(src["river_storage"], np.float64) * scale
        orc, _ = oracle_for(s, FULL, **kw)
        counts = orc.advance(1800.0, 48)
        return _discharge(orc), counts

    prev = O.set_threads(8)
    try:
        q64, c64 = run(precision="float64")
        sign = np.random.default_rng(1).choice([-1.0, 1.0], n)
        qp, cp = run(1.0 + 1e-7 * sign, precision="float64")
        q32, c32 = run(precise_surfaces=True)
    finally:
        O.set_threads(prev)
    assert np.array_equal(c64, cp) and np.array_equal(c64, c32)
    below = _reaches_below(np.asarray(src["downstream_idx"]), src["reservoir_catchment_idx"])
    out = ~below
    assert below.mean() < 0.1
    assert _rel(qp[out], q64[out]) < 1e-6
    assert _rel(qp, q64) > 1e-4
    assert _rel(q32[out], q64[out]) <= 1e-5
