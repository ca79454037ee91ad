"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests or golden vectors (its tests/ is git-ignored),
so each oracle kernel is pinned here by closed forms, invariants and
brute-force sums that a dropped term, a wrong sign or a wrong index would
break.  Every expected value below is computed independently of the oracle's
formula (in float64 from the physics), not by retyping it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = 9.8


def tiny_state(dn, L=10, **over):
    n = len(dn)
    s = {
        "downstream_idx": np.asarray(dn, np.int32),
        "river_inflow": np.zeros(n), "flood_inflow": np.zeros(n),
        "river_outflow": np.zeros(n, np.float32), "flood_outflow": np.zeros(n, np.float32),
        "river_manning": np.full(n, 0.03, np.float32), "flood_manning": np.full(n, 0.1, np.float32),
        "river_depth": np.zeros(n, np.float32), "flood_depth": np.zeros(n, np.float32),
        "protected_depth": np.zeros(n, np.float32),
        "river_width": np.full(n, 50.0, np.float32), "river_length": np.full(n, 10_000.0, np.float32),
        "river_height": np.full(n, 4.0, np.float32), "catchment_elevation": np.zeros(n, np.float32),
        "catchment_area": np.full(n, 5e8, np.float32), "downstream_distance": np.full(n, 10_000.0, np.float32),
        "river_storage": np.zeros(n), "flood_storage": np.zeros(n), "protected_storage": np.zeros(n),
        "river_cross_section_depth": np.zeros(n, np.float32),
        "flood_cross_section_depth": np.zeros(n, np.float32),
        "flood_cross_section_area": np.zeros(n, np.float32),
        "flood_fraction": np.zeros(n, np.float32), "runoff": np.zeros(n, np.float32),
        "flood_depth_table": np.tile(np.cumsum(np.full(L, 0.5, np.float32)), (n, 1)).astype(np.float32),
        "outgoing_storage": np.zeros(n),
    }
    for k, v in over.items():
        s[k] = np.ascontiguousarray(v, dtype=s[k].dtype if k in s else None)
    return s


def model(state, flags=0, L=10, precision=0):
    return O.OracleModel(state, flags=flags, num_flood_levels=L, precision=precision)


# ------------------------------------------------------------ math helpers --
def test_cbrt_correct_rounding_and_exact_cubes():
    """The float cube root rounds to the same float as the C library's double
    cbrt, for positive, negative and subnormal inputs, and cubes are exact."""
    rng = np.random.default_rng(0)
    pos = np.concatenate([rng.lognormal(0, 8, 20000), [1e-6, 1e-5, 0.125, 8.0, 27.0, 1e30, 3.0]])
    sub = np.array([1.4e-45, 1e-40, 1.1754942e-38], dtype=np.float32)
    xs = np.concatenate([pos.astype(np.float32), -pos[:2000].astype(np.float32), sub, -sub])
    for x in xs:
        got = np.float32(O.cbrtf(float(x)))
        ref = np.float32(np.cbrt(np.float64(x)))
        assert got.view(np.int32) == ref.view(np.int32), (x, got, ref)
    for k in range(1, 1001):
        assert O.cbrtf(float(k ** 3)) == float(k)
    assert O.cbrtf(0.0) == 0.0
    assert np.isinf(O.cbrtf(float("inf"))) and np.isnan(O.cbrtf(float("nan")))


def test_log_exp_within_one_ulp():
    lib = O.lib()
    rng = np.random.default_rng(1)
    for x in rng.uniform(1e-4, 1.0, 5000).astype(np.float32):
        ref = np.float32(np.log(np.float64(x)))
        got = np.float32(lib.cmf_oracle_logf(float(x)))
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1
    for y in rng.uniform(-20, 5, 5000).astype(np.float32):
        ref = np.float32(np.exp(np.float64(y)))
        got = np.float32(lib.cmf_oracle_expf(float(y)))
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1


# ---------------------------------------------------------------- outflow --
def test_outflow_steady_uniform_flow_is_a_fixed_point():
    """Setting q_new = q_old in the semi-implicit update gives Manning's law
    q = h^(5/3) S^(1/2) / n per unit width (triton/outflow.py:200-216)."""
    h, S, n_man, W, dx = 2.0, 4e-4, 0.03, 50.0, 10_000.0
    q_unit = h ** (5 / 3) * math.sqrt(S) / n_man
    # catchment 0 flows into mouth 1; water surfaces differ by S * dx
    elev = np.array([10.0 + S * dx, 10.0], np.float32)
    height = np.array([4.0, 4.0], np.float32)
    depth = np.array([h, h], np.float32)
    s = tiny_state([1, 1], river_depth=depth, catchment_elevation=elev, river_height=height,
                   river_cross_section_depth=np.array([h, h], np.float32),
                   river_outflow=np.array([W * q_unit, 0.0], np.float32),
                   river_storage=np.array([h * W * dx, h * W * dx]))
    m = model(s)
    m.kernel("outflow", time_step=600.0)
    assert s["river_outflow"][0] == pytest.approx(W * q_unit, rel=2e-5)


def test_outflow_zero_initial_state_gives_zero_flow_and_dry_flood():
    """Zero cross-section state: semi-implicit depth 1e-6 <= 1e-5, so q = 0
    exactly on the first step; no floodplain storage, no flood flow."""
    s = tiny_state([1, 2, 2], river_depth=np.array([1.0, 2.0, 0.5], np.float32),
                   catchment_elevation=np.array([30.0, 20.0, 10.0], np.float32))
    m = model(s)
    m.kernel("outflow", time_step=600.0)
    assert np.all(s["river_outflow"] == 0.0) and np.all(s["flood_outflow"] == 0.0)
    # the new cross-section depth is the depth under the higher water surface
    wse = s["catchment_elevation"] - s["river_height"] + s["river_depth"]
    expect0 = max(wse[0], wse[1]) - (30.0 - 4.0)
    assert s["river_cross_section_depth"][0] == pytest.approx(expect0, rel=1e-6)
    # a mouth uses its own depth (outflow.py:163-167)
    assert s["river_cross_section_depth"][2] == np.float32(0.5)


def test_outflow_sign_follows_slope_and_outgoing_is_positive_part():
    h, dx, W = 2.0, 10_000.0, 50.0
    for drop in (+0.5, -0.5):
        elev = np.array([10.0 + drop, 10.0, 0.0], np.float32)
        s = tiny_state([1, 2, 2], river_depth=np.full(3, h, np.float32), catchment_elevation=elev,
                       river_cross_section_depth=np.full(3, h, np.float32),
                       river_storage=np.full(3, h * W * dx))
        m = model(s)
        dt = 60.0
        m.kernel("outflow", time_step=dt)
        q0 = float(s["river_outflow"][0])
        # from rest the denominator is 1: q = W g dt h_s S, with h_s the
        # semi-implicit depth sqrt(h_new h_old), h_new measured under the
        # higher of the two water surfaces
        S = (drop) / dx
        wse = elev.astype(np.float64) - 4.0 + h
        h_new = max(wse[0], wse[1]) - (float(elev[0]) - 4.0)
        h_s = math.sqrt(h_new * h)
        assert q0 == pytest.approx(W * G * dt * h_s * S, rel=1e-5)
        og = s["outgoing_storage"]
        if q0 > 0:
            assert og[0] == pytest.approx(q0 * dt, rel=1e-6)
        else:
            # negative flow leaves catchment 1 (the downstream one): its outgoing holds -q dt
            assert og[0] == 0.0
            assert og[1] >= -q0 * dt * (1 - 1e-6)


def test_precise_surfaces_keep_the_centimetres():
    """High in the mountains (elevation 3,000 m, where FP32 spacing is
    0.24 mm) a 1.23 cm surface drop over 10 km: from rest q = W g dt h_s S,
    with S and h_s from the exact (float64) surfaces of the FP32 inputs.
    With PRECISE_SURFACE the FP32 profile forms the surfaces in double and
    meets the closed form to FP32 rounding; the reference's FP32 surfaces
    (triton/outflow.py:111-134) carry their rounding into S at the per-cent
    level.  The flag changes nothing in the FP64 profile."""
    h, dx, W, dt = 2.0, 10_000.0, 50.0, 60.0
    elev = np.array([3000.0123, 3000.0, 2990.0], np.float32)
    depth = np.array([2.00037, 1.99961, 2.0], np.float32)
    height = np.array([4.0, 4.0, 4.0], np.float32)
    wse = elev.astype(np.float64) - height + depth
    S = (wse[0] - wse[1]) / dx
    h_new = max(wse[0], wse[1]) - (float(elev[0]) - 4.0)
    q_exact = W * G * dt * math.sqrt(h_new * h) * S
    got = {}
    for tag, flags, prec in (("default", 0, 0), ("precise", O.PRECISE_SURFACE, 0), ("fp64", 0, 1),
                             ("fp64 precise", O.PRECISE_SURFACE, 1)):
        real = np.float32 if prec == 0 else np.float64
        over = dict(river_depth=depth, catchment_elevation=elev, river_height=height,
                    river_cross_section_depth=np.full(3, h, np.float32), river_storage=np.full(3, h * W * dx))
        s = tiny_state([1, 2, 2], **over)
        for k, v in s.items():
            if v.dtype == np.float32:
                s[k] = v.astype(real)
        m = model(s, flags=flags, precision=prec)
        m.kernel("outflow", time_step=dt)
        got[tag] = (float(s["river_outflow"][0]), float(s["river_cross_section_depth"][0]))
    assert got["precise"][0] == pytest.approx(q_exact, rel=2e-6)
    assert got["precise"][1] == pytest.approx(h_new, rel=1e-6)
    assert abs(got["default"][0] / q_exact - 1.0) > 1e-3
    assert got["fp64"] == got["fp64 precise"]
    assert got["fp64"][0] == pytest.approx(q_exact, rel=1e-12)


def test_backflow_is_limited_to_five_percent_of_the_storage():
    """Water surface higher downstream: the flow turns upstream, and at most
    5 % of the catchment's storage may leave backwards in one sub-step
    (triton/outflow.py:235-245), so a backflow that would move more is scaled
    to exactly 0.05 S / dt."""
    h, dx, W, dt = 2.0, 10_000.0, 50.0, 600.0
    s0 = 1e5
    s = tiny_state([1, 1], river_depth=np.full(2, h, np.float32),
                   catchment_elevation=np.array([10.0, 11.0], np.float32),
                   river_cross_section_depth=np.full(2, h, np.float32),
                   river_storage=np.array([s0, 1e8]))
    # unlimited, from rest: q = W g dt h_s S (S = -1e-4, h_s = sqrt((h + 1) h))
    q_free = W * G * dt * math.sqrt((h + 1.0) * h) * (-1.0 / dx)
    assert -q_free * dt > 0.05 * s0
    model(s).kernel("outflow", time_step=dt)
    assert float(s["river_outflow"][0]) == pytest.approx(-0.05 * s0 / dt, rel=1e-6)
    assert float(s["flood_outflow"][0]) == 0.0


# ----------------------------------------------------------------- inflow --
def test_inflow_is_the_sum_of_upstream_flows():
    rng = np.random.default_rng(3)
    n = 40
    dn = np.array([min(n - 1, i + 1 + int(rng.integers(0, 3))) for i in range(n)], np.int32)
    dn[n - 1] = n - 1
    q = rng.integers(0, 50, n).astype(np.float32)
    s = tiny_state(dn, river_outflow=q, flood_outflow=q * 0.5,
                   river_storage=np.full(n, 1e12), outgoing_storage=np.zeros(n))
    m = model(s)
    m.kernel("inflow")
    expect = np.zeros(n)
    for i in range(n):
        if dn[i] != i:
            expect[dn[i]] += q[i]
    assert np.array_equal(s["river_inflow"], expect)          # integer-valued: exact
    assert s["river_inflow"].sum() == q[dn != np.arange(n)].sum()


def test_inflow_limiter_caps_outgoing_at_storage():
    # catchment 0 wants to send 10x its storage; the limiter scales to storage
    s = tiny_state([1, 1], river_outflow=np.array([100.0, 0.0], np.float32),
                   river_storage=np.array([1000.0, 5e6]), outgoing_storage=np.array([10_000.0, 0.0]))
    m = model(s)
    m.kernel("inflow")
    assert s["river_outflow"][0] == pytest.approx(10.0, rel=1e-6)
    assert s["river_inflow"][1] == pytest.approx(10.0, rel=1e-6)
    # outgoing below 1e-8: limit 1
    s2 = tiny_state([1, 1], river_outflow=np.array([5.0, 0.0], np.float32),
                    river_storage=np.array([0.0, 0.0]), outgoing_storage=np.array([1e-9, 0.0]))
    model(s2).kernel("inflow")
    assert s2["river_outflow"][0] == np.float32(5.0)


# ----------------------------------------------------------- flood stage --
def forward_storage(depth, W, Lr, H, ca, table):
    """(storage, flooded width) at flood depth `depth`, in float64.

    The water-surface width grows linearly from the channel width W at the
    bank top to W + k * (catchment width / L) at table depth k, and stays at
    the full width above the last level; storage is the bankfull channel plus
    the length times the integral of that width over depth."""
    L = table.shape[0]
    wi = ca / Lr / L
    S = Lr * W * H
    if depth <= 0:
        return S, W
    prevH, prevW = 0.0, W
    for k in range(L):
        Hk = table[k]
        Wk = W + (k + 1) * wi
        if depth <= Hk:
            w_at = prevW + (depth - prevH) / (Hk - prevH) * (Wk - prevW)
            return S + Lr * (depth - prevH) * (prevW + w_at) / 2, w_at
        S += Lr * 0.5 * (prevW + Wk) * (Hk - prevH)
        prevH, prevW = Hk, Wk
    return S + Lr * prevW * (depth - prevH), prevW


@pytest.mark.parametrize("depth", [0.0, 0.3, 0.9, 1.7, 3.2, 4.99, 7.5])
def test_stage_round_trip_on_the_floodplain_profile(depth):
    """Store the water of a known flood depth; the diagnosis must return it."""
    W, Lr, H, ca = 80.0, 12_000.0, 3.0, 4e8
    table = np.cumsum(np.array([0.2, 0.5, 0.1, 0.7, 0.3, 0.6, 0.4, 0.8, 0.5, 0.9])).astype(np.float32)
    total, width = forward_storage(depth, W, Lr, H, ca, table.astype(np.float64))
    s = tiny_state([0], river_width=[W], river_length=[Lr], river_height=[H], catchment_area=[ca],
                   flood_depth_table=table[None, :], river_storage=np.array([total]))
    model(s).kernel("flood_stage", time_step=1.0)
    assert s["flood_depth"][0] == pytest.approx(depth, rel=1e-5, abs=1e-5)
    # river + flood storage conserve the total (FP64)
    assert s["river_storage"][0] + s["flood_storage"][0] == pytest.approx(total, rel=1e-12)
    assert s["flood_fraction"][0] == pytest.approx(min((width - W) * Lr / ca, 1.0), rel=1e-4, abs=1e-6)
    # the channel holds its bank-full volume plus the water above it over the channel width
    assert s["river_storage"][0] == pytest.approx(Lr * W * (H + depth), rel=1e-5)
    if depth == 0.0:
        assert s["flood_fraction"][0] == 0.0


def test_stage_dry_depth_is_storage_over_area():
    s = tiny_state([0], river_storage=np.array([123456.0]))
    model(s).kernel("flood_stage", time_step=1.0)
    assert s["river_depth"][0] == pytest.approx(123456.0 / (10_000.0 * 50.0), rel=1e-6)
    assert s["flood_depth"][0] == 0.0 and s["flood_storage"][0] == 0.0


def test_storage_update_global_water_balance():
    """Sum of new storage = old + dt * (runoff) - dt * (mouth outflow): the
    reference's own inflow-error check (storage.py:305), no clipping active."""
    n = 6
    dn = [1, 2, 5, 4, 5, 5]
    q = np.array([3.0, 5.0, 2.0, 1.0, 4.0, 6.0], np.float32)
    inflow = np.zeros(n)
    for i in range(n):
        if dn[i] != i:
            inflow[dn[i]] += q[i]
    runoff = np.array([0.5, 0.1, 0.2, 0.3, 0.4, 0.6], np.float32)
    s = tiny_state(dn, river_outflow=q, river_inflow=inflow, runoff=runoff,
                   river_storage=np.full(n, 1e7))
    before = s["river_storage"].sum()
    dt = 100.0
    model(s).kernel("flood_stage", time_step=dt)
    after = s["river_storage"].sum() + s["flood_storage"].sum()
    expect = before + dt * float(runoff.astype(np.float64).sum()) - dt * float(q[5])
    assert after == pytest.approx(expect, rel=1e-12)


def test_storage_update_takes_a_river_deficit_from_the_flood_storage():
    """A river outflow larger than the river storage: the deficit is taken
    from the floodplain storage (storage.py:81-89), so the catchment keeps
    exactly old storage - outflow dt, and the river storage never goes
    negative."""
    dt = 1.0
    s = tiny_state([0], river_storage=np.array([100.0]), flood_storage=np.array([1000.0]),
                   river_outflow=np.array([150.0], np.float32))
    s = {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in s.items()}
    model(s, precision=1).kernel("flood_stage", time_step=dt)
    assert s["river_storage"][0] + s["flood_storage"][0] == pytest.approx(950.0, rel=1e-15)
    assert s["river_storage"][0] >= 0.0


# -------------------------------------------------------------------- CFL --
def test_cfl_closed_form():
    depth = np.array([0.0, 0.5, 3.0, 12.0], np.float32)
    dist = np.array([20_000.0, 3_000.0, 15_000.0, 800.0], np.float32)
    s = tiny_state([1, 2, 3, 3], river_depth=depth, downstream_distance=dist)
    m = O.OracleModel(s, flags=O.HAS_ADAPTIVE, num_flood_levels=10)
    T = 1800.0
    expect = 1
    for h, dx in zip(depth, dist):
        dt = 0.7 * float(dx) / math.sqrt(9.8 * max(float(h), 0.01))
        expect = max(expect, math.floor(T / min(dt, T) - 0.01) + 1)
    assert m.adaptive_steps(T) == expect


def test_cfl_offset_and_depth_floor():
    """n = floor(T / dt - 0.01) + 1 (adaptive_time.py:46): a ratio T / dt of
    6.004 gives 6 sub-steps, not 7; and the depth is floored at 0.01 m
    (:39): a 1 mm deep catchment counts as 1 cm deep."""
    T = 1800.0
    h = 1.0
    dx = T / 6.004 * math.sqrt(9.8 * h) / 0.7
    s = tiny_state([0], river_depth=np.array([h], np.float32), downstream_distance=np.array([dx], np.float32))
    assert O.OracleModel(s, flags=O.HAS_ADAPTIVE, num_flood_levels=10).adaptive_steps(T) == 6
    s = tiny_state([0], river_depth=np.array([0.001], np.float32),
                   downstream_distance=np.array([100.0], np.float32))
    dt = 0.7 * 100.0 / math.sqrt(9.8 * 0.01)
    assert O.OracleModel(s, flags=O.HAS_ADAPTIVE, num_flood_levels=10).adaptive_steps(T) == \
        math.floor(T / dt - 0.01) + 1 == 9


# ------------------------------------------------------------ bifurcation --
def bif_state(P=8, seed=4):
    rng = np.random.default_rng(seed)
    n = 12
    dn = list(range(1, n)) + [n - 1]
    s = tiny_state(dn, river_depth=rng.uniform(0.5, 3, n).astype(np.float32),
                   catchment_elevation=np.linspace(30, 0, n).astype(np.float32),
                   river_storage=rng.uniform(1e5, 1e7, n))
    ci = rng.integers(0, n, P).astype(np.int32)
    di = ((ci + 1 + rng.integers(0, n - 1, P)) % n).astype(np.int32)
    Lb = 5
    s.update({
        "water_surface_elevation": np.zeros(n, np.float32), "total_storage": np.zeros(n),
        "bifurcation_catchment_idx": ci, "bifurcation_downstream_idx": di,
        "bifurcation_manning": np.full((P, Lb), 0.03, np.float32),
        "bifurcation_width": rng.uniform(20, 200, (P, Lb)).astype(np.float32),
        "bifurcation_length": np.full(P, 5000.0, np.float32),
        "bifurcation_elevation": (np.linspace(30, 0, n)[ci][:, None] - 4.0
                                  + np.array([0, 0.5, 1, 2, 3])[None, :]).astype(np.float32),
        "bifurcation_outflow": np.zeros((P, Lb), np.float32),
        "bifurcation_cross_section_depth": np.zeros((P, Lb), np.float32),
        "global_bifurcation_outflow": np.zeros(n), "limit_rate": np.ones(n, np.float32),
    })
    return s, n, P, Lb


def test_bifurcation_net_transfer_sums_to_zero():
    s, n, P, Lb = bif_state()
    flags = O.HAS_BIFURCATION | O.HAS_TOTAL_STORAGE | O.HAS_WATER_SURFACE
    m = O.OracleModel(s, flags=flags, num_flood_levels=10, num_bifurcation_levels=Lb)
    for _ in range(3):
        m.kernel("outflow", time_step=300.0)
        m.kernel("bifurcation_outflow", time_step=300.0)
        m.kernel("inflow")
        m.kernel("bifurcation_inflow")
        assert s["global_bifurcation_outflow"].sum() == 0.0    # antisymmetric pairs
        m.kernel("flood_stage", time_step=300.0)
    assert np.abs(s["bifurcation_outflow"]).sum() > 0


def test_bifurcation_first_step_closed_form():
    s, n, P, Lb = bif_state(P=1)
    flags = O.HAS_BIFURCATION | O.HAS_TOTAL_STORAGE | O.HAS_WATER_SURFACE
    m = O.OracleModel(s, flags=flags, num_flood_levels=10, num_bifurcation_levels=Lb)
    dt = 30.0
    m.kernel("outflow", time_step=dt)
    wse = s["water_surface_elevation"].astype(np.float64)
    ci, di = int(s["bifurcation_catchment_idx"][0]), int(s["bifurcation_downstream_idx"][0])
    slope = np.clip((wse[ci] - wse[di]) / 5000.0, -0.005, 0.005)
    m.kernel("bifurcation_outflow", time_step=dt)
    for lev in range(Lb):
        d = max(max(wse[ci], wse[di]) - float(s["bifurcation_elevation"][0, lev]), 0.0)
        h = math.sqrt(d * 0.01)  # zero previous depth: the 0.01 floor term
        q = float(s["bifurcation_width"][0, lev]) * G * dt * h * slope if h > 1e-5 else 0.0
        lim = min(0.05 * min(s["total_storage"][ci], s["total_storage"][di]) / (abs(q) * dt + 1e-300), 1.0)
        assert s["bifurcation_cross_section_depth"][0, lev] == pytest.approx(d, rel=1e-5, abs=1e-6)
        assert abs(s["bifurcation_outflow"][0, lev]) <= abs(q) * (1 + 1e-4) + 1e-6
        del lim


def test_bifurcation_limited_by_the_smaller_end_storage():
    """A path may move at most 5 % of the smaller of its two end storages in
    one sub-step (bifurcation.py:93-99): with the ends holding x and 100 x,
    a path flow between those two bounds is scaled to 0.05 x / dt."""
    flags = O.HAS_BIFURCATION | O.HAS_TOTAL_STORAGE | O.HAS_WATER_SURFACE
    dt = 300.0

    def path_sum(store_ci, store_di):
        s, n, P, Lb = bif_state(P=1)
        ci, di = int(s["bifurcation_catchment_idx"][0]), int(s["bifurcation_downstream_idx"][0])
        rs = s["river_storage"].copy()
        rs[ci], rs[di] = store_ci, store_di
        s["river_storage"] = rs
        m = O.OracleModel(s, flags=flags, num_flood_levels=10, num_bifurcation_levels=Lb)
        m.kernel("outflow", time_step=dt)
        m.kernel("bifurcation_outflow", time_step=dt)
        return float(s["bifurcation_outflow"][0].astype(np.float64).sum())

    free = path_sum(1e12, 1e12)
    assert abs(free) > 0
    x = abs(free) * dt / 0.05 / 10.0          # 0.05 x < |free| dt < 0.05 (100 x)
    for a, b in ((x, 100 * x), (100 * x, x)):
        assert abs(path_sum(a, b)) * dt == pytest.approx(0.05 * x, rel=1e-5)


# ------------------------------------------------------------------ levee --
def levee_state(total, base=1.0, crown=3.0, frac=0.3):
    s = tiny_state([0], river_storage=np.array([total]))
    s.update({"is_levee": np.array([True]), "levee_catchment_idx": np.array([0], np.int32),
              "levee_base_height": np.array([base], np.float32),
              "levee_crown_height": np.array([crown], np.float32),
              "levee_fraction": np.array([frac], np.float32)})
    return s


@pytest.mark.parametrize("total", [1e6, 2.5e6, 2e7, 6e7, 2e8, 9e8])
def test_levee_partition_conserves_the_total(total):
    s = levee_state(total)
    m = O.OracleModel(s, flags=O.HAS_LEVEE, num_flood_levels=10)
    m.kernel("flood_stage", time_step=1.0)
    before = s["river_storage"][0] + s["flood_storage"][0]
    rd, fd, ff = s["river_depth"][0], s["flood_depth"][0], s["flood_fraction"][0]
    m.kernel("levee_stage")
    after = s["river_storage"][0] + s["flood_storage"][0] + s["protected_storage"][0]
    assert after == pytest.approx(before, rel=1e-14)
    assert min(s["river_storage"][0], s["flood_storage"][0], s["protected_storage"][0]) >= 0
    if before <= 10_000 * 50 * 4.0:   # below bankfull: case 1 leaves the state as the stage left it
        assert (s["river_depth"][0], s["flood_depth"][0], s["flood_fraction"][0]) == (rd, fd, ff)
        assert s["protected_storage"][0] == 0.0


def levee_outputs(total, precision=1):
    s = levee_state(total)
    if precision == 1:
        s = {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in s.items()}
    m = O.OracleModel(s, flags=O.HAS_LEVEE, num_flood_levels=10, precision=precision)
    m.kernel("flood_stage", time_step=1.0)
    m.kernel("levee_stage")
    return (float(s["river_storage"][0]), float(s["flood_storage"][0]), float(s["protected_storage"][0]),
            float(s["flood_depth"][0]), float(s["protected_depth"][0]), float(s["flood_fraction"][0]))


def profile_storage(h, rl=10_000.0, rw=50.0, rh=4.0, ca=5e8, layer=0.5, levels=10):
    """Storage at flood depth h over the whole floodplain of levee_state's
    catchment, integrated layer by layer from the geometry (channel full,
    each 0.5 m layer a trapezoid between the widths at its bottom and top)."""
    winc = ca / rl / levels
    s = rl * rw * rh
    for k in range(levels):
        lo, hi = k * layer, (k + 1) * layer
        if h <= lo:
            break
        top = min(h, hi)
        w_lo = rw + k * winc
        w_top = w_lo + winc * (top - lo) / layer
        s += rl * 0.5 * (w_lo + w_top) * (top - lo)
    return s


@pytest.mark.parametrize("total", [1.0e8, 2.0e8, 3.0e8])
def test_levee_river_side_fill_closed_form(total):
    """Case 2 (water between levee base and crown on the river side): above
    the base the water spreads only over the channel and the floodplain up
    to the levee, so the depth is base + (total - S(base)) / ((ld + W) L)
    with S(base) integrated from the profile; the protected side stays
    dry, the channel holds its column up to that depth, and the flooded
    fraction is the levee's position."""
    base, crown, rl, rw = 1.0, 3.0, 10_000.0, 50.0
    frac = float(np.float32(0.3))          # the fraction as the state stores it
    ld = frac * 5e8 / rl
    s_base = profile_storage(base)
    assert s_base < total < s_base + (ld + rw) * (crown - base) * rl
    r, f, p, fdep, pdep, ffr = levee_outputs(total)
    depth = base + (total - s_base) / ((ld + rw) * rl)
    assert fdep == pytest.approx(depth, rel=1e-12)
    assert p == 0.0 and pdep == 0.0
    assert r == pytest.approx(rl * rw * (4.0 + depth), rel=1e-12)
    assert r + f == pytest.approx(total, rel=1e-14)
    assert ffr == pytest.approx(frac, rel=1e-6)


def test_levee_cases_are_continuous_and_monotone():
    """Sweeping the total through all four cases (below bankfull, river side
    filling, protected side filling, overtopped): the flood depth, the
    protected storage and the river + flood storage never decrease, none of
    them jumps, the total is conserved, and the protected side first holds
    water exactly above S(base) + (ld + W)(crown - base) L."""
    totals = np.linspace(1e6, 1.2e9, 4001)
    prev = O.set_threads(1)          # 8,002 one-catchment kernels: no parallel regions
    try:
        out = np.array([levee_outputs(t) for t in totals])
    finally:
        O.set_threads(prev)
    r, f, p, fdep = out[:, 0], out[:, 1], out[:, 2], out[:, 3]
    assert np.allclose(r + f + p, totals, rtol=1e-13)
    step = totals[1] - totals[0]
    for x in (fdep, p, r + f):
        d = np.diff(x)
        assert np.all(d >= -1e-9 * np.abs(x[1:]).max()), "not monotone"
    # no jumps: a step of the total moves a storage by at most that step, and
    # the flood depth by at most the step over the narrowest wetted width
    assert np.all(np.abs(np.diff(p)) <= step * (1 + 1e-9))
    assert np.all(np.abs(np.diff(r + f)) <= step * (1 + 1e-9))
    assert np.all(np.abs(np.diff(fdep)) <= step / (50.0 * 10_000.0) * (1 + 1e-9))
    s_top = profile_storage(1.0) + (float(np.float32(0.3)) * 5e8 / 10_000.0 + 50.0) * 2.0 * 10_000.0
    wet_protected = totals[p > 0]
    assert wet_protected.size and wet_protected.min() >= s_top * (1 - 1e-12)
    assert wet_protected.min() <= s_top + 2 * step
    assert fdep.max() > 3.0                     # the sweep reaches past the crown


# -------------------------------------------------------------- reservoir --
def reservoir_state(total, inflow, cv=4e7, ev=9.5e7, qn=50.0, qf=150.0):
    n = 3
    s = tiny_state([1, 2, 2], river_storage=np.array([1e6, total, 1e6]),
                   runoff=np.zeros(n, np.float32))
    fv = (ev - cv) / 0.95
    av = cv + fv * 0.1
    qa = (qn + qf) * 0.5
    s.update({
        "total_storage": np.zeros(n), "reservoir_catchment_idx": np.array([1], np.int32),
        "conservation_volume": np.array([cv], np.float32), "emergency_volume": np.array([ev], np.float32),
        "adjustment_volume": np.array([av], np.float32),
        "effective_normal_outflow": np.array([qn], np.float32),
        "adjustment_outflow": np.array([qa], np.float32),
        "flood_control_outflow": np.array([qf], np.float32),
        "reservoir_total_inflow": np.array([0.0, inflow, 0.0]),
        "is_reservoir": np.array([False, True, False]), "is_dam_upstream": np.array([True, False, False]),
        "is_dam_related": np.array([True, True, False]),
    })
    return s, dict(cv=cv, ev=ev, av=av, qn=qn, qa=qa, qf=qf)


@pytest.mark.parametrize("total,inflow", [(1e7, 10.0), (4.2e7, 10.0), (6e7, 10.0), (6e7, 400.0),
                                          (9.9e7, 10.0), (9.9e7, 400.0)])
def test_reservoir_release_regimes(total, inflow):
    """Piecewise rule of reservoir.py:114-163 evaluated in float64."""
    s, p = reservoir_state(total, inflow)
    flags = O.HAS_RESERVOIR | O.HAS_TOTAL_STORAGE
    m = O.OracleModel(s, flags=flags, num_flood_levels=10)
    dt = 600.0
    m.kernel("outflow", time_step=dt)
    m.kernel("reservoir", time_step=dt)
    S = total
    if S <= p["cv"]:
        q = p["qn"] * math.sqrt(S / p["cv"])
    elif S <= p["av"]:
        q = p["qn"] + ((S - p["cv"]) / (p["av"] - p["cv"])) ** 3 * (p["qa"] - p["qn"])
    elif S <= p["ev"]:
        tmp = p["qa"] + ((S - p["av"]) / (p["ev"] - p["av"])) ** 0.1 * (p["qf"] - p["qa"])
        if inflow >= p["qf"]:
            q = max(p["qn"] + (S - p["cv"]) / (p["ev"] - p["cv"]) * (inflow - p["qn"]), tmp)
        else:
            q = tmp
    else:
        q = inflow if inflow >= p["qf"] else p["qf"]
    q = min(max(q, 0.0), S / dt)
    assert s["river_outflow"][1] == pytest.approx(q, rel=2e-5)
    assert s["flood_outflow"][1] == 0.0
    assert s["reservoir_total_inflow"][1] == 0.0           # read and zeroed
    # outgoing of the dam catchment is exactly its release
    assert s["outgoing_storage"][1] == pytest.approx(float(s["river_outflow"][1]) * dt, rel=1e-7)


def test_reservoir_rule_is_continuous_at_the_thresholds():
    outs = []
    for S in (4e7 * (1 - 1e-6), 4e7 * (1 + 1e-6)):
        s, _ = reservoir_state(S, 10.0)
        m = O.OracleModel(s, flags=O.HAS_RESERVOIR | O.HAS_TOTAL_STORAGE, num_flood_levels=10)
        m.kernel("outflow", time_step=600.0)
        m.kernel("reservoir", time_step=600.0)
        outs.append(float(s["river_outflow"][1]))
    assert outs[0] == pytest.approx(outs[1], rel=1e-3)


# ------------------------------------------- levee bifurcation variant --
def levee_path_state(csd_prev, w=(20.0, 19.0), pw=(18.5, 18.2), bel=(15.0, 18.3, 18.4, 19.5, 30.0)):
    """One path between catchments 0 and 1 with water surfaces `w`,
    protected surfaces `pw` (a protected side lower than the river) and
    huge storages, so that the path limiter stays at 1."""
    s = tiny_state([1, 1], river_storage=np.array([1e12, 1e12]))
    Lb = len(bel)
    s.update({
        "water_surface_elevation": np.array(w, np.float32),
        "protected_water_surface_elevation": np.array(pw, np.float32),
        "total_storage": np.array([1e12, 1e12]),
        "is_levee": np.array([True, True]),
        "bifurcation_catchment_idx": np.array([0], np.int32),
        "bifurcation_downstream_idx": np.array([1], np.int32),
        "bifurcation_manning": np.full((1, Lb), 0.03, np.float32),
        "bifurcation_width": np.array([[50.0, 60.0, 70.0, 80.0, 90.0]], np.float32),
        "bifurcation_length": np.array([4000.0], np.float32),
        "bifurcation_elevation": np.array([bel], np.float32),
        "bifurcation_outflow": np.zeros((1, Lb), np.float32),
        "bifurcation_cross_section_depth": np.array([csd_prev], np.float32),
        "global_bifurcation_outflow": np.zeros(2), "limit_rate": np.ones(2, np.float32),
    })
    return s


@pytest.mark.parametrize("prec", [0, 1])
def test_levee_path_depths(prec):
    """With levees the path levels split (triton/levee.py:631-653): level 0
    flows on the river surfaces with the semi-implicit depth, falling back to
    the explicit depth when the previous depth is zero; levels above flow on
    the higher protected surface with the explicit depth, whatever their
    previous depth.  From rest q = W g dt h S, S from the river surfaces."""
    dt = 20.0
    w, pw = (20.0, 19.0), (18.5, 18.2)
    bel = (15.0, 18.3, 18.4, 19.5, 30.0)
    S = (w[0] - w[1]) / 4000.0
    widths = (50.0, 60.0, 70.0, 80.0, 90.0)
    for c0 in (0.0, 2.0):
        prev = [c0, 5.0, 5.0, 5.0, 5.0]
        s = levee_path_state(prev, w, pw, bel)
        if prec == 1:
            s = {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in s.items()}
        m = O.OracleModel(s, flags=O.HAS_BIFURCATION | O.HAS_LEVEE, num_flood_levels=10,
                          num_bifurcation_levels=5, precision=prec)
        m.kernel("levee_bifurcation_outflow", time_step=dt)
        # the inputs as stored (float32 values, widened in the FP64 profile)
        w32 = [float(np.float32(x)) for x in w]
        pw32 = [float(np.float32(x)) for x in pw]
        b32 = [float(np.float32(x)) for x in bel]
        S = (w32[0] - w32[1]) / 4000.0
        for lev in range(5):
            top = max(w32) if lev == 0 else max(pw32)
            upd = max(top - b32[lev], 0.0)
            h = upd if lev > 0 or c0 == 0.0 else math.sqrt(upd * c0)
            q = widths[lev] * G * dt * h * S if h > 1e-5 else 0.0
            rel = 2e-6 if prec == 0 else 1e-12
            assert float(s["bifurcation_cross_section_depth"][0, lev]) == pytest.approx(upd, rel=rel, abs=1e-6)
            assert float(s["bifurcation_outflow"][0, lev]) == pytest.approx(q, rel=rel, abs=1e-9), (c0, lev)
        # level 2 lies below the river surfaces but above the protected ones:
        # it carries no flow in the levee variant
        assert float(s["bifurcation_outflow"][0, 3]) == 0.0
        assert float(s["bifurcation_outflow"][0, 2]) > 0.0


# ----------------------------------------- kinematic flow above a dam --
def dam_upstream_state(elev0, elev1, rs, fs, r_dep=2.0, f_dep=0.4, dist=8000.0):
    s = tiny_state([1, 1], river_storage=np.array([rs, 1e9]), flood_storage=np.array([fs, 0.0]),
                   river_depth=np.array([r_dep, 3.0], np.float32),
                   flood_depth=np.array([f_dep, 0.0], np.float32),
                   catchment_elevation=np.array([elev0, elev1], np.float32),
                   downstream_distance=np.array([dist, dist], np.float32),
                   river_cross_section_depth=np.array([r_dep, 3.0], np.float32),
                   river_outflow=np.array([123.0, 0.0], np.float32),
                   flood_outflow=np.array([45.0, 0.0], np.float32))
    s["is_dam_upstream"] = np.array([True, False])
    return s


def kinematic_expect(elev0, elev1, rs, fs, dt, r_dep=2.0, f_dep=0.4, dist=8000.0,
                     W=50.0, L=10_000.0, n_r=0.03, n_f=0.1, min_kin=1e-5):
    S = max((elev0 - elev1) / dist, min_kin)
    q_r = W * r_dep * (1.0 / n_r) * math.sqrt(S) * r_dep ** (2.0 / 3.0)
    q_r = max(min(q_r, rs / dt), 0.0)
    Sf = min(S, 0.005)
    area = max(fs / L - f_dep * W, 0.0)
    q_f = area * (1.0 / n_f) * math.sqrt(Sf) * f_dep ** (2.0 / 3.0)
    q_f = max(min(q_f, fs / dt), 0.0)
    return q_r, q_f


@pytest.mark.parametrize("case", ["moderate", "uphill", "steep", "capped"])
def test_kinematic_flow_upstream_of_a_dam(case):
    """Upstream of a dam the momentum update is replaced by Manning's
    kinematic flow on the bed slope (triton/outflow.py:251-272):
    Q = W d n^-1 S^(1/2) d^(2/3) with S the bed slope floored at the minimum
    kinematic slope, the floodplain on the same slope capped at 0.005, each
    flow capped at storage / dt and non-negative; the previous flows and the
    water-surface slope play no part."""
    elev0, elev1, rs, fs, f_dep, dt = {
        "moderate": (18.0, 10.0, 1e9, 5e7, 0.4, 600.0),     # bed slope 1e-3
        "uphill": (10.0, 12.0, 1e9, 5e7, 0.4, 600.0),       # negative: floored at 1e-5
        "steep": (170.0, 10.0, 1e9, 5e7, 0.4, 600.0),       # 0.02: floodplain on 0.005
        "capped": (170.0, 10.0, 2e4, 1e8, 8.0, 7200.0),     # storage / dt binds
    }[case]
    s = dam_upstream_state(elev0, elev1, rs, fs, f_dep=f_dep)
    m = model(s, flags=O.HAS_RESERVOIR)
    m.kernel("outflow", time_step=dt)
    e32 = [float(np.float32(x)) for x in (elev0, elev1)]
    q_r, q_f = kinematic_expect(e32[0], e32[1], rs, fs, dt, f_dep=f_dep)
    assert float(s["river_outflow"][0]) == pytest.approx(q_r, rel=1e-5)
    assert float(s["flood_outflow"][0]) == pytest.approx(q_f, rel=1e-5)
    if case == "capped":
        assert float(s["river_outflow"][0]) == pytest.approx(rs / dt, rel=1e-6)
        assert float(s["flood_outflow"][0]) == pytest.approx(fs / dt, rel=1e-6)
    # not above a dam: the same catchment runs the momentum update
    s2 = dam_upstream_state(elev0, elev1, rs, fs, f_dep=f_dep)
    s2["is_dam_upstream"] = np.array([False, False])
    model(s2, flags=O.HAS_RESERVOIR).kernel("outflow", time_step=dt)
    assert float(s2["river_outflow"][0]) != pytest.approx(q_r, rel=1e-3)


# ---------------------------------------------- water-balance log sums --
def log_network(n=60, seed=11, levee_every=0, clip=False):
    rng = np.random.default_rng(seed)
    dn = np.array([min(n - 1, i + 1 + int(rng.integers(0, 4))) for i in range(n)], np.int32)
    dn[n - 1] = n - 1
    dn[n - 5] = n - 5
    rs = rng.uniform(2e6, 4e6, n)
    if levee_every:
        rs = rng.uniform(1e6, 8e8, n)       # every levee case, and dry catchments
    qr = rng.uniform(0.0, 20.0, n).astype(np.float32)
    qf = rng.uniform(0.0, 5.0, n).astype(np.float32)
    if clip:
        qr[3] = np.float32(1e6)             # far more than the catchment holds
    rin = np.zeros(n)
    fin = np.zeros(n)
    for i in range(n):
        if dn[i] != i:
            rin[dn[i]] += qr[i]
            fin[dn[i]] += qf[i]
    s = tiny_state(dn, river_outflow=qr, flood_outflow=qf, river_inflow=rin, flood_inflow=fin,
                   river_storage=rs, flood_storage=rng.uniform(0.0, 1e5, n),
                   protected_storage=np.where(np.arange(n) % 7 == 0, 1e4, 0.0),
                   runoff=rng.uniform(0.0, 3.0, n).astype(np.float32),
                   catchment_area=rng.uniform(1e8, 1e9, n).astype(np.float32))
    s["log_sums"] = np.zeros((11, 4))
    flags = O.HAS_LOG
    if levee_every:
        lv = np.arange(0, n, levee_every).astype(np.int32)
        isl = np.zeros(n, bool)
        isl[lv] = True
        s.update({"is_levee": isl, "levee_catchment_idx": lv,
                  "levee_base_height": np.full(lv.size, 1.0, np.float32),
                  "levee_crown_height": np.full(lv.size, 3.0, np.float32),
                  "levee_fraction": np.full(lv.size, 0.3, np.float32)})
        flags |= O.HAS_LEVEE
    s = {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in s.items()}
    return s, flags


# log slots (cmfgpu/modules/log.py:207-326 field order)
PRE, NEXT, NEW, IN_ERR, INFLOW, OUTFLOW, STAGE, STAGE_ERR, RIVER, FLOOD, AREA = range(11)


def test_log_sums_against_brute_force():
    """The eleven FP64 sums of compute_flood_stage_log (triton/storage.py:
    276-305, 395-414), each against a brute-force float64 sum of the state
    before and after, in units of 1e9 m3, added into slot current_step."""
    s, flags = log_network()
    old = {k: v.copy() for k, v in s.items()}
    s["log_sums"][:, 2] = 7.0                # the slot accumulates
    dt = 300.0
    m = O.OracleModel(s, flags=flags, num_flood_levels=10, log_buffer_size=4, precision=1)
    m.kernel("flood_stage", time_step=dt, current_step=2)
    got = s["log_sums"][:, 2] - 7.0
    assert np.all(s["log_sums"][:, [0, 1, 3]] == 0.0)
    pre = old["river_storage"] + old["flood_storage"] + old["protected_storage"]
    new = s["river_storage"] + s["flood_storage"] + s["protected_storage"]
    inflow = (old["river_inflow"] + old["flood_inflow"]) * dt
    outflow = (old["river_outflow"] + old["flood_outflow"]) * dt
    runoff = old["runoff"] * dt
    tol = dict(rel=1e-12)
    assert got[PRE] == pytest.approx(pre.sum() * 1e-9, **tol)
    assert got[INFLOW] == pytest.approx(inflow.sum() * 1e-9, **tol)
    assert got[OUTFLOW] == pytest.approx(outflow.sum() * 1e-9, **tol)
    # no clipping: next = pre + inflow + runoff - outflow = new
    expect_next = (pre + inflow + runoff - outflow).sum() * 1e-9
    assert got[NEXT] == pytest.approx(expect_next, **tol)
    assert got[NEW] == pytest.approx(new.sum() * 1e-9, **tol)
    assert got[STAGE] == pytest.approx(new.sum() * 1e-9, **tol)
    assert got[RIVER] == pytest.approx(s["river_storage"].sum() * 1e-9, **tol)
    assert got[FLOOD] == pytest.approx(s["flood_storage"].sum() * 1e-9, **tol)
    assert got[AREA] == pytest.approx((s["flood_fraction"] * s["catchment_area"]).sum() * 1e-9, **tol)
    scale = pre.sum() * 1e-9
    assert abs(got[IN_ERR]) <= 1e-13 * scale and abs(got[STAGE_ERR]) <= 1e-13 * scale


def test_log_inflow_error_holds_the_clipped_deficit():
    """A river outflow far beyond the catchment's storage is clipped at zero
    storage (storage.py:81-94): the inflow-error sum is then pre - next +
    net inflow dt, i.e. minus the water that was asked for but not there."""
    s, flags = log_network(clip=True)
    old = {k: v.copy() for k, v in s.items()}
    dt = 300.0
    m = O.OracleModel(s, flags=flags, num_flood_levels=10, log_buffer_size=4, precision=1)
    m.kernel("flood_stage", time_step=dt, current_step=0)
    c = 3
    net = (old["river_inflow"][c] + old["flood_inflow"][c] + old["runoff"][c]
           - old["river_outflow"][c] - old["flood_outflow"][c]) * dt
    pre_c = old["river_storage"][c] + old["flood_storage"][c] + old["protected_storage"][c]
    # the clipped catchment keeps its protected storage and receives its runoff
    next_c = old["protected_storage"][c] + old["runoff"][c] * dt
    deficit = pre_c - next_c + net           # < 0: the water that was not there
    assert deficit < -1e8
    assert s["log_sums"][IN_ERR, 0] == pytest.approx(deficit * 1e-9, rel=1e-9)


def test_log_stage_sums_count_levee_catchments_once():
    """compute_flood_stage_log leaves levee catchments out of the stage sums
    (storage.py:251-255) and compute_levee_stage_log adds them after the
    partition (levee.py:549-553): over both, the stage sum equals the new
    total storage of every catchment, its error is zero, and the river and
    flood storage sums leave exactly the protected storage out."""
    s, flags = log_network(levee_every=3)
    dt = 300.0
    m = O.OracleModel(s, flags=flags, num_flood_levels=10, log_buffer_size=4, precision=1)
    m.kernel("flood_stage", time_step=dt, current_step=1)
    lv = s["levee_catchment_idx"]
    non = np.ones(s["river_storage"].size, bool)
    non[lv] = False
    after_stage = s["log_sums"][:, 1].copy()
    assert after_stage[STAGE] == pytest.approx((s["river_storage"] + s["flood_storage"])[non].sum() * 1e-9,
                                               rel=1e-12)
    m.kernel("levee_stage", current_step=1)
    got = s["log_sums"][:, 1]
    new = s["river_storage"] + s["flood_storage"] + s["protected_storage"]
    assert s["protected_storage"][lv].sum() > 0            # some levee holds water behind it
    assert got[STAGE] == pytest.approx(new.sum() * 1e-9, rel=1e-12)
    assert got[STAGE] == pytest.approx(got[NEW], rel=1e-12)
    assert abs(got[STAGE_ERR]) <= 1e-13 * got[NEW]
    assert got[STAGE] - got[RIVER] - got[FLOOD] == pytest.approx(s["protected_storage"].sum() * 1e-9, rel=1e-9)
    assert got[AREA] == pytest.approx((s["flood_fraction"] * s["catchment_area"]).sum() * 1e-9, rel=1e-12)


# --------------------------------------------- path nets and rescale --
def test_path_levels_rescaled_by_the_end_limiters():
    """compute_bifurcation_inflow (triton/bifurcation.py:128-151): a level
    flowing from the upstream end is scaled by that end's storage limiter,
    one flowing back by the downstream end's; the net of the raw level sum
    likewise, added at the upstream end and taken at the downstream end."""
    s, n, P, Lb = bif_state(P=2)
    s["bifurcation_catchment_idx"] = np.array([2, 6], np.int32)
    s["bifurcation_downstream_idx"] = np.array([5, 9], np.int32)
    raw = np.array([[3.0, -1.0, 2.0, -4.0, 0.5], [-6.0, 1.0, -0.25, 2.0, 0.0]], np.float32)
    s["bifurcation_outflow"] = raw.copy()
    lim = np.ones(n, np.float32)
    lim[2], lim[5], lim[6], lim[9] = 0.5, 0.25, 0.75, 0.125
    s["limit_rate"] = lim
    m = O.OracleModel(s, flags=O.HAS_BIFURCATION, num_flood_levels=10, num_bifurcation_levels=Lb)
    m.kernel("bifurcation_inflow")
    for p, (a, b) in enumerate(((2, 5), (6, 9))):
        want = np.where(raw[p] >= 0, raw[p] * lim[a], raw[p] * lim[b])
        assert np.array_equal(s["bifurcation_outflow"][p], want.astype(np.float32))
        tot = float(raw[p].astype(np.float64).sum())
        net = tot * (lim[a] if tot >= 0 else lim[b])
        assert s["global_bifurcation_outflow"][a] == net
        assert s["global_bifurcation_outflow"][b] == -net
    assert s["global_bifurcation_outflow"].sum() == 0.0


def test_fp32_storage_profile_rounds_the_storage_update():
    """Without mixed precision the storages are float32 (the oracle's
    precision 2): above 2^24 m3 a 1 m3 gain is lost to rounding, which the
    FP64 storage of the mixed profile keeps; with exactly representable
    values both profiles give the same storage."""
    out = {}
    for prec in (0, 2):
        for start, q in ((16777216.0, 1.0), (1024.0, 3.0)):
            # a 40 m deep channel: 2e7 m3 bankfull, all the water in the river
            s = tiny_state([0], river_storage=np.array([start]), river_inflow=np.array([q]),
                           river_height=np.array([40.0], np.float32))
            if prec == 2:
                for k in ("river_storage", "flood_storage", "protected_storage", "outgoing_storage",
                          "river_inflow", "flood_inflow"):
                    s[k] = s[k].astype(np.float32)
            model(s, precision=prec).kernel("flood_stage", time_step=1.0)
            out[prec, start] = float(s["river_storage"][0]) + float(s["flood_storage"][0])
    assert out[0, 16777216.0] == 16777217.0
    assert out[2, 16777216.0] == 16777216.0
    assert out[0, 1024.0] == out[2, 1024.0] == 1027.0
