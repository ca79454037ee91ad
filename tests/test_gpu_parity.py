"""HIP path vs CPU oracle on the same seeded inputs (bit-exact).

The oracle and the kernels evaluate the same IEEE operation sequence with the
same fixed accumulation order, so every field must agree bit for bit after any
number of steps, in both device layouts and both schedules.
"""
import numpy as np
import pytest

from helpers import FULL, STATE_FIELDS, mismatch, oracle_for, small_source

pytestmark = pytest.mark.gpu


def _model(src, modules, **kw):
    from cmfgpu_amd.models import CaMaFlood
    return CaMaFlood(src, opened_modules=modules, device="cuda:0", step_duration=1800.0, **kw)


def _compare(model, orc, fields):
    import torch
    torch.cuda.synchronize()
    for name in fields:
        if name not in orc.state:
            continue
        msg = mismatch(model.get(name), orc.state[name])
        assert msg is None, f"{name}: {msg}"


@pytest.mark.parametrize("layout", ["basin", "level", "network", "keyed"])
@pytest.mark.parametrize("mode", ["fused", "ops"])
def test_plain_bitexact(layout, mode):
    net, src = small_source(paths=0, levee=0.0, reservoirs=0)
    mods = ("base", "adaptive_time")
    model = _model(src, mods, layout=layout, mode=mode)
    orc, _ = oracle_for(src, mods)
    for step in range(6):
        n_gpu = model.step_advance()
        n_cpu = int(orc.advance(1800.0, 1)[0])
        assert n_gpu == n_cpu
    _compare(model, orc, STATE_FIELDS)


@pytest.mark.parametrize("layout", ["basin", "level", "network", "keyed"])
@pytest.mark.parametrize("mode", ["fused", "ops"])
def test_full_physics_bitexact(layout, mode):
    net, src = small_source()
    model = _model(src, FULL, layout=layout, mode=mode)
    orc, _ = oracle_for(src, FULL)
    for step in range(6):
        assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    _compare(model, orc, STATE_FIELDS + ("bifurcation_outflow", "bifurcation_cross_section_depth",
                                         "global_bifurcation_outflow", "limit_rate", "total_storage",
                                         "water_surface_elevation", "protected_water_surface_elevation",
                                         "reservoir_total_inflow"))


@pytest.mark.parametrize("layout", ["basin", "network", "keyed"])
@pytest.mark.parametrize("mode", ["fused", "ops"])
def test_precise_surfaces_bitexact(layout, mode):
    """FP32 with water surfaces formed in double (CMF_PRECISE_SURFACE)
    against the oracle with the same flag; levee catchments started at 1-100
    times their river storage so that the protected surfaces are in play."""
    net, src = small_source()
    rs = np.array(src["river_storage"], dtype=np.float64)
    lid = np.asarray(src["levee_catchment_idx"])
    rs[lid] *= np.random.default_rng(3).uniform(1.0, 100.0, lid.size)
    src["river_storage"] = rs
    model = _model(src, FULL, layout=layout, mode=mode, precise_surfaces=True)
    orc, _ = oracle_for(src, FULL, precise_surfaces=True)
    ref, _ = oracle_for(src, FULL)
    for step in range(6):
        n = model.step_advance()
        assert n == int(orc.advance(1800.0, 1)[0])
        ref.advance(1800.0, 1, n)
    _compare(model, orc, STATE_FIELDS + ("bifurcation_outflow", "bifurcation_cross_section_depth",
                                         "global_bifurcation_outflow", "limit_rate", "total_storage",
                                         "water_surface_elevation", "protected_water_surface_elevation",
                                         "reservoir_total_inflow"))
    # the flag is in effect: the reference's FP32 surfaces give other flows
    assert not np.array_equal(orc.state["river_outflow"], ref.state["river_outflow"])


def test_fp64_profile_bitexact():
    net, src = small_source()
    model = _model(src, FULL, precision="float64")
    orc, _ = oracle_for(src, FULL, precision="float64")
    for step in range(4):
        assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    _compare(model, orc, STATE_FIELDS)


def test_log_sums():
    net, src = small_source()
    mods = FULL + ("log",)
    model = _model(src, mods, log_buffer_size=64)
    orc, _ = oracle_for(src, mods, log_buffer_size=64)
    n = model.step_advance()
    sums_gpu = model.t["log_sums"].cpu().numpy()  # step_advance zeroes after writing; read a fresh step
    m2 = _model(src, mods, log_buffer_size=64)
    import ctypes
    from cmfgpu_amd import _lib
    cnt = ctypes.c_int32(0)
    _lib.check(_lib.lib().cmf_hip_advance(m2.network_handle, m2.field_table_ref(), 1800.0, 1,
                                          ctypes.byref(cnt), _lib.current_stream()))
    n_cpu = int(orc.advance(1800.0, 1)[0])
    assert cnt.value == n_cpu == n
    g = m2.log_sums()[:, :n_cpu]
    c = orc.state["log_sums"][:, :n_cpu]
    # FP64 sums in a different (tree) order: relative 1e-12 of the magnitude
    scale = np.maximum(np.abs(c), 1e-9)
    assert np.all(np.abs(g - c) <= 1e-12 * scale + 1e-15), (g, c)
    del sums_gpu


def _member_sources(src, members, seed0=4000):
    from cmfgpu_amd import synthetic
    return [synthetic.Runoff(src["catchment_area"], seed0 + m).next_day() for m in range(members)]


@pytest.mark.parametrize("trial_layout", ["inner", "major"])
@pytest.mark.parametrize("modules", ["plain", "full"])
def test_batched_members_match_standalone_runs(trial_layout, modules):
    """Each member of a batched run with per-member CFL steps is bit-exact
    with a standalone oracle run on that member's runoff."""
    from cmfgpu_amd import _lib
    mods = ("base", "adaptive_time") if modules == "plain" else FULL
    net, src = small_source() if modules == "full" else small_source(paths=0, levee=0.0, reservoirs=0)
    M = 4
    runoffs = _member_sources(src, M)
    model = _model(src, mods, num_trials=M, batched={"runoff": np.stack(runoffs)},
                   per_trial_time_step=True, trial_layout=trial_layout)
    oracles = []
    for m in range(M):
        s = dict(src, runoff=runoffs[m])
        orc, _ = oracle_for(s, mods)
        oracles.append(orc)
    for step in range(5):
        model.step_advance()
        counts = model.t["max_sub_steps"].cpu().numpy()
        for m in range(M):
            assert int(orc_adv := oracles[m].advance(1800.0, 1)[0]) == counts[m], (step, m, orc_adv, counts[m])
    import torch
    torch.cuda.synchronize()
    for name in STATE_FIELDS:
        got = model.get(name)
        for m in range(M):
            msg = mismatch(got[m], oracles[m].state[name])
            assert msg is None, f"member {m} {name}: {msg}"
    del _lib


@pytest.mark.parametrize("precise", [False, True])
def test_batched_geometry_matches_standalone_runs(precise):
    """Members with their own river Manning's n, river height, catchment
    elevation (member m raised by 1,000 m per m, where FP32 surfaces are
    coarser), levee crown height and path sill elevations, and their own
    runoff, full physics, per-member CFL steps: each member is bit-exact with
    a standalone oracle run given that member's parameters, with the
    reference's FP32 surfaces and with precise surfaces."""
    net, src = small_source()
    M = 3
    rng = np.random.default_rng(9)
    n = src["river_height"].shape[0]
    batched = {
        "river_manning": np.stack([src["river_manning"] * rng.lognormal(0.0, 0.2, n) for _ in range(M)]),
        "river_height": np.stack([src["river_height"] * rng.uniform(0.8, 1.2, n) for _ in range(M)]),
        "catchment_elevation": np.stack([src["catchment_elevation"] + 1000.0 * m for m in range(M)]),
        "levee_crown_height": np.stack([src["levee_crown_height"] * rng.uniform(1.0, 1.2, src["levee_crown_height"].shape)
                                        for _ in range(M)]),
        "bifurcation_elevation": np.stack([src["bifurcation_elevation"] + 1000.0 * m for m in range(M)]),
        "runoff": np.stack(_member_sources(src, M)),
    }
    batched = {k: v.astype(np.float32) for k, v in batched.items()}
    model = _model(src, FULL, num_trials=M, batched=batched, per_trial_time_step=True, trial_layout="inner",
                   precise_surfaces=precise)
    oracles = [oracle_for(dict(src, **{k: v[m] for k, v in batched.items()}), FULL, precise_surfaces=precise)[0]
               for m in range(M)]
    for step in range(5):
        model.step_advance()
        counts = model.t["max_sub_steps"].cpu().numpy()
        for m in range(M):
            assert int(oracles[m].advance(1800.0, 1)[0]) == counts[m], (step, m)
    import torch
    torch.cuda.synchronize()
    for name in STATE_FIELDS + ("bifurcation_outflow", "bifurcation_cross_section_depth"):
        got = model.get(name)
        for m in range(M):
            msg = mismatch(got[m], oracles[m].state[name])
            assert msg is None, f"member {m} {name}: {msg}"


def test_batched_levee_heights_are_validated():
    """A trial whose levee crown falls to its base height is refused, as the
    levee module refuses it for a single run (levee.py:233-244)."""
    net, src = small_source()
    crown = np.stack([src["levee_crown_height"], src["levee_crown_height"]]).astype(np.float32)
    crown[1, 0] = src["levee_base_height"][0]
    with pytest.raises(ValueError, match="base height"):
        _model(src, FULL, num_trials=2, batched={"levee_crown_height": crown})


def test_batched_shared_time_step():
    """Trials sharing one time step (the reference's batched semantics): every
    member advances the maximum sub-step count; compare with fixed-count runs."""
    from oracle.oracle import HAS_ADAPTIVE
    net, src = small_source(paths=0, levee=0.0, reservoirs=0)
    M = 3
    runoffs = _member_sources(src, M)
    mods = ("base", "adaptive_time")
    model = _model(src, mods, num_trials=M, batched={"runoff": np.stack(runoffs)})
    oracles = []
    for m in range(M):
        orc, _ = oracle_for(dict(src, runoff=runoffs[m]), mods)
        orc.flags &= ~HAS_ADAPTIVE
        oracles.append(orc)
    for step in range(4):
        n = model.step_advance()
        for m in range(M):
            oracles[m].advance(1800.0, 1, fixed_count=n)
    got = model.get("river_outflow")
    for m in range(M):
        assert mismatch(got[m], oracles[m].state["river_outflow"]) is None


@pytest.mark.parametrize("what", ["levee_fraction", "flood_depth_table"])
def test_batched_levee_fraction_derives_per_trial_base(what):
    """Batching only the levee fraction (or only the floodplain table): each
    trial's levee base height is derived from its own fraction over its own
    table, as the levee module derives a single run's (levee.py:140-217), so
    every member is bit-exact with a standalone oracle run on its own
    fraction or table.  River storage is raised so that the levees are
    engaged."""
    net, src = small_source()
    rs = np.array(src["river_storage"], dtype=np.float64)
    lid = np.asarray(src["levee_catchment_idx"])
    rs[lid] *= np.random.default_rng(5).uniform(1.0, 60.0, lid.size)
    src["river_storage"] = rs
    M = 3
    rng = np.random.default_rng(12)
    # fractions and tables no larger than the source's keep every base height
    # below its crown (the base rises with both)
    if what == "levee_fraction":
        frac = np.asarray(src["levee_fraction"], np.float32)
        vals = np.stack([frac * rng.uniform(0.3, 1.0, lid.size) for _ in range(M)]).astype(np.float32)
    else:
        base = np.asarray(src["flood_depth_table"], np.float32)
        vals = np.stack([base * np.float32(0.6 + 0.2 * m) for m in range(M)]).astype(np.float32)
    model = _model(src, FULL, num_trials=M, batched={what: vals}, per_trial_time_step=True)
    oracles = [oracle_for(dict(src, **{what: vals[m]}), FULL)[0] for m in range(M)]
    # the members' base heights differ, and each is its standalone run's
    bases = model.get("levee_base_height").reshape(M, -1)
    for m in range(M):
        assert np.array_equal(bases[m], oracles[m].state["levee_base_height"])
    assert not np.array_equal(bases[0], bases[1])
    for step in range(4):
        model.step_advance()
        counts = model.t["max_sub_steps"].cpu().numpy()
        for m in range(M):
            assert int(oracles[m].advance(1800.0, 1)[0]) == counts[m], (step, m)
    import torch
    torch.cuda.synchronize()
    for name in STATE_FIELDS:
        got = model.get(name)
        for m in range(M):
            msg = mismatch(got[m], oracles[m].state[name])
            assert msg is None, f"member {m} {name}: {msg}"


@pytest.mark.parametrize("layout", ["keyed", "basin", "network"])
def test_levee_rows_in_device_order(layout):
    """Levee rows sit on the device in the order of their catchments' device
    positions; get() gives the parameters and indices back in the input's
    order, and a run matches the oracle (which keeps the input's rows)."""
    net, src = small_source()
    lid = np.asarray(src["levee_catchment_idx"])
    # rows in an order unrelated to the catchments' (reversed input)
    rev = np.arange(lid.size)[::-1].copy()
    for k in ("levee_catchment_idx", "levee_crown_height", "levee_fraction", "levee_base_height"):
        src[k] = np.ascontiguousarray(np.asarray(src[k])[rev])
    model = _model(src, FULL, layout=layout)
    dev = model.t["levee_catchment_idx"].cpu().numpy()
    assert np.all(np.diff(dev) > 0)
    for k in ("levee_catchment_idx", "levee_crown_height", "levee_fraction", "levee_base_height"):
        assert np.array_equal(model.get(k), np.asarray(src[k])), k
    orc, _ = oracle_for(src, FULL)
    for _ in range(3):
        assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    for name in STATE_FIELDS:
        msg = mismatch(model.get(name), orc.state[name])
        assert msg is None, f"{name}: {msg}"


def _hub_paths(src, seed=3):
    """src with extra bifurcation paths on three hub catchments (several
    paths each, in both roles) and every path row in a shuffled order."""
    from cmfgpu_amd import synthetic
    rng = np.random.default_rng(seed)
    n = np.asarray(src["downstream_idx"]).size
    hubs = rng.choice(n, 3, replace=False)
    others = rng.choice(np.setdiff1d(np.arange(n), hubs), 12, replace=False)
    pairs = ([(hubs[0], o) for o in others[:4]] + [(o, hubs[1]) for o in others[4:8]]
             + [(hubs[2], o) for o in others[8:10]] + [(o, hubs[2]) for o in others[10:12]]
             + [(hubs[0], hubs[1]), (hubs[1], hubs[2]), (hubs[2], hubs[0])])
    ci = np.array([c for c, _ in pairs], np.int32)
    di = np.array([d for _, d in pairs], np.int32)
    L = synthetic.NUM_BIFURCATION_LEVELS
    bed = src["catchment_elevation"][ci].astype(np.float64) - src["river_height"][ci].astype(np.float64)
    extra = {
        "bifurcation_catchment_idx": ci, "bifurcation_downstream_idx": di,
        "bifurcation_length": (0.5 * (src["river_length"][ci].astype(np.float64)
                                      + src["river_length"][di].astype(np.float64))).astype(np.float32),
        "bifurcation_width": np.full((ci.size, L), 120.0, np.float32),
        "bifurcation_elevation": (bed[:, None] + synthetic.BIFURCATION_LEVEL_HEIGHTS[None, :]).astype(np.float32),
        "bifurcation_manning": np.broadcast_to(synthetic.BIFURCATION_MANNING, (ci.size, L)).copy(),
    }
    out = dict(src)
    for k, v in extra.items():
        out[k] = np.concatenate([np.asarray(src[k]), v])
    P = out["bifurcation_catchment_idx"].size
    shuf = rng.permutation(P)
    for k in list(extra) :
        out[k] = np.ascontiguousarray(out[k][shuf])
    cid = np.asarray(src["catchment_id"])
    out["bifurcation_catchment_id"] = cid[out["bifurcation_catchment_idx"]]
    out["bifurcation_downstream_id"] = cid[out["bifurcation_downstream_idx"]]
    out["bifurcation_path_id"] = np.arange(P, dtype=np.int64)
    return out, hubs


@pytest.mark.parametrize("layout", ["keyed", "basin"])
def test_path_device_order_bitexact(layout):
    """Paths sit on the device by their upstream ends' device positions while
    each catchment keeps its paths' input order (its path terms are summed in
    that order): hub catchments with several paths in both roles, path rows
    shuffled; get() gives path fields and indices back in the input order,
    and the run is bit-exact with the oracle."""
    net, src = small_source()
    src, hubs = _hub_paths(src)
    model = _model(src, FULL, layout=layout)
    assert not np.array_equal(model.pperm, np.arange(model.P))
    for k in ("bifurcation_catchment_idx", "bifurcation_downstream_idx"):
        assert np.array_equal(model.get(k), src[k]), k
    for k in ("bifurcation_width", "bifurcation_elevation", "bifurcation_manning"):
        assert np.array_equal(model.get(k), src[k]), k
    orc, _ = oracle_for(src, FULL)
    for _ in range(4):
        assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    for name in STATE_FIELDS + ("bifurcation_outflow", "bifurcation_cross_section_depth"):
        msg = mismatch(model.get(name), orc.state[name])
        assert msg is None, f"{name}: {msg}"
    # the hubs' path terms are non-zero sums of several terms
    gb = np.asarray(orc.state["global_bifurcation_outflow"])[hubs]
    assert np.count_nonzero(gb) >= 2, gb
    assert np.array_equal(model.get("global_bifurcation_outflow"), orc.state["global_bifurcation_outflow"])


@pytest.mark.parametrize("mode", ["fused", "ops"])
def test_fp32_storage_profile_bitexact(mode):
    """The float32 profile without mixed precision (storages and accumulators
    in float32 too) against the oracle's (float, float) instantiation, full
    physics, both schedules."""
    net, src = small_source()
    model = _model(src, FULL, mixed_precision=False, mode=mode)
    assert model.t["river_storage"].dtype == __import__("torch").float32
    orc, host = oracle_for(src, FULL, mixed_precision=False)
    assert orc.precision == 2
    for step in range(4):
        assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    _compare(model, orc, STATE_FIELDS + ("bifurcation_outflow", "global_bifurcation_outflow", "total_storage"))


NO_RES = ("base", "adaptive_time", "bifurcation", "levee")


@pytest.mark.parametrize("case", ["plain", "paths_levees", "log", "precise", "fp64", "fp32sto", "network_levee",
                                  "reservoirs", "keyed"])
def test_schedules_bitexact(case):
    """The block-local schedule (one pass completing whole blocks, then the
    limiter and stage lists), the three-pass schedule (provisional limiters
    from the outflow pass, fixed up at path ends and below backflow) and the
    limiter-pass schedule against the oracle, bit for bit; levee catchments
    started at 1-60 times their river storage so that the levee cases and
    floodplains are in play.  Backflow must occur (non-mouth catchments with
    negative outflow), so that the fix-up of flagged limiters is exercised."""
    import torch
    kw, okw = {}, {}
    mods = NO_RES
    net, src = small_source(n=6000, seed=21, reservoirs=40 if case == "reservoirs" else 0, paths=150,
                            levee=0.08)
    if case == "plain":
        mods = ("base", "adaptive_time")
    if case == "log":
        mods = NO_RES + ("log",)
        kw = okw = dict(log_buffer_size=64)
    if case == "precise":
        kw = okw = dict(precise_surfaces=True)
    if case == "fp64":
        kw = okw = dict(precision="float64")
    if case == "fp32sto":
        kw = okw = dict(mixed_precision=False)
    if case == "reservoirs":
        mods = FULL
    if case == "keyed":
        # ranges in another order than their owners (start and end arrays)
        kw = dict(layout="keyed")
    rs = np.array(src["river_storage"], dtype=np.float64)
    lid = np.asarray(src["levee_catchment_idx"])
    rs[lid] *= np.random.default_rng(3).uniform(1.0, 60.0, lid.size)
    src["river_storage"] = rs
    # the basin layout unless the case names another (the block-local
    # schedule completes most catchments inside their block there)
    models = {name: _model(src, mods, **{"layout": "basin", **kw})
              for name in ("block_local", "three_pass", "limiter_pass")}
    for name, m in models.items():
        m.set_schedule(name)
    if case == "network_levee":
        # a separate levee pass after the stage (no levee_index)
        for m in models.values():
            m.t.pop("levee_index")
            m._table = m._build_table()
    B, nl, ns = models["block_local"].local_schedule_counts()
    assert B == 512 and 0 < nl < ns < models["block_local"].N
    orc, _ = oracle_for(src, mods, **okw)
    for m in models.values():
        m.keep_log_sums = True
    for step in range(5):
        counts = {name: m.step_advance() for name, m in models.items()}
        assert set(counts.values()) == {int(orc.advance(1800.0, 1)[0])}, counts
        if case == "log":
            # FP64 sums folded in another block order: relative 1e-12
            c = models["three_pass"].last_log_sums
            for name in ("block_local", "limiter_pass"):
                a = models[name].last_log_sums
                assert np.all(np.abs(a - c) <= 1e-12 * np.maximum(np.abs(c), 1e-9) + 1e-15), (name, step)
    torch.cuda.synchronize()
    fields = STATE_FIELDS + (("bifurcation_outflow", "bifurcation_cross_section_depth", "global_bifurcation_outflow",
                              "limit_rate") if case != "plain" else ())
    for name, m in models.items():
        for f in fields:
            if f not in orc.state:
                continue
            msg = mismatch(m.get(f), orc.state[f])
            assert msg is None, f"{name} {f}: {msg}"
    dn = np.asarray(src["downstream_idx"])
    inland = dn != np.arange(dn.size)
    back = (orc.state["river_outflow"] < 0) | (orc.state["flood_outflow"] < 0)
    assert np.count_nonzero(back & inland) > 0, "no backflow: the limiter fix-up was not exercised"


def test_fp32_storage_batched_members_match_standalone_runs():
    """Members in the float32 profile without mixed precision, trial-inner,
    per-member CFL counts, full physics: each bit-exact with a standalone
    (float, float) oracle run on its own runoff."""
    net, src = small_source()
    M = 3
    runoffs = _member_sources(src, M)
    model = _model(src, FULL, num_trials=M, batched={"runoff": np.stack(runoffs)}, per_trial_time_step=True,
                   trial_layout="inner", mixed_precision=False)
    oracles = [oracle_for(dict(src, runoff=runoffs[m]), FULL, mixed_precision=False)[0] for m in range(M)]
    for step in range(4):
        model.step_advance()
        counts = model.t["max_sub_steps"].cpu().numpy()
        for m in range(M):
            assert int(oracles[m].advance(1800.0, 1)[0]) == counts[m], (step, m)
    import torch
    torch.cuda.synchronize()
    for name in STATE_FIELDS:
        got = model.get(name)
        for m in range(M):
            msg = mismatch(got[m], oracles[m].state[name])
            assert msg is None, f"member {m} {name}: {msg}"


@pytest.mark.parametrize("members,per_trial", [(2, True), (16, True), (64, True), (16, False)])
def test_member_cfl_counts_in_the_inner_layout(members, per_trial):
    """The CFL pass of trial-inner runs whose member count divides 64 (one
    trial per lane class, block maxima folded per trial): with per-member dt
    every member's count and state equal its standalone oracle run's over
    three outer steps, full physics with reservoirs; with one shared dt the
    first count is the largest standalone count."""
    import torch
    net, src = small_source(n=1500, seed=5)
    runoffs = _member_sources(src, members, seed0=7000)
    model = _model(src, FULL, num_trials=members, batched={"runoff": np.stack(runoffs)},
                   per_trial_time_step=per_trial, trial_layout="inner")
    oracles = [oracle_for(dict(src, runoff=r), FULL)[0] for r in runoffs]
    steps = 3 if per_trial else 1
    for step in range(steps):
        model.step_advance()
        counts = model.t["max_sub_steps"].cpu().numpy()
        want = [int(o.advance(1800.0, 1)[0]) for o in oracles]
        if per_trial:
            assert counts.tolist() == want, (step, counts.tolist(), want)
        else:
            assert int(counts[0]) == max(want), (int(counts[0]), want)
    if per_trial:
        torch.cuda.synchronize()
        for name in ("river_storage", "flood_storage", "river_outflow"):
            got = model.get(name)
            for m in (0, members // 2, members - 1):
                msg = mismatch(got[m], oracles[m].state[name])
                assert msg is None, f"member {m} {name}: {msg}"
