"""The plain PyTorch-op transcription used as the FP64 autograd checker of the
adjoint (oracle/torch_step.py) against the FP64 CPU oracle: same states after
several sub-steps, to the rounding of different summation orders."""
import numpy as np
import pytest
import torch

from helpers import oracle_for

FIELDS = ("river_outflow", "flood_outflow", "river_cross_section_depth", "flood_cross_section_depth",
          "flood_cross_section_area", "river_storage", "flood_storage", "river_depth", "flood_depth")


def test_torch_transcription_matches_fp64_oracle():
    from cmfgpu_amd import synthetic
    from oracle.torch_step import substep
    net = synthetic.generate_network(2000, 1003)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    orc, host = oracle_for(src, ("base",), precision="float64")
    st = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in FIELDS}
    prm = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in (
        "river_manning", "river_height", "flood_manning", "river_width", "river_length", "catchment_elevation",
        "catchment_area", "downstream_distance", "runoff", "flood_depth_table")}
    prm["downstream_idx"] = torch.tensor(orc.state["downstream_idx"], dtype=torch.int64)
    count = 6
    wet_seen = 0
    for _ in range(12):
        orc.advance(1800.0, 1, count)
        for _ in range(count):
            st = substep(st, prm, 1800.0 / count)
        wet_seen = max(wet_seen, int((orc.state["flood_depth"] > 0).sum()))
    assert wet_seen > 20
    for k in FIELDS:
        ref = orc.state[k].astype(np.float64)
        got = st[k].numpy()
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30)
        assert err < 1e-9, (k, err)


PATH_FIELDS = ("bifurcation_outflow", "bifurcation_cross_section_depth")


def levee_flooded(net, seed=3):
    """The network's fields with runoff x4 and every levee catchment's
    initial river storage scaled by U(1, 100), so that floods reach the
    levees within a few outer steps."""
    from cmfgpu_amd import synthetic
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    rs = np.array(src["river_storage"], dtype=np.float64)
    lid = np.asarray(src["levee_catchment_idx"])
    rs[lid] *= np.random.default_rng(seed).uniform(1.0, 100.0, lid.size)
    src["river_storage"] = rs.astype(np.asarray(src["river_storage"]).dtype)
    return src


def test_torch_transcription_with_paths_matches_fp64_oracle():
    """The same with 150 bifurcation paths open: path flows and depths and
    every catchment field after 12 outer steps of 6 sub-steps."""
    from cmfgpu_amd import synthetic
    from oracle.torch_step import substep
    net = synthetic.generate_network(2000, 1003, paths=150)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    orc, host = oracle_for(src, ("base", "bifurcation"), precision="float64")
    st = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in FIELDS + PATH_FIELDS}
    prm = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in (
        "river_manning", "river_height", "flood_manning", "river_width", "river_length", "catchment_elevation",
        "catchment_area", "downstream_distance", "runoff", "flood_depth_table", "bifurcation_length",
        "bifurcation_elevation", "bifurcation_width", "bifurcation_manning")}
    for k in ("downstream_idx", "bifurcation_catchment_idx", "bifurcation_downstream_idx"):
        prm[k] = torch.tensor(orc.state[k], dtype=torch.int64)
    count = 6
    moving = 0
    for _ in range(12):
        orc.advance(1800.0, 1, count)
        for _ in range(count):
            st = substep(st, prm, 1800.0 / count)
        moving = max(moving, int((np.abs(orc.state["bifurcation_outflow"]).sum(axis=1) > 0).sum()))
    assert moving > 50
    for k in FIELDS + PATH_FIELDS:
        ref = orc.state[k].astype(np.float64)
        got = st[k].numpy()
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30)
        assert err < 1e-9, (k, err)


LEVEE_FIELDS = ("protected_storage", "protected_depth")


@pytest.mark.parametrize("paths", [0, 150])
def test_torch_transcription_with_levees_matches_fp64_oracle(paths):
    """The same with levees on 30 % of the catchments (and with 150 paths,
    whose upper levels then run on the protected surface), the levee
    catchments starting at 1-100 times their initial river storage: every
    levee case (2: river side between base and crown, 3: protected side
    filling, 4: overtopped) occurs, and all state fields agree after 12 outer
    steps."""
    from cmfgpu_amd import synthetic
    from oracle.torch_step import substep
    net = synthetic.generate_network(2000, 1003, levee=0.3, paths=paths)
    src = levee_flooded(net)
    mods = ("base", "levee") + (("bifurcation",) if paths else ())
    orc, host = oracle_for(src, mods, precision="float64")
    extra = PATH_FIELDS if paths else ()
    st = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in FIELDS + LEVEE_FIELDS + extra}
    names = ["river_manning", "river_height", "flood_manning", "river_width", "river_length",
             "catchment_elevation", "catchment_area", "downstream_distance", "runoff", "flood_depth_table",
             "levee_crown_height", "levee_fraction", "levee_base_height"]
    idx = ["downstream_idx", "levee_catchment_idx"]
    if paths:
        names += ["bifurcation_length", "bifurcation_elevation", "bifurcation_width", "bifurcation_manning"]
        idx += ["bifurcation_catchment_idx", "bifurcation_downstream_idx"]
    prm = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in names}
    for k in idx:
        prm[k] = torch.tensor(orc.state[k], dtype=torch.int64)
    lidx = orc.state["levee_catchment_idx"]
    count = 6
    cases = set()
    for _ in range(12):
        orc.advance(1800.0, 1, count)
        for _ in range(count):
            st = substep(st, prm, 1800.0 / count)
        s = orc.state
        frac = s["flood_fraction"][lidx]
        ps = s["protected_storage"][lidx]
        if np.any((ps == 0) & (np.abs(frac - s["levee_fraction"]) == 0) & (s["flood_depth"][lidx] > 0)):
            cases.add(2)
        if np.any((ps > 0) & (s["flood_depth"][lidx] == s["levee_crown_height"])):
            cases.add(3)
        if np.any((ps > 0) & (s["protected_depth"][lidx] == s["flood_depth"][lidx])):
            cases.add(4)
    assert cases == {2, 3, 4}, cases
    for k in FIELDS + LEVEE_FIELDS + extra:
        ref = orc.state[k].astype(np.float64)
        got = st[k].numpy()
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30)
        assert err < 1e-9, (k, err)


RES_FIELDS = ("reservoir_total_inflow",)


def reservoirs_filled(net, seed=5):
    """The network's fields with runoff x4 and every reservoir catchment's
    initial river storage set to U(0.1, 1.1) of its capacity, so that the
    release runs through all four storage regimes."""
    from cmfgpu_amd import synthetic
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    rs = np.array(src["river_storage"], dtype=np.float64)
    rid = np.asarray(src["reservoir_catchment_idx"])
    cap = np.asarray(src["reservoir_capacity"], dtype=np.float64)
    rs[rid] = cap * np.random.default_rng(seed).uniform(0.1, 1.1, rid.size)
    src["river_storage"] = rs.astype(np.asarray(src["river_storage"]).dtype)
    return src


def test_torch_transcription_with_reservoirs_matches_fp64_oracle():
    """The same with 60 reservoirs (the kinematic flow upstream of each dam,
    the release by storage regime, the reservoir inflow carried between
    sub-steps), levees and paths: every regime occurs, and all state fields
    agree after 12 outer steps."""
    from cmfgpu_amd import synthetic
    from oracle.torch_step import substep
    net = synthetic.generate_network(3000, 1003, levee=0.05, paths=100, reservoirs=60)
    src = reservoirs_filled(net)
    mods = ("base", "levee", "bifurcation", "reservoir")
    orc, host = oracle_for(src, mods, precision="float64")
    st = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in FIELDS + LEVEE_FIELDS + PATH_FIELDS
          + RES_FIELDS}
    names = ["river_manning", "river_height", "flood_manning", "river_width", "river_length",
             "catchment_elevation", "catchment_area", "downstream_distance", "runoff", "flood_depth_table",
             "levee_crown_height", "levee_fraction", "levee_base_height", "bifurcation_length",
             "bifurcation_elevation", "bifurcation_width", "bifurcation_manning", "conservation_volume",
             "emergency_volume", "adjustment_volume", "effective_normal_outflow", "adjustment_outflow",
             "flood_control_outflow"]
    prm = {k: torch.tensor(orc.state[k], dtype=torch.float64) for k in names}
    for k in ("downstream_idx", "levee_catchment_idx", "bifurcation_catchment_idx", "bifurcation_downstream_idx",
              "reservoir_catchment_idx"):
        prm[k] = torch.tensor(orc.state[k], dtype=torch.int64)
    for k in ("is_dam_upstream", "is_reservoir"):
        prm[k] = torch.tensor(orc.state[k], dtype=torch.bool)
    rid = orc.state["reservoir_catchment_idx"]
    cv, av, ev = (orc.state[k] for k in ("conservation_volume", "adjustment_volume", "emergency_volume"))
    regimes = set()
    count = 6
    for _ in range(12):
        s = orc.state
        total = s["river_storage"][rid] + s["flood_storage"][rid]
        regimes |= {int(r) for r in np.where(total <= cv, 1, np.where(total <= av, 2, np.where(total <= ev, 3, 4)))}
        orc.advance(1800.0, 1, count)
        for _ in range(count):
            st = substep(st, prm, 1800.0 / count)
    assert regimes == {1, 2, 3, 4}, regimes
    assert prm["is_dam_upstream"].any()
    for k in FIELDS + LEVEE_FIELDS + PATH_FIELDS + RES_FIELDS:
        ref = orc.state[k].astype(np.float64)
        got = st[k].numpy()
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30)
        assert err < 1e-9, (k, err)
