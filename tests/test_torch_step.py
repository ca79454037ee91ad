"""The plain PyTorch-op transcription used as the FP64 autograd checker of the
adjoint (oracle/torch_step.py) against the FP64 CPU oracle: same states after
several sub-steps, to the rounding of different summation orders."""
import numpy as np
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
