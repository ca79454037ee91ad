"""HIP path vs CPU oracle, bit for bit, on degenerate networks: one catchment
(with a levee, and the bifurcation module open with no paths), 500 catchments
that are all mouths (with five times as many paths as catchments, so pairs
repeat), a single 3,000-catchment river (3,000 topological levels), a 4,000-way
confluence (one mouth with 3,999 direct upstream catchments), and levees on
every catchment.  Runoff x4 and the levee catchments started at 1-100 times
their initial river storage so that floods reach the levees.  Both modes (fused, per-op),
four outer steps with the adaptive CFL count.
"""
import numpy as np
import pytest

from helpers import STATE_FIELDS, mismatch, oracle_for

pytestmark = pytest.mark.gpu

PATH_FIELDS = ("bifurcation_outflow", "bifurcation_cross_section_depth", "global_bifurcation_outflow")
EXTRA = ("protected_storage", "protected_depth", "flood_fraction", "limit_rate", "reservoir_total_inflow")


def _empty_paths(src):
    src["bifurcation_catchment_idx"] = np.zeros(0, np.int32)
    src["bifurcation_downstream_idx"] = np.zeros(0, np.int32)
    src["bifurcation_catchment_id"] = np.zeros(0, np.int64)
    src["bifurcation_downstream_id"] = np.zeros(0, np.int64)
    src["bifurcation_path_id"] = np.zeros(0, np.int64)
    src["bifurcation_length"] = np.zeros(0, np.float32)
    for k in ("bifurcation_width", "bifurcation_elevation", "bifurcation_manning"):
        src[k] = np.zeros((0, 5), np.float32)


def _case(name):
    from cmfgpu_amd import synthetic
    mods = ["base", "adaptive_time"]
    if name == "single":
        net = synthetic.generate_network(1, 1005, downstream=np.array([-1]), levee=1.0)
        src = dict(net.fields)
        _empty_paths(src)
        mods += ["bifurcation", "levee"]
    elif name == "all_mouths":
        net = synthetic.generate_network(500, 1005, downstream=-np.ones(500, np.int64), levee=0.2, paths=2500)
        src = dict(net.fields)
        mods += ["bifurcation", "levee"]
    elif name == "chain":
        net = synthetic.generate_network(3000, 1005, downstream=np.arange(-1, 2999), levee=0.2, paths=600,
                                         reservoirs=10)
        src = dict(net.fields)
        mods += ["bifurcation", "levee", "reservoir"]
    elif name == "confluence":
        net = synthetic.generate_network(4000, 1005, downstream=np.r_[-1, np.zeros(3999, np.int64)], levee=0.2,
                                         paths=800, reservoirs=10)
        src = dict(net.fields)
        mods += ["bifurcation", "levee", "reservoir"]
    else:   # levees everywhere
        net = synthetic.generate_network(2000, 1005, levee=1.0, paths=100)
        src = dict(net.fields)
        mods += ["bifurcation", "levee"]
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2005).next_day() * np.float32(4.0)
    if "levee_catchment_idx" in src:
        rs = np.array(src["river_storage"], dtype=np.float64)
        lid = np.asarray(src["levee_catchment_idx"])
        rs[lid] *= np.random.default_rng(3).uniform(1.0, 100.0, lid.size)
        src["river_storage"] = rs
    return net, src, tuple(mods)


@pytest.mark.parametrize("name", ["single", "all_mouths", "chain", "confluence", "levees_everywhere"])
@pytest.mark.parametrize("mode", ["fused", "ops"])
def test_degenerate_networks_bitexact(name, mode):
    import torch
    from cmfgpu_amd.models import CaMaFlood
    net, src, mods = _case(name)
    print(f"case {name} {mode}: {net.num_catchments} catchments, {net.num_paths} paths, "
          f"{net.num_levees} levees, {net.num_reservoirs} reservoirs", flush=True)
    model = CaMaFlood(src, opened_modules=mods, device="cuda:0", step_duration=1800.0, mode=mode)
    orc, _ = oracle_for(src, mods)
    counts = []
    for _ in range(4):
        n_gpu = model.step_advance()
        n_cpu = int(orc.advance(1800.0, 1)[0])
        assert n_gpu == n_cpu
        counts.append(n_gpu)
    torch.cuda.synchronize()
    for field in STATE_FIELDS + PATH_FIELDS + EXTRA:
        if field not in orc.state or field not in model.t:
            continue
        msg = mismatch(model.get(field), orc.state[field])
        assert msg is None, f"{name}/{mode} {field}: {msg}"
    if "levee_catchment_idx" in src:
        # the levees were reached: some protected storage or levee-side depth
        lid = np.asarray(src["levee_catchment_idx"])
        assert np.any(orc.state["flood_depth"][lid] > 0)
    model.close()
