"""Uneven basins advanced together as groups, each with its own CFL sub-step
count and time step, in one launch per pass.

Every basin must be bit-identical with a standalone oracle run of that basin
alone (which takes its own CFL step), and the per-group FP64 water-balance
sums must match the oracle's to the tolerance of a different summation order.
"""
import numpy as np
import pytest

from helpers import STATE_FIELDS, mismatch, oracle_for

pytestmark = pytest.mark.gpu

PLAIN = ("base", "adaptive_time")


def _basin_source(bs, b, runoff):
    a, e = int(bs.basin_ptr[b]), int(bs.basin_ptr[b + 1])
    n = bs.network.num_catchments
    src = {}
    for name, arr in bs.network.fields.items():
        if arr.shape[:1] == (n,):
            x = arr[a:e].copy()
            if name == "downstream_idx":
                x = (x.astype(np.int64) - a).astype(np.int32)
            src[name] = x
    src["runoff"] = runoff[a:e].copy()
    return src


def _basins():
    from cmfgpu_amd import synthetic
    bs = synthetic.build_basins(12, 50, 3000)
    runoff = synthetic.Runoff(bs.network.fields["catchment_area"], 2002).next_day()
    src = dict(bs.network.fields, runoff=runoff)
    return bs, src, runoff


def test_groups_match_standalone_basins():
    import torch
    from cmfgpu_amd.models import CaMaFlood
    bs, src, runoff = _basins()
    model = CaMaFlood(src, opened_modules=PLAIN, device="cuda:0", step_duration=1800.0, groups=bs.group)
    assert model.G == 12
    oracles = [oracle_for(_basin_source(bs, b, runoff), PLAIN)[0] for b in range(12)]
    seen = set()
    for step in range(4):
        model.step_advance()
        counts = model.group_sub_steps()
        for b in range(12):
            n_cpu = int(oracles[b].advance(1800.0, 1)[0])
            assert n_cpu == counts[b], (step, b, n_cpu, counts[b])
        seen.update(int(c) for c in counts)
    assert len(seen) > 1, "all basins took the same sub-step count; the test would not see per-group dt"
    torch.cuda.synchronize()
    for name in STATE_FIELDS:
        got = model.get(name)
        for b in range(12):
            a, e = bs.basin_ptr[b], bs.basin_ptr[b + 1]
            if name not in oracles[b].state:
                continue
            msg = mismatch(got[a:e], oracles[b].state[name])
            assert msg is None, f"basin {b} {name}: {msg}"
    model.close()


def test_group_water_balance_sums():
    from cmfgpu_amd.models import CaMaFlood
    bs, src, runoff = _basins()
    mods = PLAIN + ("log",)
    model = CaMaFlood(src, opened_modules=mods, device="cuda:0", step_duration=1800.0, groups=bs.group,
                      log_buffer_size=64)
    model.step_advance()
    counts = model.group_sub_steps()
    g = model.last_group_log
    for b in range(12):
        orc, _ = oracle_for(_basin_source(bs, b, runoff), mods, log_buffer_size=64)
        n_cpu = int(orc.advance(1800.0, 1)[0])
        assert n_cpu == counts[b]
        c = orc.state["log_sums"][:, :n_cpu]
        got = g[:, b, :n_cpu]
        scale = np.maximum(np.abs(c), 1e-9)
        assert np.all(np.abs(got - c) <= 1e-12 * scale + 1e-15), (b, got, c)
        assert np.all(g[:, b, n_cpu:] == 0.0)
    model.close()
