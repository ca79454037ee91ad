"""Sub-step loops replayed from captured HIP graphs against the eager launches.

Each graph (one per sub-step count) is replayed many times with the model's
own eager work in between (CFL pass, count read-back, runoff updates,
synchronisations); every outer step must leave every field bit-identical to
the eager run.
"""
import numpy as np
import pytest

from helpers import FULL, STATE_FIELDS, mismatch, small_source

pytestmark = pytest.mark.gpu


def _pair(src, mods, **kw):
    from cmfgpu_amd.models import CaMaFlood
    return [CaMaFlood(src, opened_modules=mods, device="cuda:0", step_duration=1800.0, use_graph=g, **kw)
            for g in (False, True)]


@pytest.mark.parametrize("modules", ["plain", "full"])
def test_graph_replays_match_eager(modules):
    import torch
    from cmfgpu_amd import synthetic
    mods = ("base", "adaptive_time") if modules == "plain" else FULL
    net, src = small_source() if modules == "full" else small_source(paths=0, levee=0.0, reservoirs=0)
    eager, graph = _pair(src, mods)
    forcing = synthetic.Runoff(src["catchment_area"], 2100)
    counts = []
    for step in range(12):
        if step % 4 == 0:
            r = forcing.next_day() * np.float32(3.0)
            eager.set_inputs(r)
            graph.set_inputs(r)
        a, b = eager.step_advance(), graph.step_advance()
        assert a == b, (step, a, b)
        counts.append(a)
        torch.cuda.synchronize()
        for name in STATE_FIELDS:
            if name in eager.t:
                msg = mismatch(graph.get(name), eager.get(name))
                assert msg is None, f"step {step} {name}: {msg}"
    assert len(counts) == 12
    eager.close()
    graph.close()


def test_graph_replays_match_eager_with_groups_and_log():
    import torch
    from cmfgpu_amd import synthetic
    bs = synthetic.build_basins(6, 50, 2000)
    src = dict(bs.network.fields, runoff=synthetic.Runoff(bs.network.fields["catchment_area"], 2101).next_day())
    mods = ("base", "adaptive_time", "log")
    eager, graph = _pair(src, mods, groups=bs.group, log_buffer_size=32)
    for step in range(8):
        eager.step_advance()
        graph.step_advance()
        torch.cuda.synchronize()
        assert np.array_equal(eager.group_sub_steps(), graph.group_sub_steps())
        for name in STATE_FIELDS:
            if name in eager.t:
                msg = mismatch(graph.get(name), eager.get(name))
                assert msg is None, f"step {step} {name}: {msg}"
        assert np.array_equal(eager.last_group_log, graph.last_group_log), step
    eager.close()
    graph.close()


@pytest.mark.parametrize("trial_layout", ["inner", "major"])
def test_graph_replays_match_eager_with_members(trial_layout):
    """Eight runoff members with their own sub-step counts, full physics:
    each outer step one graph launch (the loop for the largest count, the
    members past their count sitting out), bit-identical to the eager run."""
    import torch
    from cmfgpu_amd import synthetic
    net, src = small_source()
    M = 8
    runoff = np.stack([synthetic.Runoff(src["catchment_area"], 4000 + m).next_day() * np.float32(1 + m)
                       for m in range(M)])
    eager, graph = _pair(src, FULL, num_trials=M, batched={"runoff": runoff}, per_trial_time_step=True,
                         trial_layout=trial_layout)
    for step in range(8):
        eager.step_advance()
        graph.step_advance()
        torch.cuda.synchronize()
        counts = eager.t["max_sub_steps"].cpu().numpy()
        assert np.array_equal(counts, graph.t["max_sub_steps"].cpu().numpy()), step
        for name in STATE_FIELDS:
            if name in eager.t:
                msg = mismatch(graph.get(name), eager.get(name))
                assert msg is None, f"step {step} {name}: {msg}"
    assert len(set(counts.tolist())) > 1    # the members really took different counts
    eager.close()
    graph.close()
