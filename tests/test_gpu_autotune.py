"""Kernel block sizes change launch shapes only: results must stay bit-identical
to the oracle at every size, and the autotuner must leave the state untouched."""
import numpy as np
import pytest

from helpers import FULL, STATE_FIELDS, mismatch, oracle_for, small_source

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size", [64, 128, 512])
def test_block_sizes_keep_results_bitexact(size):
    import torch
    from cmfgpu_amd import autotune
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source()
    model = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0)
    autotune.set_block_sizes(model, [size] * autotune.NUM_TUNABLE)
    assert autotune.get_block_sizes(model) == [size] * autotune.NUM_TUNABLE
    orc, _ = oracle_for(src, FULL)
    for _ in range(3):
        n = model.step_advance()
        assert n == int(orc.advance(1800.0, 1)[0])
    torch.cuda.synchronize()
    for name in STATE_FIELDS:
        if name in orc.state:
            msg = mismatch(model.get(name), orc.state[name])
            assert msg is None, f"block {size} {name}: {msg}"
    model.close()


def test_block_sizes_rejected_outside_range():
    from cmfgpu_amd import _lib, autotune
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source(paths=0, levee=0.0, reservoirs=0)
    model = CaMaFlood(src, opened_modules=("base", "adaptive_time"), device="cuda:0")
    for bad in (0, 32, 96, 1024):
        with pytest.raises(_lib.CmfError, match="multiple of 64"):
            autotune.set_block_sizes(model, [bad] * autotune.NUM_TUNABLE)
    assert autotune.get_block_sizes(model) == [256] * autotune.NUM_TUNABLE
    model.close()


def test_autotune_leaves_state_and_applies_choice(tmp_path):
    import torch
    from cmfgpu_amd import autotune
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source()
    model = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0)
    model.step_advance()
    before = {name: model.get(name).copy() for name in STATE_FIELDS if name in model.t}
    report = model.autotune(cache_path=tmp_path / "tune.json", reps=2)
    torch.cuda.synchronize()
    for name, arr in before.items():
        assert mismatch(model.get(name), arr) is None, name
    sizes = autotune.get_block_sizes(model)
    for name, rec in report.items():
        assert sizes[autotune.KERNELS[name]] == rec["block"]
    # a second model of the same shape reads the cached choice
    other = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0)
    assert other.autotune(cache_path=tmp_path / "tune.json") == report
    assert autotune.get_block_sizes(other) == sizes
    model.close()
    other.close()
