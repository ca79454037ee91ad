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


@pytest.mark.parametrize("layout", ["level", "network"])
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


@pytest.mark.parametrize("layout", ["level", "network"])
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
