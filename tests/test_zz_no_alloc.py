"""cmf_hip_flood_stage_log_noalloc, the allocation-free compute_flood_stage_log
(partials in the library's static device array), against the network entry
point cmf_hip_flood_stage_log: identical fields and water-balance sums, on two
calls in a row, the second on a smaller network."""
import ctypes

import numpy as np
import pytest

from helpers import FULL, small_source

pytestmark = pytest.mark.gpu

OUT = ("river_storage", "flood_storage", "protected_storage", "river_depth", "flood_depth", "protected_depth",
       "flood_fraction", "outgoing_storage")


def _pair(n, seed):
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source(n=n, seed=seed)
    mods = FULL + ("log",)
    models = [CaMaFlood(src, opened_modules=mods, device="cuda:0", log_buffer_size=64, step_duration=1800.0)
              for _ in range(2)]
    for m in models:
        for _ in range(2):
            m.step_advance()
    return models


@pytest.mark.parametrize("sizes", [(3000, 1200)])
def test_noalloc_stage_log_matches_the_network_entry_point(sizes):
    import torch
    from cmfgpu_amd import _lib
    lib = _lib.lib()
    for n, seed in zip(sizes, (7, 8)):
        a, b = _pair(n, seed)
        # the previous sub-step's inflows and outflows are in place; rerun its stage
        _lib.check(lib.cmf_hip_flood_stage_log(a.network_handle, a.field_table_ref(), _lib.current_stream()),
                   "cmf_hip_flood_stage_log")
        _lib.check(lib.cmf_hip_flood_stage_log_noalloc(b.field_table_ref(), _lib.current_stream()),
                   "cmf_hip_flood_stage_log_noalloc")
        torch.cuda.synchronize()
        for name in OUT:
            assert torch.equal(a.t[name], b.t[name]), (n, name)
        la, lb = a.log_sums(), b.log_sums()
        assert np.array_equal(la, lb), n
        assert np.abs(la).sum() > 0
        a.close()
        b.close()


def test_noalloc_refuses_too_many_elements():
    from cmfgpu_amd import _lib
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source(n=500, seed=9)
    m = CaMaFlood(src, opened_modules=FULL + ("log",), device="cuda:0", log_buffer_size=8)
    f = _lib.CmfFields()
    ctypes.memmove(ctypes.byref(f), m.field_table_ref(), ctypes.sizeof(f))
    f.num_catchments = (1 << 24) + 1           # checked before any launch
    rc = _lib.lib().cmf_hip_flood_stage_log_noalloc(ctypes.byref(f), _lib.current_stream())
    assert rc == _lib.CMF_ERR_UNSUPPORTED
    assert b"static scratch" in _lib.lib().cmf_last_error()
    m.close()
