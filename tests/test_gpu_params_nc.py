"""A model built from a parameters.nc file runs the HIP path bit for bit with
the oracle run from the same file (the reference's user loop starts from
that file: scripts/run_daily_bin.py:65, schema cmfgpu/params/schema.py:1-140).
"""
import numpy as np
import pytest

from cmfgpu_amd import synthetic
from cmfgpu_amd.params_nc import read_parameters, write_parameters
from helpers import FULL, STATE_FIELDS, mismatch, oracle_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["basin", "network"])
def test_model_from_a_cdf5_file_matches_the_oracle(tmp_path, layout):
    import torch
    from cmfgpu_amd.models import CaMaFlood

    net = synthetic.generate_network(3000, 1007, levee=0.08, paths=60, reservoirs=12)
    p = tmp_path / "parameters.nc"
    write_parameters(p, dict(net.fields), version=5)
    src = read_parameters(p)
    # the ids are int64 in the file and the index fields are derived from them
    assert src["downstream_id"].dtype == np.int64
    assert np.array_equal(src["downstream_idx"], net.fields["downstream_idx"])
    forcing = synthetic.Runoff(src["catchment_area"], 2008, rho=0.8)
    src["runoff"] = forcing.next_day() * np.float32(3)
    model = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0, layout=layout)
    orc, _ = oracle_for(src, FULL)
    for day in range(2):
        if day:
            r = forcing.next_day() * np.float32(3)
            model.set_inputs(r)
            orc.state["runoff"][:] = r
        for _ in range(3):
            assert model.step_advance() == int(orc.advance(1800.0, 1)[0])
    torch.cuda.synchronize()
    for name in STATE_FIELDS + ("bifurcation_outflow", "reservoir_total_inflow"):
        msg = mismatch(model.get(name), orc.state[name])
        assert msg is None, f"{name}: {msg}"
    model.close()
