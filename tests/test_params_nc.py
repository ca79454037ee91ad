"""Parameter files in the reference's parameters.nc schema, NetCDF classic
encodings (CPU only).

The parser is checked against an independent writer (scipy.io.netcdf_file,
CDF-1 and CDF-2), the writer against an independent reader (scipy, CDF-2),
the CDF-5 path (int64 ids) by round trip, and the whole chain by running the
oracle from the file's fields: bit-identical to the run from the fields that
were written.
"""
import numpy as np
import pytest

from cmfgpu_amd import synthetic
from cmfgpu_amd.params_nc import ParameterFileError, read_nc, read_parameters, write_parameters
from helpers import FULL, oracle_for


@pytest.fixture(scope="module")
def src():
    net = synthetic.generate_network(600, 1006, levee=0.1, paths=20, reservoirs=6)
    return dict(net.fields)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.astype(np.float64), b.astype(np.float64))


def test_cdf5_round_trip_keeps_values_ids_and_indices(src, tmp_path):
    p = tmp_path / "parameters.nc"
    write_parameters(p, src, version=5)
    got = read_parameters(p)
    for k, v in src.items():
        assert k in got, k
        assert _same(got[k], v), k
        if not k.endswith("_idx"):
            assert np.asarray(got[k]).dtype == np.asarray(v).dtype, k
    # the indices are derived from the ids, not stored
    assert "downstream_idx" not in read_nc(p)[0]


@pytest.mark.parametrize("version", [1, 2])
def test_reader_against_scipy_writer(src, tmp_path, version):
    from scipy.io import netcdf_file
    p = tmp_path / f"scipy_v{version}.nc"
    n, L = src["river_width"].shape[0], src["flood_depth_table"].shape[1]
    with netcdf_file(p, "w", version=version) as f:
        f.title = b"written by scipy"
        f.createDimension("catchment", n)
        f.createDimension("flood_level", L)
        for k in ("catchment_id", "downstream_id"):
            v = f.createVariable(k, "i4", ("catchment",))
            v[:] = src[k].astype(np.int32)
        for k in ("river_width", "river_height", "catchment_elevation"):
            v = f.createVariable(k, "f4", ("catchment",))
            v[:] = src[k]
        v = f.createVariable("river_storage", "f8", ("catchment",))
        v[:] = src["river_storage"]
        v = f.createVariable("flood_depth_table", "f4", ("catchment", "flood_level"))
        v[:] = src["flood_depth_table"]
    got = read_parameters(p)
    for k in ("catchment_id", "downstream_id", "river_width", "river_height", "catchment_elevation",
              "river_storage", "flood_depth_table", "downstream_idx"):
        assert _same(got[k], src[k]), k
    arrays, dims, gattrs = read_nc(p)
    assert dims == {"catchment": n, "flood_level": L}
    assert gattrs["title"] == "written by scipy"


def test_writer_against_scipy_reader(src, tmp_path):
    from scipy.io import netcdf_file
    p = tmp_path / "ours_v2.nc"
    write_parameters(p, src, version=2)
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.dimensions["catchment"] == src["river_width"].shape[0]
        assert f.dimensions["bifurcation_level"] == src["bifurcation_width"].shape[1]
        for k, v in src.items():
            if k.endswith("_idx"):
                assert k not in f.variables
                continue
            assert _same(f.variables[k][:], v), k


def test_oracle_from_the_file_matches_the_written_fields(src, tmp_path):
    p = tmp_path / "parameters.nc"
    write_parameters(p, src)
    src2 = read_parameters(p)
    src = dict(src, runoff=synthetic.Runoff(src["catchment_area"], 2006).next_day() * np.float32(4))
    src2["runoff"] = src["runoff"].copy()
    a, _ = oracle_for(src, FULL)
    b, _ = oracle_for(src2, FULL)
    for _ in range(3):
        assert int(a.advance(1800.0, 1)[0]) == int(b.advance(1800.0, 1)[0])
    for k in ("river_outflow", "flood_outflow", "river_storage", "flood_storage", "bifurcation_outflow",
              "protected_storage", "reservoir_total_inflow"):
        assert np.array_equal(a.state[k], b.state[k]), k


def test_hdf5_and_broken_files_are_refused_with_a_reason(src, tmp_path):
    h = tmp_path / "netcdf4.nc"
    h.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(ParameterFileError, match="nccopy -k cdf5"):
        read_parameters(h)
    bad = dict(src)
    bad["downstream_id"] = bad["downstream_id"].copy()
    bad["downstream_id"][5] = 10 ** 12
    p = tmp_path / "bad.nc"
    write_parameters(p, bad)
    with pytest.raises(ParameterFileError, match="downstream_id"):
        read_parameters(p)
    with pytest.raises(ParameterFileError, match="CDF-5"):
        write_parameters(tmp_path / "v2.nc", bad, version=2)


def test_a_file_without_state_starts_cold(src, tmp_path):
    """The reference's parameters.nc holds parameters only: the model starts
    from zero storage outside the reservoirs (whose cold start fills them)
    and runs (oracle, full physics)."""
    params = {k: v for k, v in src.items() if k not in ("river_storage", "river_depth")}
    p = tmp_path / "parameters.nc"
    write_parameters(p, params)
    got = read_parameters(p)
    assert "river_storage" not in got
    got["runoff"] = synthetic.Runoff(got["catchment_area"], 2007).next_day()
    orc, host = oracle_for(got, FULL)
    rs = np.array(host.fields["river_storage"])
    rs[got["reservoir_catchment_idx"]] = 0.0
    assert not np.any(rs)
    orc.advance(1800.0, 2)
    assert np.all(np.isfinite(orc.state["river_outflow"])) and np.any(orc.state["river_storage"] > 0)
