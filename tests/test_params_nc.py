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


def test_map_metadata_fields_keep_their_schema_dimensions(src, tmp_path):
    """The map-construction metadata of the schema (schema.py:44-73) round-trip
    on their own dimensions: scalars nx, ny, num_basins; basin_sizes on
    'basin' (here as long as the catchment list, which must not make it a
    catchment field); gauge fields on 'gauge'; output_catchment_id on
    'saved_points'."""
    n = src["river_width"].shape[0]
    rng = np.random.default_rng(2)
    meta = dict(src)
    meta.update({
        "nx": np.int64(1440), "ny": np.int64(720), "num_basins": np.int64(n),
        "basin_sizes": np.ones(n, np.int64),
        "gauge_catchment_id": src["catchment_id"][:7].copy(), "gauge_station_id": np.arange(7, dtype=np.int64),
        "gauge_reported_area_km2": rng.uniform(1e3, 1e5, 7).astype(np.float32),
        "output_catchment_id": src["catchment_id"][::50].copy(),
    })
    for version in (2, 5):
        p = tmp_path / f"meta_v{version}.nc"
        write_parameters(p, meta, version=version)
        arrays, dims, _ = read_nc(p)
        assert dims["basin"] == n and dims["gauge"] == 7 and dims["saved_points"] == meta["output_catchment_id"].size
        assert arrays["nx"].shape == () and int(arrays["nx"]) == 1440 and int(arrays["num_basins"]) == n
        got = read_parameters(p)
        for k in ("basin_sizes", "gauge_catchment_id", "gauge_reported_area_km2", "output_catchment_id",
                  "river_width"):
            assert _same(got[k], meta[k]), (version, k)
        # written back after reading (index fields derived by the reader are left out)
        write_parameters(tmp_path / f"again_v{version}.nc", got, version=version)
    with pytest.raises(ParameterFileError, match="dimensions"):
        write_parameters(tmp_path / "bad.nc", dict(meta, flood_depth_table=np.zeros(n, np.float32)))


def test_bool_fields_in_cdf2_are_nc_byte(src, tmp_path):
    """CDF-1 and CDF-2 have type codes 1-6 only: a bool field is written as
    NC_BYTE, which scipy's independent reader takes."""
    from scipy.io import netcdf_file
    n = src["river_width"].shape[0]
    flag = np.zeros(n, np.bool_)
    flag[::3] = True
    p = tmp_path / "bool_v2.nc"
    write_parameters(p, dict(src, is_outlet=flag), version=2)
    with netcdf_file(p, "r", mmap=False) as f:
        v = f.variables["is_outlet"]
        assert v.typecode() == "b"
        assert np.array_equal(v[:].astype(bool), flag)
    assert np.array_equal(read_parameters(p)["is_outlet"].astype(bool), flag)


def _cdf2_with_type(code: int, n: int = 4) -> bytes:
    import struct
    name = b"x"
    h = b"CDF\x02" + struct.pack(">i", 0)
    h += struct.pack(">ii", 10, 1) + struct.pack(">i", 1) + b"c\0\0\0" + struct.pack(">i", n)
    h += struct.pack(">ii", 0, 0)                            # no global attributes
    h += struct.pack(">ii", 11, 1) + struct.pack(">i", 1) + name + b"\0\0\0"
    h += struct.pack(">ii", 1, 0) + struct.pack(">ii", 0, 0)   # one dimension (id 0), no attributes
    h += struct.pack(">ii", code, n) + struct.pack(">q", len(h) + 4 + 4 + 8)
    return h + bytes(n)


def test_reader_refuses_cdf5_types_in_cdf1_and_cdf2(tmp_path):
    p = tmp_path / "ubyte_v2.nc"
    p.write_bytes(_cdf2_with_type(7))
    with pytest.raises(ParameterFileError, match="only in CDF-5"):
        read_nc(p)
    p.write_bytes(_cdf2_with_type(1))
    arrays, dims, _ = read_nc(p)
    assert arrays["x"].dtype == np.int8 and dims == {"c": 4}
