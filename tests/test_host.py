"""Host-side checks that need no GPU: the C ABI library, the op registry and
spec restatement, the module validators and derived parameters."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from cmfgpu_amd import _lib, synthetic
from cmfgpu_amd.modules import LeveeModule, ReservoirModule
from cmfgpu_amd.modules.levee import interp_lookup
from cmfgpu_amd.phys import specs
from cmfgpu_amd.phys.registry import BackendRegistry
from cmfgpu_amd.state import assemble

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "cmf_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cmf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"{name} declared in cmf_hip.h but not exported"


def test_field_table_abi_matches():
    assert _lib.lib().cmf_fields_size() == ctypes.sizeof(_lib.CmfFields)
    assert _lib.lib().cmf_version() >= 10000


def test_network_create_rejects_bad_topology_without_touching_the_device():
    h = ctypes.c_void_p()
    bad = np.array([0, 0, 2], np.int32)      # catchment 1 drains upstream
    rc = _lib.lib().cmf_network_create(3, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 0, None, None, 1,
                                       ctypes.byref(h))
    assert rc == _lib.CMF_ERR_INVALID
    assert b"Flow direction" in _lib.lib().cmf_last_error()


def test_ops_reject_null_tables():
    for sym in _lib.OP_SYMBOLS.values():
        rc = getattr(_lib.lib(), sym)(None, None, None)
        assert rc == _lib.CMF_ERR_INVALID


def test_specs_match_reference_construction_rules():
    assert len(specs.ALL_SPECS) == 11
    assert specs.OUTFLOW.buffers["outgoing_storage_ptr"] == "atomic_add"
    assert specs.ADAPTIVE_TIME.buffers["max_sub_steps_ptr"] == "atomic_max"
    assert "is_levee" not in specs.OUTFLOW.required_fields(set())
    assert "is_levee" in specs.OUTFLOW.required_fields({"HAS_LEVEE"})
    with pytest.raises(ValueError, match="conflicting access"):
        specs._spec("x", ("a_ptr",), "n", read=("a_ptr",), write=("a_ptr",))
    with pytest.raises(ValueError, match="requires one explicit access mode"):
        specs._spec("x", ("a_ptr", "b_ptr"), "n", read=("a_ptr",))
    with pytest.raises(ValueError, match="undefined runtime scalar"):
        specs._spec("x", ("a_ptr", "mystery"), "n", read=("a_ptr",))


def test_registry_has_hip_only_and_needs_a_bound_model():
    from cmfgpu_amd.phys import compute_outflow
    with pytest.raises(RuntimeError, match="no CaMaFlood model is bound"):
        compute_outflow()
    with pytest.raises(ValueError, match="unsupported backends"):
        BackendRegistry({"triton": lambda: None}, name="compute_outflow", spec=specs.OUTFLOW)


def test_base_validators_raise_like_the_reference():
    net = synthetic.generate_network(500, 1)
    src = dict(net.fields)
    bad = dict(src, river_width=np.where(np.arange(500) == 3, -1.0, src["river_width"]).astype(np.float32))
    with pytest.raises(ValueError, match="river_width=1"):
        assemble(bad)
    table = src["flood_depth_table"].copy()
    table[5, 4] = table[5, 3] - 0.5
    with pytest.raises(ValueError, match="monotonically non-decreasing"):
        assemble(dict(src, flood_depth_table=table))
    with pytest.raises(ValueError, match="finite non-negative"):
        assemble(dict(src, river_storage=-np.ones(500)))


def test_levee_base_height_interpolates_the_table():
    table = np.array([[1.0, 2.0, 4.0, 4.5]], np.float32)
    pos = np.array([0.0, 0.5, 1.0, 2.5, 4.0], np.float32)
    got = interp_lookup(np.repeat(table, 5, 0), np.arange(5), pos)
    assert got.tolist() == [0.0, 0.5, 1.0, 3.0, 4.5]


def test_levee_validators():
    net = synthetic.generate_network(300, 2)
    base = assemble(dict(net.fields)).fields
    with pytest.raises(ValueError, match=r"\[0, 1\)"):
        LeveeModule({"levee_catchment_idx": [1], "levee_fraction": [1.0], "levee_crown_height": [9.0]}, base)
    with pytest.raises(ValueError, match="base height greater than or equal"):
        LeveeModule({"levee_catchment_idx": [1], "levee_fraction": [0.5], "levee_crown_height": [0.0]}, base)


def test_reservoir_derived_volumes_closed_form():
    net = synthetic.generate_network(300, 3)
    base = assemble(dict(net.fields)).fields
    src = {"reservoir_catchment_idx": [10], "reservoir_capacity": [1e8], "conservation_volume": [4e7],
           "emergency_volume": [9.5e7], "normal_outflow": [100.0], "flood_control_outflow": [300.0],
           "reservoir_area": [1e6]}
    r = ReservoirModule(src, base)
    fv = (9.5e7 - 4e7) / 0.95
    assert r.fields["adjustment_volume"][0] == pytest.approx(4e7 + 0.1 * fv, rel=1e-6)
    qsto = (4e7 * 0.7 + 100.0 * 365 * 86400 / 4) / (180 * 86400)
    qn = min(100.0, qsto) * 1.5
    assert r.fields["effective_normal_outflow"][0] == pytest.approx(qn, rel=1e-6)
    assert r.fields["adjustment_outflow"][0] == pytest.approx((qn + 300.0) / 2, rel=1e-6)
    dn = base["downstream_idx"]
    ups = np.flatnonzero((dn == 10) & (np.arange(300) != 10))
    assert r.fields["is_dam_upstream"][ups].all() and r.fields["is_dam_related"][10]
    with pytest.raises(ValueError, match="Conservation volume exceeds"):
        ReservoirModule(dict(src, conservation_volume=[9.9e7]), base)


def test_assemble_flags_follow_opened_modules():
    net, src = synthetic.generate_network(400, 4, levee=0.05, paths=10), None
    src = dict(net.fields)
    h = assemble(src, ("base", "adaptive_time", "bifurcation", "levee"))
    want = (_lib.HAS_BIFURCATION | _lib.HAS_LEVEE | _lib.HAS_ADAPTIVE_TIME | _lib.HAS_TOTAL_STORAGE
            | _lib.HAS_WATER_SURFACE | _lib.HAS_PROTECTED_WATER_SURFACE)
    assert h.flags == want
    with pytest.raises(ValueError, match="unsupported modules"):
        assemble(src, ("base", "inflow"))


def test_grouped_network_rejects_links_across_groups():
    lib = _lib.lib()
    P32 = ctypes.POINTER(ctypes.c_int32)
    P64 = ctypes.POINTER(ctypes.c_int64)
    dn = np.array([1, 1, 3, 3], np.int32)            # basins {0, 1} and {2, 3}
    h = ctypes.c_void_p()
    bad = np.array([0, 3, 4], np.int64)               # splits basin {0, 1}
    rc = lib.cmf_network_create_grouped(4, dn.ctypes.data_as(P32), 0, None, None, 1, 2,
                                        bad.ctypes.data_as(P64), ctypes.byref(h))
    assert rc == _lib.CMF_ERR_INVALID and b"drains into another group" in lib.cmf_last_error()
    empty = np.array([0, 0, 4], np.int64)
    rc = lib.cmf_network_create_grouped(4, dn.ctypes.data_as(P32), 0, None, None, 1, 2,
                                        empty.ctypes.data_as(P64), ctypes.byref(h))
    assert rc == _lib.CMF_ERR_INVALID and b"empty" in lib.cmf_last_error()
    ci, di = np.array([0], np.int32), np.array([2], np.int32)
    good = np.array([0, 2, 4], np.int64)
    rc = lib.cmf_network_create_grouped(4, dn.ctypes.data_as(P32), 1, ci.ctypes.data_as(P32),
                                        di.ctypes.data_as(P32), 1, 2, good.ctypes.data_as(P64), ctypes.byref(h))
    assert rc == _lib.CMF_ERR_INVALID and b"joins two groups" in lib.cmf_last_error()
    rc = lib.cmf_network_create_grouped(4, dn.ctypes.data_as(P32), 0, None, None, 2, 2,
                                        good.ctypes.data_as(P64), ctypes.byref(h))
    assert rc == _lib.CMF_ERR_UNSUPPORTED


def test_group_basins_orders_groups_contiguously():
    from cmfgpu_amd.models.cama_flood_model import _group_basins
    from cmfgpu_amd import network
    dn = np.array([1, 1, 3, 3, 5, 5], np.int32)      # three basins
    order = network.basin_level_order(dn)
    labels = np.array([7, 7, 3, 3, 7, 7])
    perm, gptr, lab = _group_basins(order, labels, dn)
    assert lab.tolist() == [3, 7]
    assert gptr.tolist() == [0, 2, 6]
    assert perm.tolist() == [2, 3, 0, 1, 4, 5]
    with pytest.raises(ValueError, match="union of basins"):
        _group_basins(order, np.array([1, 2, 3, 3, 4, 4]), dn)
