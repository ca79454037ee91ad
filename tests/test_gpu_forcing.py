"""Runoff mapped from a grid on the device (cmf_hip_map_runoff) and the
device-side time integral of total outflow (the "mean" output statistics)."""
import ctypes

import numpy as np
import pytest

from helpers import small_source

pytestmark = pytest.mark.gpu


def _mapping(n, cells, seed=5):
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, 5, n)
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum(counts)
    indices = rng.integers(0, cells, indptr[-1]).astype(np.int32)
    weights = rng.uniform(1e6, 4e8, indptr[-1]).astype(np.float32)
    return indptr, indices, weights


def test_mapped_runoff_is_the_csr_product():
    import scipy.sparse as sp
    from cmfgpu_amd.forcing import RunoffMapping
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source(paths=0, levee=0.0, reservoirs=0)
    n = net.num_catchments
    cells = 90 * 180
    indptr, indices, weights = _mapping(n, cells)
    scale = 1e-3 / 86400.0
    grid = np.random.default_rng(9).lognormal(0.0, 1.0, (90, 180)).astype(np.float32)
    model = CaMaFlood(src, opened_modules=("base", "adaptive_time"), device="cuda:0")
    rm = RunoffMapping(indptr, indices, weights, cells, scale).bind(model)
    rm.apply(grid)
    m = sp.csr_matrix((weights.astype(np.float64), indices, indptr), shape=(n, cells))
    want = (scale * m.dot(grid.reshape(-1).astype(np.float64))).astype(np.float32)
    got = model.get("runoff")
    assert np.array_equal(got, want)
    model.step_advance()
    model.close()


def test_mapping_from_reference_npz(tmp_path):
    from cmfgpu_amd.forcing import RunoffMapping
    n, cells = 500, 64 * 32
    indptr, indices, weights = _mapping(n, cells, seed=3)
    ids = np.arange(n, dtype=np.int64) * 13 + 7
    order = np.random.default_rng(1).permutation(n)          # the file lists catchments in another order
    counts = np.diff(indptr)[order]
    fptr = np.zeros(n + 1, np.int64)
    fptr[1:] = np.cumsum(counts)
    fidx = np.concatenate([indices[indptr[r]:indptr[r + 1]] for r in order])
    fw = np.concatenate([weights[indptr[r]:indptr[r + 1]] for r in order])
    path = tmp_path / "runoff_mapping.npz"
    np.savez(path, catchment_ids=ids[order], sparse_data=fw, sparse_indices=fidx, sparse_indptr=fptr,
             matrix_shape=np.array([n, cells]), coord_lon=np.linspace(-180, 180, 64),
             coord_lat=np.linspace(-90, 90, 32))
    rm = RunoffMapping.from_npz(path, ids)
    assert np.array_equal(rm.indptr, indptr)
    assert np.array_equal(rm.indices, indices)
    assert np.array_equal(rm.weights, weights)
    rm2 = RunoffMapping.from_npz(path, np.array([7, 999999, 20]))
    assert np.diff(rm2.indptr).tolist() == [np.diff(indptr)[0], 0, np.diff(indptr)[1]]


def test_outflow_mean_is_the_substep_time_integral():
    import torch
    from cmfgpu_amd import _lib
    from cmfgpu_amd.models import CaMaFlood
    from helpers import FULL
    net, src = small_source()
    model = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0)
    ref = CaMaFlood(src, opened_modules=FULL, device="cuda:0", step_duration=1800.0)
    model.enable_outflow_mean()
    lib = _lib.lib()
    acc = np.zeros(net.num_catchments)
    for _ in range(3):
        c = model.step_advance()
        # the same outer step on the twin, sub-step by sub-step, summed on the host
        dt = np.float32(1800.0 / c)
        ref.t["time_step"].fill_(float(dt))
        for k in range(c):
            _lib.check(lib.cmf_hip_substep(ref.network_handle, ref.field_table_ref(), k, 1 if k == c - 1 else 0,
                                           _lib.current_stream()), "cmf_hip_substep")
            torch.cuda.synchronize()
            q = ref.get("river_outflow") + ref.get("flood_outflow")
            acc += q.astype(np.float64) * np.float64(dt)
    raw = model.t["acc_outflow"].cpu().numpy()[model.inv]
    assert np.array_equal(raw, acc)                  # the same sums, bit for bit
    mean = model.outflow_mean()
    assert np.allclose(mean, acc / (3 * 1800.0), rtol=1e-15, atol=0.0)
    assert np.abs(mean).max() > 0
    assert np.all(model.outflow_mean() == 0.0)       # reset
    model.close()
    ref.close()


@pytest.mark.parametrize("trial_layout", ["inner", "major"])
def test_set_inputs_from_a_device_tensor(trial_layout):
    """Runoff given as a device tensor is permuted into the device layout on
    the GPU and lands exactly where the host path puts the same values, for
    one run and for [T, N] members in both trial layouts."""
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.models import CaMaFlood
    net, src = small_source(paths=0, levee=0.0, reservoirs=0)
    n = net.num_catchments
    T = 3
    r = np.stack([synthetic.Runoff(src["catchment_area"], 2100 + t).next_day() for t in range(T)])
    kw = dict(opened_modules=("base", "adaptive_time"), device="cuda:0")
    one = CaMaFlood(src, **kw)
    many = CaMaFlood(src, num_trials=T, batched={"runoff": r}, trial_layout=trial_layout, **kw)
    for model, x in ((one, r[1]), (many, r[::-1].copy())):
        model.set_inputs(x)
        host = model.t["runoff"].clone()
        model.t["runoff"].zero_()
        model.set_inputs(torch.as_tensor(x, device="cuda:0"))
        assert torch.equal(model.t["runoff"], host)
        # float64 input: converted on the device
        model.set_inputs(torch.as_tensor(x, dtype=torch.float64, device="cuda:0"))
        assert torch.equal(model.t["runoff"], host)
    assert np.array_equal(many.get("runoff"), r[::-1])
    with pytest.raises(ValueError, match="shape"):
        one.set_inputs(torch.zeros(n + 1, device="cuda:0"))


@pytest.mark.parametrize("placed", [False, True])
def test_map_runoff_entry_point_with_and_without_out_index(placed):
    """cmf_hip_map_runoff through the C ABI: rows summed in index order in
    FP64 and written to out[r] (out_index NULL) or to out[out_index[r]] (a
    permutation), against the FP64 CSR product of scipy."""
    import scipy.sparse as sp
    import torch
    from cmfgpu_amd import _lib
    n, cells = 4000, 64 * 64
    indptr, indices, weights = _mapping(n, cells, seed=11)
    grid = np.random.default_rng(4).lognormal(0.0, 1.0, cells).astype(np.float32)
    scale = 2.5e-4
    want = (scale * sp.csr_matrix((weights.astype(np.float64), indices, indptr), shape=(n, cells))
            .dot(grid.astype(np.float64))).astype(np.float32)
    dev = "cuda:0"
    t = {k: torch.from_numpy(v).to(dev) for k, v in
         (("indptr", indptr), ("indices", indices), ("weights", weights), ("grid", grid))}
    out = torch.full((n,), -1.0, dtype=torch.float32, device=dev)
    place = np.random.default_rng(2).permutation(n).astype(np.int32) if placed else None
    d_place = torch.from_numpy(place).to(dev) if placed else None
    _lib.check(_lib.lib().cmf_hip_map_runoff(n, _lib.ptr(t["indptr"]), _lib.ptr(t["indices"]),
                                             _lib.ptr(t["weights"]), _lib.ptr(t["grid"]), scale,
                                             _lib.ptr(d_place) if placed else None, _lib.PRECISION_F32_F64,
                                             _lib.ptr(out), _lib.current_stream()), "cmf_hip_map_runoff")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    if placed:
        got = got[place]          # row r was written at place[r]
    assert np.array_equal(got, want)
