"""Network indexing (integer work): the C-ABI routines vs the pure-Python
restatement in oracle/ordering.py, bit for bit, plus textbook invariants."""
import numpy as np
import pytest

from cmfgpu_amd import network, synthetic
from oracle import ordering


def random_forest(n, seed, p_outlet=0.1):
    """Generation-order forest: node g links to an earlier node (or is an outlet)."""
    rng = np.random.default_rng(seed)
    dn = np.full(n, -1, np.int64)
    for g in range(1, n):
        if rng.random() >= p_outlet:
            dn[g] = rng.integers(max(0, g - 7), g)
    ids = rng.permutation(n).astype(np.int64) * 3 + 5
    down_ids = np.where(dn >= 0, ids[np.maximum(dn, 0)], -1)
    return ids, down_ids


@pytest.mark.parametrize("seed", range(8))
def test_trace_topo_reorder_match_restatement(seed):
    ids, down = random_forest(200, seed)
    assert np.array_equal(network.trace_outlets(ids, down), ordering.trace_outlets(ids, down))
    topo = network.topological_sort(ids, down)
    assert np.array_equal(topo, ordering.topological_sort(ids, down))
    basin = network.trace_outlets(ids, down)
    o1, s1 = network.reorder_by_basin_size(topo, basin)
    o2, s2 = ordering.reorder_by_basin_size(topo, basin)
    assert np.array_equal(o1, o2) and np.array_equal(s1, s2)


def test_topological_sort_is_kahn_fifo_on_a_textbook_graph():
    # 0->2, 1->2, 2->4, 3->4 (ids = indices); Kahn FIFO from headwaters: 0 1 3 2 4
    ids = np.arange(5)
    down = np.array([2, 2, 4, 4, -1])
    assert network.topological_sort(ids, down).tolist() == [0, 1, 3, 2, 4]


def test_ring_raises_like_the_reference():
    ids = np.array([1, 2, 3])
    down = np.array([2, 3, 1])
    with pytest.raises(ValueError, match="Rings exist"):
        network.topological_sort(ids, down)


def test_finalize_connectivity_invariants():
    ids, down = random_forest(500, 3)
    c = network.finalize_connectivity(ids, down)
    n = ids.shape[0]
    idx = np.arange(n)
    mouth = c.downstream_idx == idx
    assert np.all(c.downstream_idx[~mouth] > idx[~mouth])          # merit_map.py:1063-1070
    assert np.array_equal(np.sort(c.order), np.arange(n))          # permutation
    assert np.array_equal(c.downstream_id[mouth], c.catchment_id[mouth])
    # basins are contiguous and sized in descending order
    b = c.catchment_basin_id
    assert np.all(np.diff(b) >= 0)
    assert np.all(np.diff(np.bincount(b)) <= 0)


def test_check_flow_direction_raises():
    with pytest.raises(ValueError, match="Flow direction error"):
        network.check_flow_direction(np.array([1, 0, 2], np.int32))


@pytest.mark.parametrize("seed", range(5))
def test_level_order_matches_restatement(seed):
    ids, down = random_forest(300, seed, p_outlet=0.05)
    c = network.finalize_connectivity(ids, down)
    lo = network.level_order(c.downstream_idx)
    perm, inv, up_ptr, dn, nl = ordering.level_order(c.downstream_idx.tolist())
    assert np.array_equal(lo.perm, perm)
    assert np.array_equal(lo.inv, inv)
    assert np.array_equal(lo.up_ptr, up_ptr)
    assert np.array_equal(lo.dn, dn)
    assert lo.num_levels == nl


def test_level_order_properties_at_config0_size():
    net = synthetic.build_config(0)
    d = net.fields["downstream_idx"]
    lo = network.level_order(d)
    n = d.shape[0]
    k = np.arange(n)
    mouth = lo.dn == k
    assert np.all(lo.dn[~mouth] > k[~mouth])
    assert np.array_equal(lo.inv[lo.perm], k)
    # upstream ranges are exactly the catchments whose downstream is k
    counts = np.bincount(lo.dn[~mouth], minlength=n)
    assert np.array_equal(np.diff(lo.up_ptr), counts)
    owner = np.repeat(k, np.diff(lo.up_ptr))
    assert np.array_equal(lo.dn[:lo.up_ptr[-1]], owner)
    # inside each range the network order is ascending (accumulation order)
    up_net = lo.perm[: lo.up_ptr[-1]]
    same = owner[1:] == owner[:-1]
    assert np.all(up_net[1:][same] > up_net[:-1][same])


@pytest.mark.parametrize("seed", range(5))
def test_basin_level_order_matches_restatement(seed):
    ids, down = random_forest(300, seed, p_outlet=0.05)
    c = network.finalize_connectivity(ids, down)
    bo = network.basin_level_order(c.downstream_idx)
    perm, inv, up_ptr, dn, basin_ptr = ordering.basin_level_order(c.downstream_idx.tolist())
    assert np.array_equal(bo.perm, perm)
    assert np.array_equal(bo.inv, inv)
    assert np.array_equal(bo.up_ptr, up_ptr)
    assert np.array_equal(bo.dn, dn)
    assert np.array_equal(bo.basin_ptr, basin_ptr)


def test_basin_level_order_properties_at_config0_size():
    net = synthetic.build_config(0)
    d = net.fields["downstream_idx"]
    bo = network.basin_level_order(d)
    n = d.shape[0]
    k = np.arange(n)
    mouth = bo.dn == k
    assert np.all(bo.dn[~mouth] > k[~mouth])
    assert np.array_equal(bo.inv[bo.perm], k)
    # each basin is contiguous and ends with its mouth
    assert bo.basin_ptr[0] == 0 and bo.basin_ptr[-1] == n
    ends = bo.basin_ptr[1:] - 1
    assert np.array_equal(np.flatnonzero(mouth), ends)
    basin = np.repeat(np.arange(bo.num_basins), np.diff(bo.basin_ptr))
    assert np.array_equal(basin[bo.dn], basin)
    # ranges tile [0, up_ptr[n]): the catchments draining to k, plus k itself
    # at the end of a mouth's range
    assert bo.up_ptr[0] == 0 and np.all(np.diff(bo.up_ptr) >= 0)
    owner = np.repeat(k, np.diff(bo.up_ptr))
    members = np.arange(bo.up_ptr[-1])
    own = members == owner
    assert np.all(mouth[owner[own]])
    assert np.all(bo.up_ptr[owner[own] + 1] - 1 == members[own])
    assert np.array_equal(bo.dn[members[~own]], owner[~own])
    assert np.array_equal(np.bincount(owner[~own], minlength=n), np.bincount(bo.dn[~mouth], minlength=n))
    up_net = bo.perm[members[~own]]
    same = owner[~own][1:] == owner[~own][:-1]
    assert np.all(up_net[1:][same] > up_net[:-1][same])


def test_find_indices_and_accumulate():
    ref = np.array([40, 10, 30, 20])
    assert network.find_indices_in(np.array([10, 20, 99, 40]), ref).tolist() == [1, 3, -1, 0]
    d = np.array([2, 2, 3, 3], np.int32)
    out = network.accumulate_downstream(d, np.array([1.0, 2.0, 4.0, 8.0]))
    assert out.tolist() == [1.0, 2.0, 7.0, 15.0]


@pytest.mark.parametrize("seed", range(6))
def test_keyed_level_order_matches_restatement(seed):
    ids, down = random_forest(400, seed, p_outlet=0.05)
    c = network.finalize_connectivity(ids, down)
    d = c.downstream_idx
    rng = np.random.default_rng(100 + seed)
    key = rng.choice(np.array([0, 0, 0, 1, 2, 3, 6], np.uint8), size=d.shape[0])
    ko = network.keyed_level_order(d, key)
    perm, inv, dn, levels = ordering.keyed_level_order(d.tolist(), key.tolist())
    assert np.array_equal(ko.perm, perm)
    assert np.array_equal(ko.inv, inv)
    assert np.array_equal(ko.dn, dn)
    assert ko.num_levels == levels
    # no key: the plain forest-wide level order
    lo = network.level_order(d)
    assert np.array_equal(network.keyed_level_order(d).perm, lo.perm)


def test_keyed_level_order_on_a_small_tree():
    """Mouth 6 with upstream catchments 3 (levee), 4, 5 (path end); 3 drains
    0 and 1, 5 drains 2 (levee).  Depth 1 (odd, classes in the order 0, 2, 3,
    1) holds one group, {3, 4, 5}, of class 1 | 2 = 3, kept whole and in
    network order.  Depth 2 (even: 1, 3, 2, 0) is formed as {0, 1} (from 3)
    then {2} (from 5), and lists {2} (class 1) before {0, 1} (class 0).
    Levels deepest first."""
    d = np.array([3, 3, 5, 6, 6, 6, 6], np.int32)
    key = np.array([0, 0, 1, 1, 0, 2, 0], np.uint8)
    ko = network.keyed_level_order(d, key)
    assert ko.perm.tolist() == [2, 0, 1, 3, 4, 5, 6]
    assert ko.level_ptr.tolist() == [0, 3, 6, 7]
    # without keys: breadth-first formation order
    assert network.keyed_level_order(d).perm.tolist() == [0, 1, 2, 3, 4, 5, 6]


def test_keyed_level_order_properties_at_config1_size():
    """Every catchment's upstream catchments are one contiguous range in
    ascending network index (the kernels' summation order), downstream(k) > k,
    and the levee catchments fill far fewer 64-catchment waves than in the
    plain level order."""
    net = synthetic.build_config(1)
    f = net.fields
    d = f["downstream_idx"]
    n = d.shape[0]
    key = np.zeros(n, np.uint8)
    key[f["levee_catchment_idx"]] |= 1
    key[f["bifurcation_catchment_idx"]] |= 2
    key[f["bifurcation_downstream_idx"]] |= 2
    ko = network.keyed_level_order(d, key)
    k = np.arange(n)
    mouth = ko.dn == k
    assert np.all(ko.dn[~mouth] > k[~mouth])
    assert np.array_equal(ko.inv[ko.perm], k)
    up = k[~mouth]
    owner = ko.dn[up]
    order = np.lexsort((up, owner))
    up, owner = up[order], owner[order]
    first = np.r_[True, owner[1:] != owner[:-1]]
    gid = np.cumsum(first) - 1
    rank_in_group = np.arange(up.size) - np.flatnonzero(first)[gid]
    assert np.array_equal(up, up[first][gid] + rank_in_group)        # consecutive internal indices
    net_order = ko.perm[up]
    same = ~first[1:]
    assert np.all(net_order[1:][same] > net_order[:-1][same])         # ascending network index

    def levee_waves(perm):
        lev = (key[perm] & 1).astype(bool)[: (n // 64) * 64].reshape(-1, 64)
        return lev.any(1).mean()
    plain = levee_waves(network.level_order(d).perm)
    keyed = levee_waves(ko.perm)
    assert plain > 0.9 and keyed < 0.3, (plain, keyed)


def test_path_device_order():
    """The device order of bifurcation paths: a permutation in which every
    catchment's paths keep their input order (the order its path terms are
    summed in), otherwise ascending by the upstream end's device position;
    without shared ends it is the sort by upstream end."""
    from cmfgpu_amd.models.cama_flood_model import _path_order
    rng = np.random.default_rng(11)
    for trial in range(20):
        n, P = int(rng.integers(5, 60)), int(rng.integers(1, 120))
        ci = rng.integers(0, n, P)
        di = (ci + rng.integers(1, n, P)) % n
        o = _path_order(ci, di)
        assert np.array_equal(np.sort(o), np.arange(P))
        pos = np.empty(P, np.int64)
        pos[o] = np.arange(P)
        for c in range(n):
            inc = np.flatnonzero((ci == c) | (di == c))
            assert np.all(np.diff(pos[inc]) > 0), (trial, c)
    # distinct ends everywhere: plain sort by upstream end
    ci = rng.permutation(200)[:50]
    di = 200 + rng.permutation(50)
    o = _path_order(ci, di)
    assert np.array_equal(o, np.argsort(ci, kind="stable"))
    assert _path_order(np.zeros(0, np.int32), np.zeros(0, np.int32)).size == 0
