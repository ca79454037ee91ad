// Host-side network indexing for the CaMa-Flood hot path.
//
// Integer work, bit-exact by construction:
//   cmf_trace_outlets          cmfgpu/params/utils.py:20-80   (trace_outlets)
//   cmf_topological_sort       cmfgpu/params/utils.py:83-138  (topological_sort, Kahn FIFO)
//   cmf_reorder_by_basin_size  cmfgpu/params/utils.py:483-503 (reorder_by_basin_size)
//   cmf_find_indices_in        hydroforge find_indices_in, used at merit_map.py:450
//   cmf_check_flow_direction   cmfgpu/params/merit_map.py:1063-1070
//   cmf_accumulate_downstream  cmfgpu/params/estimate_river_geometry.py:76-104
// and the build's own renumbering:
//   cmf_level_order            breadth-first levels from the mouths; see the
//                              comment on the function.
//   cmf_basin_level_order      the same levels laid out basin by basin.
//   cmf_keyed_level_order      the forest-wide levels with each level's
//                              sibling groups arranged by the work they carry.

#include "cmf_internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using cmf::set_error;

extern "C" {

int cmf_trace_outlets(int64_t n, const int64_t* catchment_id,
                      const int64_t* downstream_id, int64_t* out_mouth) {
    if (n < 0 || (n > 0 && (!catchment_id || !downstream_id || !out_mouth)))
        return set_error(CMF_ERR_INVALID, "cmf_trace_outlets: null pointer or negative size");
    int64_t max_id = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (catchment_id[i] < 0) return set_error(CMF_ERR_INVALID, "cmf_trace_outlets: negative catchment id");
        max_id = std::max(max_id, catchment_id[i]);
        max_id = std::max(max_id, downstream_id[i]);
    }
    // down[id]: >= 0 downstream id (unresolved), -1 mouth, -2 not a cell,
    // <= -3 resolved with outlet -(value + 3)
    std::vector<int64_t> down((size_t)max_id + 1, -2);
    for (int64_t i = 0; i < n; ++i) down[catchment_id[i]] = downstream_id[i];
    for (int64_t i = 0; i < n; ++i) {
        const int64_t cid = catchment_id[i];
        if (down[cid] <= -3) { out_mouth[i] = -(down[cid] + 3); continue; }
        int64_t cur = cid;
        int64_t guard = 0;
        while (down[cur] >= 0) {
            cur = down[cur];
            if (++guard > n + 1) return set_error(CMF_ERR_INVALID, "cmf_trace_outlets: ring in network");
        }
        const int64_t v = down[cur];
        const int64_t mouth = (v <= -3) ? -(v + 3) : cur;
        cur = cid;
        while (down[cur] >= 0) {
            const int64_t nxt = down[cur];
            down[cur] = -(mouth + 3);
            cur = nxt;
        }
        down[cur] = -(mouth + 3);
        out_mouth[i] = mouth;
    }
    return CMF_OK;
}

int cmf_topological_sort(int64_t n, const int64_t* catchment_id,
                         const int64_t* downstream_id, int32_t* out_idx) {
    if (n < 0 || n > INT32_MAX || (n > 0 && (!catchment_id || !downstream_id || !out_idx)))
        return set_error(CMF_ERR_INVALID, "cmf_topological_sort: bad arguments");
    int64_t max_id = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (catchment_id[i] < 0) return set_error(CMF_ERR_INVALID, "cmf_topological_sort: negative catchment id");
        max_id = std::max(max_id, catchment_id[i]);
    }
    const int64_t total = max_id + 1;
    std::vector<int64_t> grid_to_idx((size_t)total, -1);
    for (int64_t i = 0; i < n; ++i) grid_to_idx[catchment_id[i]] = i;
    auto down_of = [&](int64_t u) -> int64_t {
        const int64_t d = downstream_id[u];
        if (d >= 0 && d < total) return grid_to_idx[d];
        return -1;
    };
    std::vector<int32_t> indeg((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t d = down_of(i);
        if (d >= 0) indeg[d]++;
    }
    std::vector<int32_t> q((size_t)n);
    int64_t front = 0, back = 0;
    for (int64_t i = 0; i < n; ++i) if (indeg[i] == 0) q[back++] = (int32_t)i;
    int64_t count = 0;
    while (front < back) {
        const int32_t u = q[front++];
        out_idx[count++] = u;
        const int64_t d = down_of(u);
        if (d >= 0) {
            if (--indeg[d] == 0) q[back++] = (int32_t)d;
        }
    }
    if (count != n) return set_error(CMF_ERR_INVALID, "Rings exist and topological sorting is not possible");
    return CMF_OK;
}

int cmf_reorder_by_basin_size(int64_t n, const int64_t* topo_idx, const int64_t* basin_id,
                              int64_t* out_order, int64_t* out_basin_sizes,
                              int64_t* out_num_basins) {
    if (n < 0 || (n > 0 && (!topo_idx || !basin_id || !out_order || !out_num_basins)))
        return set_error(CMF_ERR_INVALID, "cmf_reorder_by_basin_size: bad arguments");
    // topo_basin = basin_id[topo_idx]; unique labels (sorted) and counts
    std::vector<int64_t> topo_basin((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (topo_idx[i] < 0 || topo_idx[i] >= n)
            return set_error(CMF_ERR_INVALID, "cmf_reorder_by_basin_size: index out of range");
        topo_basin[i] = basin_id[topo_idx[i]];
    }
    std::vector<int64_t> labels(topo_basin);
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
    const int64_t nb = (int64_t)labels.size();
    std::vector<int64_t> inverse((size_t)n), counts((size_t)nb, 0);
    for (int64_t i = 0; i < n; ++i) {
        inverse[i] = std::lower_bound(labels.begin(), labels.end(), topo_basin[i]) - labels.begin();
        counts[inverse[i]]++;
    }
    // basins by descending size; ties by ascending label (stable)
    std::vector<int64_t> size_order((size_t)nb);
    std::iota(size_order.begin(), size_order.end(), 0);
    std::stable_sort(size_order.begin(), size_order.end(),
                     [&](int64_t a, int64_t b) { return counts[a] > counts[b]; });
    std::vector<int64_t> rank((size_t)nb);
    for (int64_t r = 0; r < nb; ++r) rank[size_order[r]] = r;
    // stable sort of the topological order by basin rank
    std::vector<int64_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int64_t a, int64_t b) { return rank[inverse[a]] < rank[inverse[b]]; });
    for (int64_t i = 0; i < n; ++i) out_order[i] = topo_idx[perm[i]];
    if (out_basin_sizes)
        for (int64_t r = 0; r < nb; ++r) out_basin_sizes[r] = counts[size_order[r]];
    *out_num_basins = nb;
    return CMF_OK;
}

int cmf_find_indices_in(int64_t n_query, const int64_t* query, int64_t n_ref,
                        const int64_t* ref, int64_t* out) {
    if (n_query < 0 || n_ref < 0 || (n_query > 0 && (!query || !out)) || (n_ref > 0 && !ref))
        return set_error(CMF_ERR_INVALID, "cmf_find_indices_in: bad arguments");
    std::vector<int64_t> order((size_t)n_ref);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ref[a] < ref[b]; });
    for (int64_t k = 1; k < n_ref; ++k)
        if (ref[order[k]] == ref[order[k - 1]])
            return set_error(CMF_ERR_INVALID, "cmf_find_indices_in: duplicate reference id");
    for (int64_t i = 0; i < n_query; ++i) {
        const int64_t v = query[i];
        int64_t lo = 0, hi = n_ref;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (ref[order[mid]] < v) lo = mid + 1; else hi = mid;
        }
        out[i] = (lo < n_ref && ref[order[lo]] == v) ? order[lo] : -1;
    }
    return CMF_OK;
}

int cmf_check_flow_direction(int64_t n, const int32_t* downstream_idx) {
    if (n < 0 || (n > 0 && !downstream_idx))
        return set_error(CMF_ERR_INVALID, "cmf_check_flow_direction: bad arguments");
    for (int64_t i = 0; i < n; ++i) {
        const int64_t d = downstream_idx[i];
        if (d == i) continue;
        if (d <= i || d >= n)
            return set_error(CMF_ERR_INVALID,
                             "Flow direction error: downstream catchment should have higher index");
    }
    return CMF_OK;
}

int cmf_accumulate_downstream(int64_t n, const int32_t* downstream_idx,
                              const double* local, double* out) {
    if (n < 0 || (n > 0 && (!downstream_idx || !local || !out)))
        return set_error(CMF_ERR_INVALID, "cmf_accumulate_downstream: bad arguments");
    std::memmove(out, local, sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t d = downstream_idx[i];
        if (d >= 0 && d < n && d != i) out[d] += out[i];
    }
    return CMF_OK;
}

// Breadth-first level renumbering.
//
// depth(mouth) = 0, depth(i) = depth(downstream(i)) + 1.  Level d holds the
// catchments of depth d; level 0 lists the mouths in network order, and level
// d + 1 lists, for each catchment of level d in turn, its upstream catchments in
// ascending network index.  The internal order places the deepest level first
// and the mouths last, so downstream(k) > k for every non-mouth k, and the
// upstream catchments of internal catchment j are exactly the contiguous range
// [up_ptr[j], up_ptr[j + 1]), with up_ptr monotone over the whole network.
//
//   out_perm[k]       network index of internal catchment k
//   out_inv[i]        internal index of network catchment i
//   out_up_ptr[k]     upstream range start of internal catchment k (n + 1 entries)
//   out_dn[k]         internal downstream index of internal catchment k
//   out_level_ptr     start of each level in internal order, deepest first
//                     (num_levels + 1 entries; pass NULL to skip)
int cmf_level_order(int64_t n, const int32_t* downstream_idx, int32_t* out_perm,
                    int32_t* out_inv, int32_t* out_up_ptr, int32_t* out_dn,
                    int64_t* out_num_levels, int64_t level_ptr_capacity,
                    int64_t* out_level_ptr) {
    if (n < 0 || n > INT32_MAX - 1 ||
        (n > 0 && (!downstream_idx || !out_perm || !out_inv || !out_up_ptr || !out_dn)) ||
        !out_num_levels)
        return set_error(CMF_ERR_INVALID, "cmf_level_order: bad arguments");
    int rc = cmf_check_flow_direction(n, downstream_idx);
    if (rc != CMF_OK) return rc;
    // children lists in ascending network index
    std::vector<int32_t> cptr((size_t)n + 1, 0), cidx;
    for (int64_t i = 0; i < n; ++i) if (downstream_idx[i] != i) cptr[downstream_idx[i] + 1]++;
    for (int64_t i = 0; i < n; ++i) cptr[i + 1] += cptr[i];
    cidx.resize((size_t)cptr[n]);
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            if (downstream_idx[i] != i) cidx[fill[downstream_idx[i]]++] = (int32_t)i;
    }
    // breadth-first traversal from the mouths
    std::vector<int32_t> bfs;
    bfs.reserve((size_t)n);
    std::vector<int64_t> lstart;
    for (int64_t i = 0; i < n; ++i) if (downstream_idx[i] == i) bfs.push_back((int32_t)i);
    lstart.push_back(0);
    int64_t head = 0;
    while (head < (int64_t)bfs.size()) {
        const int64_t end = (int64_t)bfs.size();
        lstart.push_back(end);
        for (int64_t k = head; k < end; ++k) {
            const int32_t u = bfs[k];
            for (int32_t c = cptr[u]; c < cptr[u + 1]; ++c) bfs.push_back(cidx[c]);
        }
        head = end;
    }
    if ((int64_t)bfs.size() != n)
        return set_error(CMF_ERR_INVALID, "cmf_level_order: network is not a forest");
    // lstart holds level boundaries in BFS order; the last entry repeats n
    while (lstart.size() >= 2 && lstart[lstart.size() - 1] == lstart[lstart.size() - 2]) lstart.pop_back();
    const int64_t num_levels = (int64_t)lstart.size() - 1;
    *out_num_levels = num_levels;
    // internal order: levels from deepest to 0, BFS order inside a level
    int64_t pos = 0;
    std::vector<int64_t> ilevel_start((size_t)num_levels + 1);
    for (int64_t d = num_levels - 1; d >= 0; --d) {
        ilevel_start[num_levels - 1 - d] = pos;
        for (int64_t k = lstart[d]; k < lstart[d + 1]; ++k) out_perm[pos++] = bfs[k];
    }
    ilevel_start[num_levels] = n;
    for (int64_t k = 0; k < n; ++k) out_inv[out_perm[k]] = (int32_t)k;
    for (int64_t k = 0; k < n; ++k) out_dn[k] = out_inv[downstream_idx[out_perm[k]]];
    // upstream ranges: the children of consecutive catchments are consecutive
    int64_t acc = 0;
    for (int64_t k = 0; k < n; ++k) {
        out_up_ptr[k] = (int32_t)acc;
        const int32_t u = out_perm[k];
        const int32_t cnt = cptr[u + 1] - cptr[u];
        if (cnt > 0 && out_inv[cidx[cptr[u]]] != acc)
            return set_error(CMF_ERR_INTERNAL, "cmf_level_order: upstream ranges are not contiguous");
        acc += cnt;
    }
    out_up_ptr[n] = (int32_t)acc;
    if (out_level_ptr) {
        if (level_ptr_capacity < num_levels + 1)
            return set_error(CMF_ERR_INVALID, "cmf_level_order: level_ptr buffer too small");
        for (int64_t d = 0; d <= num_levels; ++d) out_level_ptr[d] = ilevel_start[d];
    }
    return CMF_OK;
}

// Breadth-first levels inside each basin.
//
// A basin is the tree that drains to one mouth.  Basins are laid out one after
// another in ascending network index of their mouths; inside a basin the
// levels of that tree follow deepest first, each built as in cmf_level_order
// (upstream catchments of consecutive catchments are consecutive, in
// ascending network index).  So downstream(k) > k, upstream ranges stay
// contiguous with up_ptr monotone over the whole network, and a catchment and
// its downstream neighbour lie one tree level apart instead of one level of the
// whole forest apart.
//
//   out_basin_ptr     start of each basin in internal order (num_basins + 1
//                     entries, capacity basin_ptr_capacity; NULL to skip)
int cmf_basin_level_order(int64_t n, const int32_t* downstream_idx, int32_t* out_perm,
                          int32_t* out_inv, int32_t* out_up_ptr, int32_t* out_dn,
                          int64_t* out_num_basins, int64_t basin_ptr_capacity,
                          int64_t* out_basin_ptr) {
    if (n < 0 || n > INT32_MAX - 1 ||
        (n > 0 && (!downstream_idx || !out_perm || !out_inv || !out_up_ptr || !out_dn)) ||
        !out_num_basins)
        return set_error(CMF_ERR_INVALID, "cmf_basin_level_order: bad arguments");
    int rc = cmf_check_flow_direction(n, downstream_idx);
    if (rc != CMF_OK) return rc;
    std::vector<int32_t> cptr((size_t)n + 1, 0), cidx;
    for (int64_t i = 0; i < n; ++i) if (downstream_idx[i] != i) cptr[downstream_idx[i] + 1]++;
    for (int64_t i = 0; i < n; ++i) cptr[i + 1] += cptr[i];
    cidx.resize((size_t)cptr[n]);
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            if (downstream_idx[i] != i) cidx[fill[downstream_idx[i]]++] = (int32_t)i;
    }
    std::vector<int32_t> bfs;
    std::vector<int64_t> lstart;
    std::vector<int64_t> bstart;
    int64_t pos = 0;
    for (int64_t m = 0; m < n; ++m) {
        if (downstream_idx[m] != m) continue;
        bstart.push_back(pos);
        bfs.clear();
        lstart.clear();
        bfs.push_back((int32_t)m);
        lstart.push_back(0);
        int64_t head = 0;
        while (head < (int64_t)bfs.size()) {
            const int64_t end = (int64_t)bfs.size();
            lstart.push_back(end);
            for (int64_t k = head; k < end; ++k) {
                const int32_t u = bfs[k];
                for (int32_t c = cptr[u]; c < cptr[u + 1]; ++c) bfs.push_back(cidx[c]);
            }
            head = end;
        }
        if (pos + (int64_t)bfs.size() > n)
            return set_error(CMF_ERR_INVALID, "cmf_basin_level_order: network is not a forest");
        // lstart[d]..lstart[d+1] is level d; emit deepest first
        for (int64_t d = (int64_t)lstart.size() - 2; d >= 0; --d)
            for (int64_t k = lstart[d]; k < lstart[d + 1]; ++k) out_perm[pos++] = bfs[k];
    }
    if (pos != n)
        return set_error(CMF_ERR_INVALID, "cmf_basin_level_order: network is not a forest");
    const int64_t nb = (int64_t)bstart.size();
    *out_num_basins = nb;
    for (int64_t k = 0; k < n; ++k) out_inv[out_perm[k]] = (int32_t)k;
    for (int64_t k = 0; k < n; ++k) out_dn[k] = out_inv[downstream_idx[out_perm[k]]];
    // A mouth sits right after its upstream range, so its range also covers
    // the mouth itself (the catchment after it starts the next basin); users
    // of the ranges skip the owner.
    int64_t acc = 0;
    for (int64_t k = 0; k < n; ++k) {
        out_up_ptr[k] = (int32_t)acc;
        const int32_t u = out_perm[k];
        const int32_t cnt = cptr[u + 1] - cptr[u];
        if (cnt > 0 && out_inv[cidx[cptr[u]]] != acc)
            return set_error(CMF_ERR_INTERNAL, "cmf_basin_level_order: upstream ranges are not contiguous");
        acc += cnt;
        if (out_dn[k] == k && acc == k) acc += 1;
    }
    out_up_ptr[n] = (int32_t)acc;
    if (out_basin_ptr) {
        if (basin_ptr_capacity < nb + 1)
            return set_error(CMF_ERR_INVALID, "cmf_basin_level_order: basin_ptr buffer too small");
        for (int64_t b = 0; b < nb; ++b) out_basin_ptr[b] = bstart[b];
        out_basin_ptr[nb] = n;
    }
    return CMF_OK;
}

// Breadth-first levels of the whole forest, each level's sibling groups (the
// upstream catchments of one catchment, in ascending network index) arranged
// by the class of work they carry.  key[i] (network order; NULL: all 0) is a
// small class code of catchment i (the model passes bit 0 for a levee, bit 1
// for a bifurcation path end); a group's class is the OR of its members'
// (a mouth is a group of its own).  Each level lists its groups class by
// class, stable inside a class, in the order
//     1, 3, 2, 0, then every other class value ascending    (even depth)
// and the reverse of that on odd depths, so that the class a level ends with
// is the one the next level in memory (one nearer the mouths: deepest level
// first) starts with.  A wave of consecutive catchments then mostly holds one
// class, and the kernels' per-class work (the levee partition, the path
// nets) runs in few waves.  Level d + 1 is formed from level d's order as in
// cmf_level_order, so every catchment's upstream catchments are still one
// contiguous range in ascending network index, but the ranges no longer
// follow the order of their owners (there is no monotone up_ptr: the device
// network keeps a start and an end per catchment).
//
//   out_perm, out_inv, out_dn, out_num_levels, out_level_ptr: as cmf_level_order
int cmf_keyed_level_order(int64_t n, const int32_t* downstream_idx, const uint8_t* key, int32_t* out_perm,
                          int32_t* out_inv, int32_t* out_dn, int64_t* out_num_levels,
                          int64_t level_ptr_capacity, int64_t* out_level_ptr) {
    if (n < 0 || n > INT32_MAX - 1 || (n > 0 && (!downstream_idx || !out_perm || !out_inv || !out_dn)) ||
        !out_num_levels)
        return set_error(CMF_ERR_INVALID, "cmf_keyed_level_order: bad arguments");
    int rc = cmf_check_flow_direction(n, downstream_idx);
    if (rc != CMF_OK) return rc;
    std::vector<int32_t> cptr((size_t)n + 1, 0), cidx;
    for (int64_t i = 0; i < n; ++i) if (downstream_idx[i] != i) cptr[downstream_idx[i] + 1]++;
    for (int64_t i = 0; i < n; ++i) cptr[i + 1] += cptr[i];
    cidx.resize((size_t)cptr[n]);
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            if (downstream_idx[i] != i) cidx[fill[downstream_idx[i]]++] = (int32_t)i;
    }
    // class rank: 1, 3, 2, 0, then 4..255 ascending
    int rank[256];
    {
        const int lead[4] = {1, 3, 2, 0};
        int r = 0;
        for (int c : lead) rank[c] = r++;
        for (int c = 4; c < 256; ++c) rank[c] = r++;
    }
    auto cls = [&](int32_t i) -> int { return key ? (int)key[i] : 0; };
    // a group: [begin, end) in `members`; sorted by (rank, position) per level
    struct Grp { int64_t begin, end; int c; };
    std::vector<std::vector<int32_t>> levels;
    std::vector<int32_t> members;
    std::vector<Grp> groups;
    // level 0: each mouth a group of its own, in ascending network index
    for (int64_t i = 0; i < n; ++i) {
        if (downstream_idx[i] != i) continue;
        groups.push_back({(int64_t)members.size(), (int64_t)members.size() + 1, cls((int32_t)i)});
        members.push_back((int32_t)i);
    }
    int64_t placed = 0;
    for (int64_t d = 0; !groups.empty(); ++d) {
        const bool odd = (d & 1) != 0;
        std::stable_sort(groups.begin(), groups.end(), [&](const Grp& a, const Grp& b) {
            return odd ? rank[a.c] > rank[b.c] : rank[a.c] < rank[b.c];
        });
        std::vector<int32_t> level;
        level.reserve((size_t)(members.size()));
        for (const Grp& g : groups)
            for (int64_t k = g.begin; k < g.end; ++k) level.push_back(members[(size_t)k]);
        placed += (int64_t)level.size();
        if (placed > n) return set_error(CMF_ERR_INVALID, "cmf_keyed_level_order: network is not a forest");
        // the next level's groups: the upstream catchments of this level's, in its order
        members.clear();
        groups.clear();
        for (int32_t u : level) {
            if (cptr[u + 1] == cptr[u]) continue;
            int c = 0;
            const int64_t b = (int64_t)members.size();
            for (int32_t q = cptr[u]; q < cptr[u + 1]; ++q) {
                members.push_back(cidx[q]);
                c |= cls(cidx[q]);
            }
            groups.push_back({b, (int64_t)members.size(), c});
        }
        levels.push_back(std::move(level));
    }
    if (placed != n) return set_error(CMF_ERR_INVALID, "cmf_keyed_level_order: network is not a forest");
    const int64_t num_levels = (int64_t)levels.size();
    *out_num_levels = num_levels;
    if (out_level_ptr && level_ptr_capacity < num_levels + 1)
        return set_error(CMF_ERR_INVALID, "cmf_keyed_level_order: level_ptr buffer too small");
    int64_t pos = 0;
    for (int64_t d = num_levels - 1; d >= 0; --d) {
        if (out_level_ptr) out_level_ptr[num_levels - 1 - d] = pos;
        for (int32_t c : levels[(size_t)d]) out_perm[pos++] = c;
    }
    if (out_level_ptr) out_level_ptr[num_levels] = n;
    for (int64_t k = 0; k < n; ++k) out_inv[out_perm[k]] = (int32_t)k;
    for (int64_t k = 0; k < n; ++k) out_dn[k] = out_inv[downstream_idx[out_perm[k]]];
    return CMF_OK;
}

}  // extern "C"
