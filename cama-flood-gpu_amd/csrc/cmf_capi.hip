// C ABI of libcmfhip: network context, per-op entry points and the fused
// sub-step schedule.  Declarations and contracts are in include/cmf_hip.h.

#include "cmf_kernels.hip"
#include "cmf_adjoint.hip"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace cmf {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define CMF_HIP_CHECK(expr)                                                            \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return set_error(CMF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

}  // namespace cmf

using namespace cmf;

struct cmf_network {
    int64_t n = 0, T = 0, P = 0;
    int contiguous = 0;
    int device = 0;
    int32_t* up_ptr = nullptr;
    int2* up_rng = nullptr;           // {start, end} where ranges do not follow their owners
    int32_t* up_idx = nullptr;
    int32_t* ev_ptr = nullptr;
    int32_t* ev_code = nullptr;
    int32_t* ev_other = nullptr;
    int2* ev1 = nullptr;
    uint8_t* ev_flag = nullptr;   // 1 where the catchment has a bifurcation event
    void* pos_term = nullptr;
    void* neg_term = nullptr;
    void* qraw_r = nullptr;
    void* qraw_f = nullptr;
    void* lim = nullptr;
    void* path_pos = nullptr;
    void* path_neg = nullptr;
    void* path_net = nullptr;
    void* path_raw = nullptr;
    double* partials = nullptr;
    int64_t partial_blocks = 0;
    int32_t* host_counts = nullptr;   // pinned, max(T, G) entries
    int32_t* host_counts_dev = nullptr;   // its device address (null: copied instead)
    // block size of each tunable kernel of the fused schedule (CMF_TUNE_*)
    int32_t blocks[CMF_TUNE_COUNT] = {256, 256, 256, 256};
    // groups (G > 0): group of each catchment and the block table of the
    // catchment kernels (blocks never straddle two groups)
    int64_t G = 0;
    int64_t nblk = 0;
    int32_t* cgroup = nullptr;
    // block records {first catchment, end, table entry, group} in memory
    // order (rec_id); group g holds records [gblk_ptr[g], gblk_ptr[g + 1])
    int4* blk_rec_id = nullptr;
    int32_t* gblk_ptr = nullptr;
    std::vector<int32_t> gblk_host;
    // adaptive grouped advance: one table per sub-step k of the outer step,
    // the records of the groups with count > k in memory order, at
    // level_rec + level_off[k], sub_blocks[k] records long; built on the
    // device from level_pos[k * G + g] (where group g's records start in
    // level k's table), staged through host_level_pos
    int4* level_rec = nullptr;
    int64_t level_cap = 0;
    int32_t* level_pos = nullptr;
    int32_t* host_level_pos = nullptr;
    int64_t level_pos_cap = 0;
    hipEvent_t order_copied = nullptr;
    std::vector<int64_t> sub_blocks, level_off;
    // plain adaptive advance: the event that marks the sub-step count's arrival
    // in host_counts, stored there by the CFL pass or copied (the first
    // sub-step is enqueued before it is waited on)
    hipEvent_t count_ready = nullptr;
    // captured sub-step loops (cmf_hip_advance_graph): one executable graph
    // per sub-step count, valid for the field table they were captured with
    hipStream_t capture_stream = nullptr;
    cmf_fields graph_fields;
    bool graph_fields_set = false;
    std::vector<std::pair<int, hipGraphExec_t>> graphs;
    // output statistics accumulator (cmf_network_set_outflow_accumulator)
    double* acc_outflow = nullptr;
    // block-local schedule (k_fused_local; one trial, ungrouped, contiguous
    // ranges): its block size (0: not available), the per-catchment flags and
    // the catchments left to the list passes; `schedule` 1 uses it
    int64_t local_block = 0;
    uint8_t* lflags = nullptr;
    int32_t* lim_list = nullptr;
    int64_t lim_count = 0;
    int32_t* stage_list = nullptr;
    int64_t stage_count = 0;
    int schedule = CMF_SCHEDULE_THREE_PASS;
    // three-pass schedule: the flags of the provisional limiters (one bit
    // per element, zero between sub-steps), the path-end catchments, whose
    // limiters the fix-up pass always forms ({catchment, upstream range
    // start, end, first event} in ascending catchment order, and a last
    // record whose .w ends the events), the event slots of each path's two
    // ends and the path terms at those slots (2 P per trial)
    uint32_t* bmask = nullptr;
    int64_t bmask_words = 0;
    int4* pend_list = nullptr;
    int64_t pend_count = 0;
    int2* pev = nullptr;
    void* eterm = nullptr;
    // adjoint scratch (cmf_hip_substep_adjoint), allocated on first use:
    // 14 arrays of N elements of up to 8 bytes
    void* adj_scratch = nullptr;
    // per-block, per-trial CFL maxima of k_cfl_inner (trial counts dividing 64)
    int32_t* cfl_part = nullptr;
    void* adj_path_scratch = nullptr;   // with paths: [P, LB] + [P] x 9
    void* adj_res_rows = nullptr;       // with reservoirs: catchment -> reservoir row
};

namespace {

void free_network(cmf_network* net) {
    if (!net) return;
    void* ptrs[] = {net->up_ptr, net->up_rng, net->up_idx, net->ev_ptr, net->ev_code, net->ev_other, net->ev1, net->pos_term,
                    net->neg_term, net->qraw_r, net->qraw_f, net->lim, net->path_pos,
                    net->path_neg, net->path_net, net->path_raw, net->partials, net->cgroup,
                    net->blk_rec_id, net->level_rec, net->level_pos, net->gblk_ptr, net->lflags, net->lim_list,
                    net->stage_list, net->bmask, net->pend_list, net->pev, net->eterm, net->ev_flag};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (net->host_counts) (void)hipHostFree(net->host_counts);
    if (net->host_level_pos) (void)hipHostFree(net->host_level_pos);
    if (net->order_copied) (void)hipEventDestroy(net->order_copied);
    if (net->count_ready) (void)hipEventDestroy(net->count_ready);
    for (auto& g : net->graphs) (void)hipGraphExecDestroy(g.second);
    if (net->capture_stream) (void)hipStreamDestroy(net->capture_stream);
    if (net->adj_scratch) (void)hipFree(net->adj_scratch);
    if (net->cfl_part) (void)hipFree(net->cfl_part);
    if (net->adj_path_scratch) (void)hipFree(net->adj_path_scratch);
    if (net->adj_res_rows) (void)hipFree(net->adj_res_rows);
    delete net;
}

int dev_alloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(CMF_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return CMF_OK;
}

int upload(void** p, const void* host, size_t bytes) {
    int rc = dev_alloc(p, bytes);
    if (rc) return rc;
    if (bytes) CMF_HIP_CHECK(hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
    return CMF_OK;
}

int64_t blocks_for(int64_t count) { return (count + kBlock - 1) / kBlock; }

// Block size of the block-local schedule (k_fused_local).
constexpr int64_t kLocalBlock = 512;

// Flags of the block-local schedule for blocks of B catchments (contiguous
// upstream ranges [up_ptr, up_end), bifurcation events `ev_ptr`), and the catchments
// it leaves to the limiter and stage list passes (ascending).
//   LL(j): j's upstream range lies in j's block and j is no path end: its
//          limiter comes out of the block's LDS;
//   SL(j): LL(j), LL of every upstream catchment, and j is a mouth or its
//          downstream catchment is in the block with LL: its stage too.
// The scratch each catchment writes is what the list passes read: its
// positive term when not LL (the limiter list), its negative term when its
// downstream catchment is not LL, its unlimited flows when it or its
// downstream catchment is not SL (the stage list), its limiter when LL and
// it, its downstream catchment or one of its upstream catchments is not SL.
void local_tables(int64_t n, const int32_t* dn, const std::vector<int32_t>& up_ptr,
                  const std::vector<int32_t>& up_end, const std::vector<int32_t>& ev_ptr, int64_t B,
                  std::vector<uint8_t>& flags,
                  std::vector<int32_t>& lim_list, std::vector<int32_t>& stage_list) {
    std::vector<uint8_t> ll((size_t)n), sl((size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t b0 = up_ptr[j], b1 = up_end[j];
        const bool in_block = (b0 == b1) || (b0 / B == j / B && (b1 - 1) / B == j / B);
        ll[j] = in_block && ev_ptr[j] == ev_ptr[j + 1];
    }
    for (int64_t j = 0; j < n; ++j) {
        bool ok = ll[j] && (dn[j] == j || ((int64_t)dn[j] / B == j / B && ll[dn[j]]));
        for (int64_t u = up_ptr[j]; ok && u < up_end[j]; ++u)
            if (u != j && !ll[u]) ok = false;
        sl[j] = ok;
    }
    flags.assign((size_t)n, 0);
    for (int64_t j = 0; j < n; ++j) {
        const bool mouth = dn[j] == j;
        uint8_t f = 0;
        if (ll[j]) f |= kLocalLim;
        if (sl[j]) f |= kLocalStage;
        if (!ll[j]) f |= kPosG;
        if (!mouth && !ll[dn[j]]) f |= kNegG;
        const bool stage_reader = !sl[j] || (!mouth && !sl[dn[j]]);
        if (stage_reader) f |= kQrawG;
        bool up_not_sl = false;
        for (int64_t u = up_ptr[j]; u < up_end[j]; ++u)
            if (u != j && !sl[u]) up_not_sl = true;
        if (ll[j] && (stage_reader || up_not_sl)) f |= kLimG;
        flags[j] = f;
        if (!ll[j]) lim_list.push_back((int32_t)j);
        if (!sl[j]) stage_list.push_back((int32_t)j);
    }
}

// --------------------------------------------------------------- view ---
// The field part of a view: everything a cmf_fields table holds; the
// network part (upstream ranges, events, scratch) stays null.
template <class REAL, class STO>
int view_of_fields(const cmf_fields* f, View<REAL, STO>& v) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->num_catchments < 0 || f->num_trials < 1)
        return set_error(CMF_ERR_INVALID, "num_catchments must be >= 0 and num_trials >= 1");
    std::memset(&v, 0, sizeof(v));
    v.N = f->num_catchments;
    v.T = f->num_trials;
    v.L = f->num_flood_levels;
    v.P = (f->flags & CMF_HAS_BIFURCATION) ? f->num_bifurcation_paths : 0;
    v.LB = f->num_bifurcation_levels;
    v.V = (f->flags & CMF_HAS_LEVEE) ? f->num_levees : 0;
    v.R = (f->flags & CMF_HAS_RESERVOIR) ? f->num_reservoirs : 0;
    v.logsize = f->log_buffer_size;
    v.ts = f->trial_stride ? f->trial_stride : v.N;
    v.cs = f->catchment_stride ? f->catchment_stride : 1;
    v.pts = f->path_trial_stride ? f->path_trial_stride : v.P;
    v.pps = f->path_stride ? f->path_stride : 1;
    if (v.T == 1 && (v.cs != 1 || (v.P > 0 && v.pps != 1)))
        return set_error(CMF_ERR_INVALID, "one trial: catchment_stride and path_stride must be 1");
    if (v.T > 1) {
        const bool major = (v.cs == 1 && v.ts == v.N);
        const bool inner = (v.ts == 1 && v.cs == v.T);
        if (!major && !inner)
            return set_error(CMF_ERR_INVALID, "trial layout must be [trials, catchments] or [catchments, trials]");
        const bool pmajor = (v.pps == 1 && v.pts == v.P);
        const bool pinner = (v.pts == 1 && v.pps == v.T);
        if (v.P > 0 && !pmajor && !pinner)
            return set_error(CMF_ERR_INVALID, "bifurcation trial layout must be [trials, paths] or [paths, trials]");
    }
    if (v.L < 1) return set_error(CMF_ERR_INVALID, "num_flood_levels must be >= 1");
    v.flags = f->flags;
    v.batched = f->batched;
    v.gravity = (REAL)f->gravity;
    v.adaptive_factor = (REAL)f->adaptive_time_factor;
    v.min_kin = (REAL)f->min_kinematic_slope;
    v.time_step = static_cast<const REAL*>(f->time_step);
    v.outer_time_step = static_cast<const REAL*>(f->outer_time_step);
    v.current_step = f->current_step;
    v.max_sub_steps = f->max_sub_steps;
    v.dn = f->downstream_idx;
#define CMF_SET(dst, src, type) v.dst = static_cast<type>(f->src)
    CMF_SET(river_inflow, river_inflow, STO*);
    CMF_SET(flood_inflow, flood_inflow, STO*);
    CMF_SET(river_outflow, river_outflow, REAL*);
    CMF_SET(flood_outflow, flood_outflow, REAL*);
    CMF_SET(river_manning, river_manning, const REAL*);
    CMF_SET(flood_manning, flood_manning, const REAL*);
    CMF_SET(river_depth, river_depth, REAL*);
    CMF_SET(flood_depth, flood_depth, REAL*);
    CMF_SET(protected_depth, protected_depth, REAL*);
    CMF_SET(river_width, river_width, const REAL*);
    CMF_SET(river_length, river_length, const REAL*);
    CMF_SET(river_height, river_height, const REAL*);
    CMF_SET(elev, catchment_elevation, const REAL*);
    CMF_SET(catchment_area, catchment_area, const REAL*);
    CMF_SET(dist, downstream_distance, const REAL*);
    CMF_SET(river_storage, river_storage, STO*);
    CMF_SET(flood_storage, flood_storage, STO*);
    CMF_SET(protected_storage, protected_storage, STO*);
    CMF_SET(rcsd, river_cross_section_depth, REAL*);
    CMF_SET(fcsd, flood_cross_section_depth, REAL*);
    CMF_SET(fcsa, flood_cross_section_area, REAL*);
    CMF_SET(flood_fraction, flood_fraction, REAL*);
    CMF_SET(runoff, runoff, const REAL*);
    CMF_SET(table, flood_depth_table, const REAL*);
    CMF_SET(outgoing, outgoing_storage, STO*);
    CMF_SET(total_storage, total_storage, STO*);
    CMF_SET(wse, water_surface_elevation, REAL*);
    CMF_SET(pwse, protected_water_surface_elevation, REAL*);
    v.is_levee = f->is_levee;
    v.bif_ci = f->bifurcation_catchment_idx;
    v.bif_di = f->bifurcation_downstream_idx;
    CMF_SET(bif_manning, bifurcation_manning, const REAL*);
    CMF_SET(bif_width, bifurcation_width, const REAL*);
    CMF_SET(bif_length, bifurcation_length, const REAL*);
    CMF_SET(bif_elev, bifurcation_elevation, const REAL*);
    CMF_SET(bif_outflow, bifurcation_outflow, REAL*);
    CMF_SET(bif_csd, bifurcation_cross_section_depth, REAL*);
    CMF_SET(gbif, global_bifurcation_outflow, STO*);
    CMF_SET(limit_rate, limit_rate, REAL*);
    v.levee_idx = f->levee_catchment_idx;
    CMF_SET(levee_base, levee_base_height, const REAL*);
    CMF_SET(levee_crown, levee_crown_height, const REAL*);
    CMF_SET(levee_frac, levee_fraction, const REAL*);
    v.res_idx = f->reservoir_catchment_idx;
    CMF_SET(res_cons, conservation_volume, const REAL*);
    CMF_SET(res_emer, emergency_volume, const REAL*);
    CMF_SET(res_adj, adjustment_volume, const REAL*);
    CMF_SET(res_qn, effective_normal_outflow, const REAL*);
    CMF_SET(res_qa, adjustment_outflow, const REAL*);
    CMF_SET(res_qf, flood_control_outflow, const REAL*);
    CMF_SET(res_inflow, reservoir_total_inflow, STO*);
#undef CMF_SET
    v.is_reservoir = f->is_reservoir;
    v.is_dam_up = f->is_dam_upstream;
    v.is_dam_related = f->is_dam_related;
    v.levee_index = f->levee_index;
    for (int k = 0; k < CMF_LOG_COUNT; ++k) v.log_sums[k] = f->log_sums[k];
    return CMF_OK;
}

template <class REAL, class STO>
int make_view(const cmf_network* net, const cmf_fields* f, View<REAL, STO>& v) {
    if (!net || !f) return set_error(CMF_ERR_INVALID, "null network or field table");
    if (f->num_catchments != net->n)
        return set_error(CMF_ERR_INVALID, "num_catchments %lld does not match the network (%lld)",
                         (long long)f->num_catchments, (long long)net->n);
    if (f->num_trials < 1 || f->num_trials > net->T)
        return set_error(CMF_ERR_INVALID, "num_trials %lld outside [1, %lld] of the network",
                         (long long)f->num_trials, (long long)net->T);
    if ((f->flags & CMF_HAS_BIFURCATION) && f->num_bifurcation_paths != net->P)
        return set_error(CMF_ERR_INVALID, "num_bifurcation_paths does not match the network");
    const int rc = view_of_fields<REAL, STO>(f, v);
    if (rc) return rc;
    v.acc_outflow = net->acc_outflow;
    v.G = net->G;
    v.cgroup = net->cgroup;
    v.blk_rec = net->blk_rec_id;
    v.ordered = false;
    v.up_ptr = net->up_ptr;
    v.up_idx = net->contiguous ? nullptr : net->up_idx;
    v.up_rng = net->up_rng;
    v.ev_ptr = net->ev_ptr;
    v.ev_code = net->ev_code;
    v.ev_other = net->ev_other;
    v.ev1 = net->ev1;
    v.ev_flag = net->ev_flag;
    v.pos_term = static_cast<REAL*>(net->pos_term);
    v.neg_term = static_cast<REAL*>(net->neg_term);
    v.qraw_r = static_cast<REAL*>(net->qraw_r);
    v.qraw_f = static_cast<REAL*>(net->qraw_f);
    v.lim = static_cast<REAL*>(net->lim);
    v.path_pos = static_cast<REAL*>(net->path_pos);
    v.path_neg = static_cast<REAL*>(net->path_neg);
    v.path_net = static_cast<REAL*>(net->path_net);
    v.path_raw = static_cast<REAL*>(net->path_raw);
    v.partials = net->partials;
    v.bmask = net->bmask;
    v.pev = net->pev;
    v.eterm = nullptr;
    return CMF_OK;
}

// Required-pointer checks, grouped by the fields each op touches.
struct Need {
    const char* name;
    const void* p;
};
int require(std::initializer_list<Need> needs, const char* op) {
    for (const Need& n : needs)
        if (!n.p) return set_error(CMF_ERR_INVALID, "%s: field %s is NULL", op, n.name);
    return CMF_OK;
}
#define NEED(field) Need{#field, (const void*)f->field}

int need_base(const cmf_fields* f, const char* op) {
    return require({NEED(time_step), NEED(downstream_idx), NEED(river_inflow), NEED(flood_inflow),
                    NEED(river_outflow), NEED(flood_outflow), NEED(river_manning), NEED(flood_manning),
                    NEED(river_depth), NEED(flood_depth), NEED(protected_depth), NEED(river_width),
                    NEED(river_length), NEED(river_height), NEED(catchment_elevation),
                    NEED(catchment_area), NEED(downstream_distance), NEED(river_storage),
                    NEED(flood_storage), NEED(protected_storage), NEED(river_cross_section_depth),
                    NEED(flood_cross_section_depth), NEED(flood_cross_section_area),
                    NEED(flood_fraction), NEED(runoff), NEED(flood_depth_table)},
                   op);
}
int need_features(const cmf_fields* f, const char* op) {
    int rc;
    if (f->flags & CMF_HAS_TOTAL_STORAGE) if ((rc = require({NEED(total_storage)}, op))) return rc;
    if (f->flags & CMF_HAS_WATER_SURFACE) if ((rc = require({NEED(water_surface_elevation)}, op))) return rc;
    if (f->flags & CMF_HAS_PROTECTED_WATER_SURFACE)
        if ((rc = require({NEED(protected_water_surface_elevation)}, op))) return rc;
    if (f->flags & CMF_HAS_LEVEE)
        if ((rc = require({NEED(is_levee), NEED(levee_catchment_idx), NEED(levee_base_height),
                           NEED(levee_crown_height), NEED(levee_fraction)}, op))) return rc;
    if (f->flags & CMF_HAS_BIFURCATION) {
        if ((rc = require({NEED(global_bifurcation_outflow), NEED(limit_rate), NEED(water_surface_elevation),
                           NEED(total_storage)}, op)))
            return rc;
        // the per-path arrays (empty, so possibly NULL, without paths)
        if (f->num_bifurcation_paths > 0 &&
            (rc = require({NEED(bifurcation_catchment_idx), NEED(bifurcation_downstream_idx),
                           NEED(bifurcation_manning), NEED(bifurcation_outflow), NEED(bifurcation_width),
                           NEED(bifurcation_length), NEED(bifurcation_elevation),
                           NEED(bifurcation_cross_section_depth)}, op)))
            return rc;
        if (f->num_bifurcation_levels < 1)
            return set_error(CMF_ERR_INVALID, "%s: num_bifurcation_levels must be >= 1", op);
    }
    if (f->flags & CMF_HAS_RESERVOIR)
        if ((rc = require({NEED(reservoir_catchment_idx), NEED(conservation_volume), NEED(emergency_volume),
                           NEED(adjustment_volume), NEED(effective_normal_outflow),
                           NEED(adjustment_outflow), NEED(flood_control_outflow),
                           NEED(reservoir_total_inflow), NEED(is_reservoir), NEED(is_dam_upstream),
                           NEED(is_dam_related), NEED(total_storage)}, op))) return rc;
    if (f->flags & CMF_HAS_LOG) {
        for (int k = 0; k < CMF_LOG_COUNT; ++k)
            if (!f->log_sums[k]) return set_error(CMF_ERR_INVALID, "%s: log_sums[%d] is NULL", op, k);
        if ((rc = require({NEED(current_step)}, op))) return rc;
        if (f->num_trials > 1) return set_error(CMF_ERR_UNSUPPORTED, "%s: the log module does not support trials", op);
    }
    if ((f->flags & CMF_PER_TRIAL_DT) && !f->max_sub_steps)
        return set_error(CMF_ERR_INVALID, "%s: per-trial time steps need max_sub_steps", op);
    return CMF_OK;
}

// Kernel timing: while a Timing is installed, every launch is bracketed by
// a pair of HIP events on its stream (cmf_hip_substep_timed).
struct Timing {
    static constexpr int kMax = 32;
    hipEvent_t ev[2 * kMax];
    const char* name[kMax];
    int count = 0;
};
thread_local Timing* g_timing = nullptr;

// Launch `kernel`.  While a Timing is installed the launch goes through
// hipExtLaunchKernelGGL with a pair of the Timing's events, which the dispatch
// itself stamps when the kernel starts and ends: the duration is the
// kernel's own, without the dispatch latency before it.
#define CMF_DISPATCH(label, kernel, grid, block, smem, stream, ...)                           \
    do {                                                                                      \
        Timing* _t = g_timing;                                                                \
        if (_t && _t->count < Timing::kMax) {                                                 \
            _t->name[_t->count] = label;                                                      \
            hipExtLaunchKernelGGL(kernel, grid, block, smem, stream, _t->ev[2 * _t->count],   \
                                  _t->ev[2 * _t->count + 1], 0, __VA_ARGS__);                 \
            _t->count++;                                                                      \
        } else {                                                                              \
            hipLaunchKernelGGL(kernel, grid, block, smem, stream, __VA_ARGS__);               \
        }                                                                                     \
    } while (0)

#define LAUNCH(kernel, count, stream, ...)                                                   \
    do {                                                                                     \
        const int64_t _cnt = (count);                                                        \
        if (_cnt > 0) {                                                                      \
            CMF_DISPATCH(#kernel, kernel, dim3((unsigned)blocks_for(_cnt)), dim3(kBlock), 0,  \
                         (hipStream_t)(stream), __VA_ARGS__);                                \
            hipError_t _e = hipGetLastError();                                               \
            if (_e != hipSuccess)                                                            \
                return set_error(CMF_ERR_HIP, "launch of %s failed: %s", #kernel,            \
                                 hipGetErrorString(_e));                                     \
        }                                                                                    \
    } while (0)

constexpr uint32_t kLogAll = (1u << CMF_LOG_COUNT) - 1;
constexpr uint32_t kLogStage = (1u << CMF_LOG_STORAGE_STAGE) | (1u << CMF_LOG_RIVER_STORAGE) |
                               (1u << CMF_LOG_FLOOD_STORAGE) | (1u << CMF_LOG_FLOOD_AREA) |
                               (1u << CMF_LOG_STAGE_ERROR);

template <class REAL, class STO>
int log_final(const View<REAL, STO>& v, int64_t nblocks, uint32_t mask, void* stream) {
    CMF_DISPATCH("k_log_final", (k_log_final<REAL, STO>), dim3(1), dim3(kBlock), 0, (hipStream_t)stream, v,
                 nblocks, mask);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CMF_ERR_HIP, "launch of k_log_final failed: %s", hipGetErrorString(e));
    return CMF_OK;
}

// Launch with an explicit block size (a multiple of 64, at most kMaxBlock).
#define LAUNCH_B(kernel, count, block, stream, ...)                                         \
    do {                                                                                     \
        const int64_t _cnt = (count);                                                        \
        const int64_t _blk = (block);                                                        \
        if (_cnt > 0) {                                                                      \
            CMF_DISPATCH(#kernel, kernel, dim3((unsigned)((_cnt + _blk - 1) / _blk)), dim3((unsigned)_blk), 0, \
                         (hipStream_t)(stream), __VA_ARGS__);                                \
            hipError_t _e = hipGetLastError();                                               \
            if (_e != hipSuccess)                                                            \
                return set_error(CMF_ERR_HIP, "launch of %s failed: %s", #kernel,            \
                                 hipGetErrorString(_e));                                     \
        }                                                                                    \
    } while (0)

// Block size of a tunable kernel; catchment kernels of a grouped network run
// the group block table (kGroupRecord threads per entry).
inline int64_t block_of(const cmf_network* net, int id, bool catchment_kernel) {
    return (catchment_kernel && net->G > 0) ? kGroupRecord : net->blocks[id];
}

// Whether the contiguous CFL kernel applies: one trial, or trials in the
// inner layout whose count divides 64.
template <class REAL, class STO>
bool cfl_inner_ok(const cmf_network* net, const View<REAL, STO>& v) {
    const bool contiguous = (v.T == 1 && v.cs == 1) || (v.ts == 1 && v.cs == v.T);
    return contiguous && net->cfl_part && v.T == net->T && v.T <= 64 && 64 % v.T == 0;
}

// The CFL sub-step count(s) into max_sub_steps (cleared before): the
// contiguous kernel (cfl_inner_ok), else k_cfl (chunks of the catchments per
// trial).  With `tail` (contiguous kernel only) the fold also writes the
// counts whole, the time steps and the rest of CflTail: nothing is cleared
// before and nothing follows.
template <class REAL, class STO>
int launch_cfl(const cmf_network* net, const View<REAL, STO>& v, hipStream_t s, const CflTail<REAL>* tail = nullptr) {
    if (cfl_inner_ok(net, v)) {
        int log2T = 0;
        while ((int64_t(1) << log2T) < v.T) ++log2T;
        const int64_t nb = std::min<int64_t>(kCflInnerBlocks, (v.N * v.T + 8 * kBlock - 1) / (8 * kBlock));
        if (tail) {
            hipLaunchKernelGGL((k_cfl_inner<REAL, STO, true>), dim3((unsigned)nb), dim3(kBlock), 0, s, v, log2T,
                               net->cfl_part, tail->outer);
            hipLaunchKernelGGL((k_cfl_fold_tail<REAL>), dim3(1), dim3(kBlock), 0, s, (const int32_t*)net->cfl_part,
                               nb, log2T, *tail);
        } else {
            hipLaunchKernelGGL((k_cfl_inner<REAL, STO>), dim3((unsigned)nb), dim3(kBlock), 0, s, v, log2T,
                               net->cfl_part, 0.0);
            hipLaunchKernelGGL(k_cfl_fold, dim3((unsigned)v.T), dim3(kBlock), 0, s, (const int32_t*)net->cfl_part, nb,
                               v.T, v.max_sub_steps, v.has(CMF_PER_TRIAL_DT) ? 1 : 0);
        }
    } else if (tail) {
        return set_error(CMF_ERR_INTERNAL, "CFL pass: the one-launch form needs the contiguous kernel");
    } else {
        const int64_t chunks = std::min<int64_t>(blocks_for(v.N), 512);
        hipLaunchKernelGGL((k_cfl<REAL, STO>), dim3((unsigned)chunks, (unsigned)v.T), dim3(kBlock), 0, s, v);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CMF_ERR_HIP, "launch of the CFL pass failed: %s", hipGetErrorString(e));
    return CMF_OK;
}

// ----------------------------------------------------------- per-op -----
enum Op {
    OP_OUTFLOW, OP_INFLOW, OP_STAGE, OP_STAGE_LOG, OP_CFL, OP_BIF_OUT, OP_LEVEE_BIF_OUT,
    OP_BIF_IN, OP_LEVEE, OP_LEVEE_LOG, OP_RESERVOIR
};

template <class REAL, class STO>
int run_op(cmf_network* net, const cmf_fields* f, int op, void* stream) {
    View<REAL, STO> v;
    int rc = make_view<REAL, STO>(net, f, v);
    if (rc) return rc;
    if (net->G > 0)
        return set_error(CMF_ERR_UNSUPPORTED,
                         "per-op entry points take ungrouped networks; advance grouped networks with cmf_hip_advance");
    const int64_t NT = v.N * v.T;
    const bool contig = net->contiguous != 0;
    switch (op) {
        case OP_OUTFLOW:
            if ((rc = need_base(f, "compute_outflow")) || (rc = need_features(f, "compute_outflow"))) return rc;
            if (!f->outgoing_storage) return set_error(CMF_ERR_INVALID, "compute_outflow: outgoing_storage is NULL");
            LAUNCH((k_outflow<REAL, STO, false>), NT, stream, v, -1, 1);
            if (contig) LAUNCH((k_outgoing_sum<REAL, STO, true>), NT, stream, v);
            else LAUNCH((k_outgoing_sum<REAL, STO, false>), NT, stream, v);
            return CMF_OK;
        case OP_INFLOW:
            if ((rc = require({NEED(downstream_idx), NEED(river_outflow), NEED(flood_outflow),
                               NEED(river_storage), NEED(flood_storage), NEED(outgoing_storage),
                               NEED(river_inflow), NEED(flood_inflow)}, "compute_inflow")))
                return rc;
            if ((f->flags & CMF_HAS_BIFURCATION) && !f->limit_rate)
                return set_error(CMF_ERR_INVALID, "compute_inflow: limit_rate is NULL");
            if ((f->flags & CMF_HAS_RESERVOIR) && (!f->reservoir_total_inflow || !f->is_reservoir))
                return set_error(CMF_ERR_INVALID, "compute_inflow: reservoir fields are NULL");
            LAUNCH((k_inflow_limit<REAL, STO>), NT, stream, v);
            if (contig) LAUNCH((k_inflow_sum<REAL, STO, true>), NT, stream, v);
            else LAUNCH((k_inflow_sum<REAL, STO, false>), NT, stream, v);
            return CMF_OK;
        case OP_STAGE:
        case OP_STAGE_LOG: {
            const bool log = (op == OP_STAGE_LOG);
            const char* name = log ? "compute_flood_stage_log" : "compute_flood_stage";
            if ((rc = need_base(f, name))) return rc;
            if (!f->outgoing_storage) return set_error(CMF_ERR_INVALID, "%s: outgoing_storage is NULL", name);
            if ((f->flags & CMF_HAS_BIFURCATION) && !f->global_bifurcation_outflow)
                return set_error(CMF_ERR_INVALID, "%s: global_bifurcation_outflow is NULL", name);
            if ((f->flags & CMF_HAS_LEVEE) && log && !f->is_levee)
                return set_error(CMF_ERR_INVALID, "%s: is_levee is NULL", name);
            if (log) {
                cmf_fields g = *f;
                g.flags |= CMF_HAS_LOG;
                if ((rc = need_features(&g, name))) return rc;
                if (blocks_for(NT) > net->partial_blocks)
                    return set_error(CMF_ERR_INTERNAL, "%s: partial buffer too small", name);
                if (v.L == 10) LAUNCH((k_stage<REAL, STO, true, 10>), NT, stream, v);
                else LAUNCH((k_stage<REAL, STO, true, 0>), NT, stream, v);
                return log_final(v, blocks_for(NT), kLogAll, stream);
            }
            if (v.L == 10) LAUNCH((k_stage<REAL, STO, false, 10>), NT, stream, v);
            else LAUNCH((k_stage<REAL, STO, false, 0>), NT, stream, v);
            return CMF_OK;
        }
        case OP_CFL: {
            if ((rc = require({NEED(river_depth), NEED(downstream_distance), NEED(max_sub_steps),
                               NEED(outer_time_step)}, "compute_adaptive_time_step")))
                return rc;
            if ((f->flags & CMF_HAS_RESERVOIR) && !f->is_dam_related)
                return set_error(CMF_ERR_INVALID, "compute_adaptive_time_step: is_dam_related is NULL");
            if (v.N == 0) return CMF_OK;
            return launch_cfl(net, v, (hipStream_t)stream);
        }
        case OP_BIF_OUT:
        case OP_LEVEE_BIF_OUT: {
            const char* name = op == OP_BIF_OUT ? "compute_bifurcation_outflow" : "compute_levee_bifurcation_outflow";
            if (!(f->flags & CMF_HAS_BIFURCATION)) return set_error(CMF_ERR_INVALID, "%s: bifurcation is off", name);
            if ((rc = need_features(f, name))) return rc;
            if ((rc = require({NEED(outgoing_storage), NEED(time_step)}, name))) return rc;
            if (op == OP_LEVEE_BIF_OUT) {
                if ((rc = require({NEED(protected_water_surface_elevation)}, name))) return rc;
                if (v.LB == 5) LAUNCH((k_bif_outflow<REAL, STO, true, 5>), v.P * v.T, stream, v, -1);
                else LAUNCH((k_bif_outflow<REAL, STO, true, 0>), v.P * v.T, stream, v, -1);
            } else {
                if (v.LB == 5) LAUNCH((k_bif_outflow<REAL, STO, false, 5>), v.P * v.T, stream, v, -1);
                else LAUNCH((k_bif_outflow<REAL, STO, false, 0>), v.P * v.T, stream, v, -1);
            }
            LAUNCH((k_bif_outgoing<REAL, STO>), NT, stream, v);
            return CMF_OK;
        }
        case OP_BIF_IN:
            if (!(f->flags & CMF_HAS_BIFURCATION))
                return set_error(CMF_ERR_INVALID, "compute_bifurcation_inflow: bifurcation is off");
            if ((rc = need_features(f, "compute_bifurcation_inflow"))) return rc;
            LAUNCH((k_bif_inflow<REAL, STO>), v.P * v.T, stream, v, -1);
            LAUNCH((k_gbif_sum<REAL, STO>), NT, stream, v);
            return CMF_OK;
        case OP_LEVEE:
        case OP_LEVEE_LOG: {
            const bool log = (op == OP_LEVEE_LOG);
            const char* name = log ? "compute_levee_stage_log" : "compute_levee_stage";
            if (!(f->flags & CMF_HAS_LEVEE)) return set_error(CMF_ERR_INVALID, "%s: levee is off", name);
            if ((rc = need_base(f, name)) || (rc = need_features(f, name))) return rc;
            if (log) {
                cmf_fields g = *f;
                g.flags |= CMF_HAS_LOG;
                if ((rc = need_features(&g, name))) return rc;
                if (v.L == 10) LAUNCH((k_levee<REAL, STO, true, 10>), v.V * v.T, stream, v, -1);
                else LAUNCH((k_levee<REAL, STO, true, 0>), v.V * v.T, stream, v, -1);
                if (v.V * v.T > 0) return log_final(v, blocks_for(v.V * v.T), kLogStage, stream);
                return CMF_OK;
            }
            if (v.L == 10) LAUNCH((k_levee<REAL, STO, false, 10>), v.V * v.T, stream, v, -1);
            else LAUNCH((k_levee<REAL, STO, false, 0>), v.V * v.T, stream, v, -1);
            return CMF_OK;
        }
        case OP_RESERVOIR:
            if (!(f->flags & CMF_HAS_RESERVOIR))
                return set_error(CMF_ERR_INVALID, "compute_reservoir_outflow: reservoir is off");
            if ((rc = need_base(f, "compute_reservoir_outflow")) ||
                (rc = need_features(f, "compute_reservoir_outflow")))
                return rc;
            if (!f->outgoing_storage) return set_error(CMF_ERR_INVALID, "compute_reservoir_outflow: outgoing_storage is NULL");
            LAUNCH((k_reservoir<REAL, STO, false>), v.R * v.T, stream, v, -1);
            if (contig) LAUNCH((k_res_regather<REAL, STO, true>), v.R * v.T, stream, v);
            else LAUNCH((k_res_regather<REAL, STO, false>), v.R * v.T, stream, v);
            return CMF_OK;
    }
    return set_error(CMF_ERR_INVALID, "unknown op");
}

// ------------------------------------------- allocation-free variant ----
// cmf_hip_flood_stage_log_noalloc keeps its FP64 per-block log partials in
// this file-scope device array instead of a network's scratch: room for the
// partials of kNoAllocMaxElements catchment elements (kBlock per block, 11
// sums each), 5.8 MB held on every GPU the library runs on until the process
// ends.  Host code never passes `scratch` itself to a kernel: on the host the
// name stands for a host-side symbol, so each device's own address is taken
// with hipGetSymbolAddress, once per device, and kept in g_scratch_dev.
constexpr int64_t kNoAllocMaxElements = int64_t(1) << 24;
constexpr int64_t kNoAllocBlocks = kNoAllocMaxElements / kBlock;
static __device__ __align__(256) double scratch[kNoAllocBlocks * CMF_LOG_COUNT];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The current device's copy of `scratch` (its address looked up on the first
// call on that device).
int scratch_of_current_device(double** out) {
    int dev = -1;
    CMF_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(CMF_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(CMF_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return CMF_OK;
}

// compute_flood_stage_log without a network context and without any
// allocation: k_stage with its partials in `scratch`, then k_log_final.
template <class REAL, class STO>
int run_stage_log_noalloc(const cmf_fields* f, void* stream) {
    const char* name = "cmf_hip_flood_stage_log_noalloc";
    View<REAL, STO> v;
    int rc = view_of_fields<REAL, STO>(f, v);
    if (rc) return rc;
    if ((rc = need_base(f, name))) return rc;
    if (!f->outgoing_storage) return set_error(CMF_ERR_INVALID, "%s: outgoing_storage is NULL", name);
    if ((f->flags & CMF_HAS_BIFURCATION) && !f->global_bifurcation_outflow)
        return set_error(CMF_ERR_INVALID, "%s: global_bifurcation_outflow is NULL", name);
    if ((f->flags & CMF_HAS_LEVEE) && !f->is_levee) return set_error(CMF_ERR_INVALID, "%s: is_levee is NULL", name);
    cmf_fields g = *f;
    g.flags |= CMF_HAS_LOG;
    if ((rc = need_features(&g, name))) return rc;
    const int64_t NT = v.N * v.T;
    if (NT > kNoAllocMaxElements)
        return set_error(CMF_ERR_UNSUPPORTED, "%s: %lld catchment elements exceed the %lld its static scratch holds",
                         name, (long long)NT, (long long)kNoAllocMaxElements);
    if ((rc = scratch_of_current_device(&v.partials))) return rc;
    if (v.L == 10) LAUNCH((k_stage<REAL, STO, true, 10>), NT, stream, v);
    else LAUNCH((k_stage<REAL, STO, true, 0>), NT, stream, v);
    if (NT == 0) return CMF_OK;
    return log_final(v, blocks_for(NT), kLogAll, stream);
}

int dispatch_op(cmf_network* net, const cmf_fields* f, int op, void* stream) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64) return run_op<float, double>(net, f, op, stream);
    if (f->precision == CMF_PRECISION_F64_F64) return run_op<double, double>(net, f, op, stream);
    if (f->precision == CMF_PRECISION_F32_F32) return run_op<float, float>(net, f, op, stream);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

// ------------------------------------------------------------- fused ----
// The stage pass k_stage_fused, specialised on the table level count and the
// fused levee stage.  `nt`: threads of a catchment kernel (N*T, or the group
// block table).
template <class REAL, class STO, bool CONTIG, bool LOG, class VT>
int launch_stage(const cmf_network* net, const VT& v, int64_t nt, bool lev, int substep,
                  int store_inflow, void* stream) {
    const int64_t b = block_of(net, CMF_TUNE_STAGE, true);
    if (v.L == 10) {
        if (lev) LAUNCH_B((k_stage_fused<REAL, STO, CONTIG, LOG, 10, true, VT>), nt, b, stream, v, substep, store_inflow);
        else LAUNCH_B((k_stage_fused<REAL, STO, CONTIG, LOG, 10, false, VT>), nt, b, stream, v, substep, store_inflow);
    } else {
        if (lev) LAUNCH_B((k_stage_fused<REAL, STO, CONTIG, LOG, 0, true, VT>), nt, b, stream, v, substep, store_inflow);
        else LAUNCH_B((k_stage_fused<REAL, STO, CONTIG, LOG, 0, false, VT>), nt, b, stream, v, substep, store_inflow);
    }
    return CMF_OK;
}

// The block-local schedule of one sub-step: k_fused_local, the limiter list,
// the stage list (see k_fused_local).
template <class REAL, class STO, bool LOG, bool LEV, bool PATHS, bool PLEVEE>
int local_substep_t(cmf_network* net, const View<REAL, STO>& v, int substep, int store_inflow, void* stream) {
    const int64_t B = net->local_block;
    const int64_t pb = PATHS ? (v.P + B - 1) / B : 0;
    const int64_t nf = (v.N + B - 1) / B;
    LAUNCH_B((k_fused_local<REAL, STO, true, LOG, 10, LEV, PATHS, PLEVEE, 5>), (pb + nf) * B, B, stream, v, substep,
             pb, store_inflow, (const uint8_t*)net->lflags);
    if (net->lim_count > 0)
        LAUNCH_B((k_limiter_list<REAL, STO, true>), net->lim_count, block_of(net, CMF_TUNE_LIMITER, true), stream, v,
                 store_inflow, (const int32_t*)net->lim_list, net->lim_count);
    const int64_t bs = block_of(net, CMF_TUNE_STAGE, true);
    if (net->stage_count > 0)
        LAUNCH_B((k_stage_list<REAL, STO, true, LOG, 10, LEV>), net->stage_count, bs, stream, v, substep,
                 store_inflow, (const int32_t*)net->stage_list, net->stage_count, nf);
    if (LOG) {
        const int64_t nblocks = nf + (net->stage_count + bs - 1) / bs;
        if (nblocks > net->partial_blocks) return set_error(CMF_ERR_INTERNAL, "local sub-step: partial buffer too small");
        const int rc = log_final(v, nblocks, kLogAll, stream);
        if (rc) return rc;
    }
    return CMF_OK;
}

template <class REAL, class STO>
int local_substep(cmf_network* net, const View<REAL, STO>& v, int substep, int store_inflow, bool lev_fused,
                  void* stream) {
    const bool log = v.has(CMF_HAS_LOG), paths = v.has(CMF_HAS_BIFURCATION), levee = v.has(CMF_HAS_LEVEE);
#define CMF_LOCAL(LOGV, LEVV, PV, PLV) local_substep_t<REAL, STO, LOGV, LEVV, PV, PLV>(net, v, substep, store_inflow, stream)
#define CMF_LOCAL_P(LOGV)                                                                 \
    (!paths ? (lev_fused ? CMF_LOCAL(LOGV, true, false, false) : CMF_LOCAL(LOGV, false, false, false)) \
            : (lev_fused ? CMF_LOCAL(LOGV, true, true, true)                              \
                         : (levee ? CMF_LOCAL(LOGV, false, true, true) : CMF_LOCAL(LOGV, false, true, false))))
    return log ? CMF_LOCAL_P(true) : CMF_LOCAL_P(false);
#undef CMF_LOCAL_P
#undef CMF_LOCAL
}

// The outflow launch of a fused sub-step: the catchment outflow pass, with
// bifurcation paths in leading blocks of its own.  PROV: provisional
// limiters instead of outgoing terms (k_limiter_fix).
template <class REAL, class STO, class VT, bool PROV>
int launch_outflow(const cmf_network* net, const VT& v, int64_t NT, int substep, int store_inflow, void* stream) {
    const int64_t b_out = block_of(net, CMF_TUNE_OUTFLOW, true);
    if (v.has(CMF_HAS_BIFURCATION)) {
        // the path pass rides in the outflow launch, in blocks of its own first
        const int64_t pb = (v.P * v.T + b_out - 1) / b_out;
        const int64_t cnt = (pb + (NT + b_out - 1) / b_out) * b_out;
        const bool levee = v.has(CMF_HAS_LEVEE), l5 = v.LB == 5;
        if constexpr (PROV) {
            if (levee && l5) LAUNCH_B((k_outflow_paths_prov<REAL, STO, true, 5, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else if (levee) LAUNCH_B((k_outflow_paths_prov<REAL, STO, true, 0, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else if (l5) LAUNCH_B((k_outflow_paths_prov<REAL, STO, false, 5, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else LAUNCH_B((k_outflow_paths_prov<REAL, STO, false, 0, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
        } else {
            if (levee && l5) LAUNCH_B((k_outflow_paths<REAL, STO, true, 5, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else if (levee) LAUNCH_B((k_outflow_paths<REAL, STO, true, 0, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else if (l5) LAUNCH_B((k_outflow_paths<REAL, STO, false, 5, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
            else LAUNCH_B((k_outflow_paths<REAL, STO, false, 0, VT>), cnt, b_out, stream, v, substep, pb, store_inflow);
        }
    } else if constexpr (PROV) {
        LAUNCH_B((k_outflow_prov<REAL, STO, VT>), NT, b_out, stream, v, substep, store_inflow);
    } else {
        LAUNCH_B((k_outflow<REAL, STO, true, VT>), NT, b_out, stream, v, substep, store_inflow);
    }
    return CMF_OK;
}

// The three passes of one fused sub-step, on view type VT (PlainView for one
// plain trial, where the per-trial and per-group indexing folds away).
// THREE_PASS: outflow with provisional limiters, [reservoir release], the
// limiter fix-up (path ends, reservoir catchments, flagged catchments), stage
// (taped: the unlimited flows and limiters land in the tape, the same values
// as from the limiter pass); LIMITER_PASS: outflow, [reservoir], limiter,
// stage.
template <class REAL, class STO, class VT>
int three_pass_substep(cmf_network* net, const VT& v_in, int substep, int store_inflow, void* stream) {
    const bool prov = net->schedule == CMF_SCHEDULE_THREE_PASS && net->bmask;
    VT v = v_in;
    if (prov) v.eterm = static_cast<REAL*>(net->eterm);
    // catchment kernels: N*T threads, or one block per group table entry
    const bool ordered = net->G > 0 && v.ordered && substep >= 0 && substep < (int)net->sub_blocks.size();
    const int64_t NT = net->G > 0 ? (ordered ? net->sub_blocks[substep] : net->nblk) * kGroupRecord : v.N * v.T;
    const bool contig = net->contiguous != 0;
    const bool bif = v.has(CMF_HAS_BIFURCATION), levee = v.has(CMF_HAS_LEVEE);
    const bool log = v.has(CMF_HAS_LOG);
    const bool lev_fused = levee && v.levee_index != nullptr;
    const int64_t b_lim = block_of(net, CMF_TUNE_LIMITER, true), b_stage = block_of(net, CMF_TUNE_STAGE, true);
    int rc = prov ? launch_outflow<REAL, STO, VT, true>(net, v, NT, substep, store_inflow, stream)
                  : launch_outflow<REAL, STO, VT, false>(net, v, NT, substep, store_inflow, stream);
    if (rc) return rc;
    if (v.has(CMF_HAS_RESERVOIR))
        LAUNCH_B((k_reservoir<REAL, STO, true>), v.R * v.T, block_of(net, CMF_TUNE_RESERVOIR, false), stream, v, substep);
    // limiters of every catchment, then the stage pass (which also forms the
    // path nets from the limiters of both path ends)
    if (prov) {
        // path ends (per trial) in leading blocks, reservoir catchments
        // next, then the flag words: the whole network, or one wave per
        // group record of the sub-step
        const int64_t list_n = bif ? net->pend_count : 0;
        const int64_t nres = v.has(CMF_HAS_RESERVOIR) ? v.R * v.T : 0;
        const int4* list = net->pend_list;
        if (net->G > 0) {
            const View<REAL, STO> gv = v;
            const int64_t lb = (list_n * v.T + 63) / 64, rb = (nres + 63) / 64, nrec = NT / kGroupRecord;
            if (contig) LAUNCH_B((k_limiter_fix<REAL, STO, true, View<REAL, STO>>), (lb + rb + nrec) * 64, 64, stream,
                                 gv, substep, store_inflow, list, list_n, lb, rb, (const int4*)gv.blk_rec);
            else LAUNCH_B((k_limiter_fix<REAL, STO, false, View<REAL, STO>>), (lb + rb + nrec) * 64, 64, stream,
                          gv, substep, store_inflow, list, list_n, lb, rb, (const int4*)gv.blk_rec);
        } else {
            const int64_t lb = (list_n * v.T + b_lim - 1) / b_lim, rb = (nres + b_lim - 1) / b_lim;
            const int64_t words = (v.N * v.T + 31) / 32;
            const int64_t cnt = (lb + rb + (words + b_lim - 1) / b_lim) * b_lim;
            if (contig) LAUNCH_B((k_limiter_fix<REAL, STO, true, VT>), cnt, b_lim, stream, v, substep, store_inflow, list,
                                 list_n, lb, rb, (const int4*)nullptr);
            else LAUNCH_B((k_limiter_fix<REAL, STO, false, VT>), cnt, b_lim, stream, v, substep, store_inflow, list,
                          list_n, lb, rb, (const int4*)nullptr);
        }
    } else if constexpr (VT::kGrouped) {
        // one block of kGroupRecord / 2 threads per group table entry
        if (contig)
            LAUNCH_B((k_limiter_pair<REAL, STO, true, VT>), NT / 2, kGroupRecord / 2, stream, v, substep, store_inflow);
        else
            LAUNCH_B((k_limiter_pair<REAL, STO, false, VT>), NT / 2, kGroupRecord / 2, stream, v, substep, store_inflow);
    } else if (net->G > 0) {
        // the group block table maps one catchment per thread
        if (contig) LAUNCH_B((k_limiter<REAL, STO, true>), NT, b_lim, stream, v, substep, store_inflow);
        else LAUNCH_B((k_limiter<REAL, STO, false>), NT, b_lim, stream, v, substep, store_inflow);
    } else {
        const int64_t threads = ((NT + 2 * b_lim - 1) / (2 * b_lim)) * b_lim;
        if (contig) LAUNCH_B((k_limiter_pair<REAL, STO, true, VT>), threads, b_lim, stream, v, substep, store_inflow);
        else LAUNCH_B((k_limiter_pair<REAL, STO, false, VT>), threads, b_lim, stream, v, substep, store_inflow);
    }
    rc = log ? (contig ? launch_stage<REAL, STO, true, true, VT>(net, v, NT, lev_fused, substep, store_inflow, stream)
                           : launch_stage<REAL, STO, false, true, VT>(net, v, NT, lev_fused, substep, store_inflow, stream))
                 : (contig ? launch_stage<REAL, STO, true, false, VT>(net, v, NT, lev_fused, substep, store_inflow, stream)
                           : launch_stage<REAL, STO, false, false, VT>(net, v, NT, lev_fused, substep, store_inflow, stream));
    if (rc) return rc;
    if (log) {
        if (net->G > 0) {
            CMF_DISPATCH("k_log_final_grouped", (k_log_final_grouped<REAL, STO>), dim3((unsigned)net->G),
                         dim3(kBlock), 0, (hipStream_t)stream, v, net->gblk_ptr, kLogAll, substep);
            CMF_HIP_CHECK(hipGetLastError());
        } else {
            // partials: one per block of the stage pass
            const int rf = log_final(v, (NT + b_stage - 1) / b_stage, kLogAll, stream);
            if (rf) return rf;
        }
    }
    if (levee && !lev_fused) {
        if (log) {
            if (v.L == 10) LAUNCH((k_levee<REAL, STO, true, 10>), v.V * v.T, stream, v, substep);
            else LAUNCH((k_levee<REAL, STO, true, 0>), v.V * v.T, stream, v, substep);
            if (v.V * v.T > 0) {
                rc = log_final(v, blocks_for(v.V * v.T), kLogStage, stream);
                if (rc) return rc;
            }
        } else {
            if (v.L == 10) LAUNCH((k_levee<REAL, STO, false, 10>), v.V * v.T, stream, v, substep);
            else LAUNCH((k_levee<REAL, STO, false, 0>), v.V * v.T, stream, v, substep);
        }
    }
    return CMF_OK;
}

// The group block table of sub-step `substep` of an ordered (adaptive,
// eager) grouped advance: the records of the groups still sub-stepping, in
// memory order; else the view's own table.
template <class REAL, class STO>
const int4* level_table(const cmf_network* net, const View<REAL, STO>& v, int substep) {
    if (net->G > 0 && v.ordered && substep >= 0 && substep < (int)net->sub_blocks.size())
        return net->level_rec + net->level_off[(size_t)substep];
    return v.blk_rec;
}

// One fused sub-step: outflow [+ bifurcation paths], [reservoir], limiters,
// stage (inflow sums, limited flows, bifurcation nets, storage, flood stage
// [, levee stage]).
//
// dev_last: a plain run's first sub-step, enqueued before the host has read
// the outer step's count (early_advance): whether it is also the last is
// read on the device.
template <class REAL, class STO>
int fused_substep(cmf_network* net, const View<REAL, STO>& v, const cmf_fields* f, int substep,
                  int store_inflow, void* stream, bool dev_last = false) {
    if (dev_last) {
        PlainView<REAL, STO, true> pv;
        static_cast<View<REAL, STO>&>(pv) = v;
        return three_pass_substep<REAL, STO, PlainView<REAL, STO, true>>(net, pv, substep, 0, stream);
    }
    if (net->local_block > 0 && net->schedule == CMF_SCHEDULE_BLOCK_LOCAL && net->G == 0 && v.T == 1 && v.R == 0 && v.L == 10 &&
        (v.P == 0 || v.LB == 5) && !v.tape_q_r && !v.has(CMF_PER_TRIAL_DT) && net->contiguous) {
        const bool levee = v.has(CMF_HAS_LEVEE);
        const bool lev_fused = levee && v.levee_index != nullptr;
        int rc = local_substep<REAL, STO>(net, v, substep, store_inflow, lev_fused, stream);
        if (rc) return rc;
        if (levee && !lev_fused) {
            const bool log = v.has(CMF_HAS_LOG);
            if (log) {
                LAUNCH((k_levee<REAL, STO, true, 10>), v.V * v.T, stream, v, substep);
                if (v.V * v.T > 0) {
                    rc = log_final(v, blocks_for(v.V * v.T), kLogStage, stream);
                    if (rc) return rc;
                }
            } else {
                LAUNCH((k_levee<REAL, STO, false, 10>), v.V * v.T, stream, v, substep);
            }
        }
        return CMF_OK;
    }
    const bool plain = net->G == 0 && v.T == 1 && v.batched == 0 && !v.has(CMF_PER_TRIAL_DT);
    (void)f;
    if (plain) {
        PlainView<REAL, STO> pv;
        static_cast<View<REAL, STO>&>(pv) = v;
        return three_pass_substep<REAL, STO, PlainView<REAL, STO>>(net, pv, substep, store_inflow, stream);
    }
    // members in the trial-inner layout, a power of two (one catchment per
    // wave from 64 on)
    if (net->G == 0 && v.T >= 2 && (v.T & (v.T - 1)) == 0 && v.ts == 1 && v.cs == v.T &&
        (v.P == 0 || (v.pts == 1 && v.pps == v.T)) && !v.tape_q_r) {
        int log2T = 0;
        while ((int64_t(1) << log2T) < v.T) ++log2T;
        if (v.T >= 64) {
            MemberView<REAL, STO, true> mv;
            static_cast<View<REAL, STO>&>(mv) = v;
            mv.log2T = log2T;
            return three_pass_substep<REAL, STO, MemberView<REAL, STO, true>>(net, mv, substep, store_inflow, stream);
        }
        MemberView<REAL, STO, false> mv;
        static_cast<View<REAL, STO>&>(mv) = v;
        mv.log2T = log2T;
        return three_pass_substep<REAL, STO, MemberView<REAL, STO, false>>(net, mv, substep, store_inflow, stream);
    }
    // grouped, without the path pass (whose blocks are no group records)
    if (net->G > 0 && v.T == 1 && v.batched == 0 && !v.has(CMF_PER_TRIAL_DT) && !v.has(CMF_HAS_BIFURCATION) &&
        !v.tape_q_r) {
        GroupView<REAL, STO> gv;
        static_cast<View<REAL, STO>&>(gv) = v;
        gv.blk_rec = level_table(net, v, substep);
        return three_pass_substep<REAL, STO, GroupView<REAL, STO>>(net, gv, substep, store_inflow, stream);
    }
    if (net->G > 0) {
        View<REAL, STO> w = v;
        w.blk_rec = level_table(net, v, substep);
        return three_pass_substep<REAL, STO, View<REAL, STO>>(net, w, substep, store_inflow, stream);
    }
    return three_pass_substep<REAL, STO, View<REAL, STO>>(net, v, substep, store_inflow, stream);
}

// Grouped networks: one trial; with the log module the levee stage must run
// inside the stage pass (levee_index), whose blocks follow the group table.
int need_grouped(const cmf_network* net, const cmf_fields* f, const char* name) {
    if (net->G == 0) return CMF_OK;
    if (f->num_trials != 1) return set_error(CMF_ERR_UNSUPPORTED, "%s: grouped networks take one trial", name);
    if ((f->flags & CMF_HAS_LOG) && (f->flags & CMF_HAS_LEVEE) && !f->levee_index)
        return set_error(CMF_ERR_UNSUPPORTED, "%s: grouped networks with log and levee need levee_index", name);
    if (!f->max_sub_steps || !f->time_step)
        return set_error(CMF_ERR_INVALID, "%s: grouped networks need max_sub_steps and time_step (one per group)", name);
    return CMF_OK;
}

template <class REAL, class STO>
int run_substep(cmf_network* net, const cmf_fields* f, int substep, int store_inflow, void* stream) {
    View<REAL, STO> v;
    int rc = make_view<REAL, STO>(net, f, v);
    if (rc) return rc;
    if ((rc = need_base(f, "substep")) || (rc = need_features(f, "substep")) || (rc = need_grouped(net, f, "substep")))
        return rc;
    if (v.has(CMF_HAS_LOG) && (net->G > 0 ? net->nblk : (v.N * v.T + 63) / 64) > net->partial_blocks)
        return set_error(CMF_ERR_INTERNAL, "substep: partial buffer too small");
    return fused_substep(net, v, f, substep, store_inflow, stream);
}

template <class REAL>
__global__ void k_prepare(REAL* outer_time_step, double outer) {
    if (threadIdx.x == 0 && blockIdx.x == 0) outer_time_step[0] = (REAL)outer;
}
// the same and the n sub-step counts cleared, in one launch
template <class REAL>
__global__ void k_prepare_counts(REAL* outer_time_step, double outer, int32_t* counts, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) outer_time_step[0] = (REAL)outer;
    if (t < n) counts[t] = 0;
}
template <class REAL>
__global__ void k_set_dt_fixed(REAL* time_step, int64_t n, double value) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) time_step[t] = (REAL)value;
}
__global__ void k_fill_int(int32_t* p, int64_t n, int32_t value) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = value;
}

// The sub-step loop of one outer step: `count` fused sub-steps, the inflow
// stored on the last, current_step left at count - 1.
template <class REAL, class STO>
int substep_loop(cmf_network* net, const View<REAL, STO>& v, const cmf_fields* f, int count, hipStream_t s) {
    for (int k = 0; k < count; ++k) {
        if (v.has(CMF_HAS_LOG)) {
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, f->current_step, (int64_t)1, (int32_t)k);
        }
        // with per-trial counts each trial stores its inflow on its own last
        // sub-step (k_stage checks); a shared count stores on the last one
        int rc = fused_substep(net, v, f, k, (k == count - 1) ? 1 : 0, s);
        if (rc) return rc;
    }
    if (f->current_step)
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, f->current_step, (int64_t)1, (int32_t)(count - 1));
    CMF_HIP_CHECK(hipGetLastError());
    return CMF_OK;
}

// The same loop replayed from a graph captured once per sub-step count (on a
// private stream) and launched on `s`.  A changed field table drops the cache.
template <class REAL, class STO>
int substep_loop_graph(cmf_network* net, const View<REAL, STO>& v, const cmf_fields* f, int count, hipStream_t s) {
    if (!net->graph_fields_set || std::memcmp(&net->graph_fields, f, sizeof(cmf_fields)) != 0) {
        for (auto& g : net->graphs) (void)hipGraphExecDestroy(g.second);
        net->graphs.clear();
        std::memcpy(&net->graph_fields, f, sizeof(cmf_fields));
        net->graph_fields_set = true;
    }
    hipGraphExec_t exec = nullptr;
    for (auto& g : net->graphs)
        if (g.first == count) exec = g.second;
    if (!exec) {
        if (!net->capture_stream) CMF_HIP_CHECK(hipStreamCreateWithFlags(&net->capture_stream, hipStreamNonBlocking));
        CMF_HIP_CHECK(hipStreamBeginCapture(net->capture_stream, hipStreamCaptureModeThreadLocal));
        int rc = substep_loop(net, v, f, count, net->capture_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(net->capture_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return set_error(CMF_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_error(CMF_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        if (net->graphs.size() >= 64) {
            (void)hipGraphExecDestroy(net->graphs.front().second);
            net->graphs.erase(net->graphs.begin());
        }
        net->graphs.emplace_back(count, exec);
    }
    CMF_HIP_CHECK(hipGraphLaunch(exec, s));
    return CMF_OK;
}

// The per-sub-step group block tables of an adaptive grouped advance whose
// group counts are in host_counts and, on the device, `counts` (count: their
// maximum).  Level k lists the
// records of the groups with count > k in memory order, so a sub-step
// launches only those groups' blocks and walks memory as an ungrouped launch
// does.  The host places each group in each level (count x G positions); a
// device pass copies the records.  Growing a table waits for the device
// first (a launch on another stream may still read the old one); the staging
// buffer is rewritten only after the previous copy from it has completed.
int build_level_tables(cmf_network* net, const int32_t* counts, int count, hipStream_t s) {
    const int64_t G = net->G;
    const int32_t* hc = net->host_counts;
    net->sub_blocks.assign((size_t)count, 0);
    net->level_off.assign((size_t)count, 0);
    for (int64_t g = 0; g < G; ++g) {
        const int64_t nb = net->gblk_host[g + 1] - net->gblk_host[g];
        for (int k = 0; k < hc[g] && k < count; ++k) net->sub_blocks[(size_t)k] += nb;
    }
    int64_t total = 0;
    for (int k = 0; k < count; ++k) {
        net->level_off[(size_t)k] = total;
        total += net->sub_blocks[(size_t)k];
    }
    const int64_t npos = (int64_t)count * G;
    if (total > net->level_cap) {
        // the old tables may still be read by launches on another stream
        if (net->level_rec) {
            CMF_HIP_CHECK(hipDeviceSynchronize());
            CMF_HIP_CHECK(hipFree(net->level_rec));
        }
        net->level_rec = nullptr;
        net->level_cap = 0;
        const int64_t cap = std::max<int64_t>(total, net->nblk * 4);
        int rc = dev_alloc((void**)&net->level_rec, (size_t)cap * sizeof(int4));
        if (rc) return rc;
        net->level_cap = cap;
    }
    CMF_HIP_CHECK(hipEventSynchronize(net->order_copied));
    if (npos > net->level_pos_cap) {
        if (net->level_pos) {
            CMF_HIP_CHECK(hipDeviceSynchronize());
            CMF_HIP_CHECK(hipFree(net->level_pos));
        }
        if (net->host_level_pos) CMF_HIP_CHECK(hipHostFree(net->host_level_pos));
        net->level_pos = nullptr;
        net->host_level_pos = nullptr;
        net->level_pos_cap = 0;
        const int64_t cap = std::max<int64_t>(npos, G * 32);
        int rc = dev_alloc((void**)&net->level_pos, (size_t)cap * sizeof(int32_t));
        if (rc) return rc;
        hipError_t e = hipHostMalloc((void**)&net->host_level_pos, (size_t)cap * sizeof(int32_t), hipHostMallocDefault);
        if (e != hipSuccess) {
            net->host_level_pos = nullptr;
            return set_error(CMF_ERR_NOMEM, "hipHostMalloc failed: %s", hipGetErrorString(e));
        }
        net->level_pos_cap = cap;
    }
    std::vector<int64_t> run(net->level_off);
    for (int64_t g = 0; g < G; ++g) {
        const int64_t nb = net->gblk_host[g + 1] - net->gblk_host[g];
        for (int k = 0; k < hc[g] && k < count; ++k) {
            net->host_level_pos[(int64_t)k * G + g] = (int32_t)run[(size_t)k];
            run[(size_t)k] += nb;
        }
    }
    CMF_HIP_CHECK(hipMemcpyAsync(net->level_pos, net->host_level_pos, (size_t)npos * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
    CMF_HIP_CHECK(hipEventRecord(net->order_copied, s));
    hipLaunchKernelGGL(k_level_tables, dim3((unsigned)blocks_for(net->nblk)), dim3(kBlock), 0, s,
                       (const int4*)net->blk_rec_id, net->nblk, (const int32_t*)net->gblk_ptr, counts,
                       (const int32_t*)net->level_pos, G, count, (int4*)net->level_rec);
    CMF_HIP_CHECK(hipGetLastError());
    return CMF_OK;
}

// A plain adaptive advance need not wait for its sub-step count before
// launching its first sub-step: one trial, ungrouped, shared parameters, the
// three-pass schedule, and no pass that runs on the generic view (reservoirs,
// the separate levee pass, the log's per-sub-step partial folding).
template <class REAL, class STO>
bool early_ok(const cmf_network* net, const View<REAL, STO>& v) {
    const bool plain = net->G == 0 && v.T == 1 && v.batched == 0 && !v.has(CMF_PER_TRIAL_DT);
    const bool lev_separate = v.has(CMF_HAS_LEVEE) && !v.levee_index;
    const bool local = net->local_block > 0 && net->schedule == CMF_SCHEDULE_BLOCK_LOCAL;
    return plain && !local && v.R == 0 && !lev_separate && !v.has(CMF_HAS_LOG) && !v.tape_q_r;
}

// Waits for the n sub-step counts that k_cfl_fold_tail stores in pinned host
// memory (-1 until then): a busy read for about a millisecond, then reads
// between stream queries, so that the wait ends with the stream even if a
// store were never seen.  No event is recorded on the stream (its marker
// costs the GPU several microseconds of idle time per outer step).
int wait_counts(cmf_network* net, int64_t n, hipStream_t s) {
    const volatile int32_t* hc = net->host_counts;
    auto ready = [&]() {
        for (int64_t t = 0; t < n; ++t)
            if (hc[t] < 0) return false;
        return true;
    };
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t k = 0;; ++k) {
        if (ready()) break;
        if ((k & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(1)) {
            const hipError_t q = hipStreamQuery(s);
            if (q == hipSuccess) {
                if (ready()) break;
                return set_error(CMF_ERR_INTERNAL, "advance: the sub-step counts did not reach the host");
            }
            if (q != hipErrorNotReady) return set_error(CMF_ERR_HIP, "advance: %s", hipGetErrorString(q));
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return CMF_OK;
}

// The sub-step loop of a plain adaptive advance whose count is on its way to
// the host (stored by the CFL pass or copied, enqueued on `s`): the first
// sub-step, which every outer step has, is enqueued at once (whether it is
// also the last is read on the device, where its stores happen), then the
// host waits for the count, which arrives while the GPU runs that sub-step,
// and enqueues the rest.  The GPU does not idle on the count's round trip.
// current_step was set to count - 1 with the time step.
template <class REAL, class STO>
int early_advance(cmf_network* net, const View<REAL, STO>& v, const cmf_fields* f, int32_t* out_count,
                  hipStream_t s, bool stored) {
    if (!stored) {
        if (!net->count_ready) CMF_HIP_CHECK(hipEventCreateWithFlags(&net->count_ready, hipEventDisableTiming));
        CMF_HIP_CHECK(hipEventRecord(net->count_ready, s));
    }
    int rc = fused_substep(net, v, f, 0, 0, s, true);
    if (rc) return rc;
    if (stored) {
        if ((rc = wait_counts(net, 1, s))) return rc;
    } else {
        CMF_HIP_CHECK(hipEventSynchronize(net->count_ready));
    }
    const int count = std::max(1, (int)net->host_counts[0]);
    for (int k = 1; k < count; ++k)
        if ((rc = fused_substep(net, v, f, k, (k == count - 1) ? 1 : 0, s))) return rc;
    if (out_count) *out_count = count;
    return CMF_OK;
}

template <class REAL, class STO>
int run_advance(cmf_network* net, const cmf_fields* f, double outer, int fixed_count,
                int32_t* out_count, void* stream, bool use_graph) {
    View<REAL, STO> v;
    int rc = make_view<REAL, STO>(net, f, v);
    if (rc) return rc;
    if ((rc = need_base(f, "advance")) || (rc = need_features(f, "advance")) || (rc = need_grouped(net, f, "advance")))
        return rc;
    if (!(outer > 0.0)) return set_error(CMF_ERR_INVALID, "advance: outer step must be > 0");
    hipStream_t s = (hipStream_t)stream;
    const bool per_trial = v.has(CMF_PER_TRIAL_DT) || net->G > 0;
    const int64_t nslots = net->G > 0 ? net->G : (v.has(CMF_PER_TRIAL_DT) ? v.T : 1);
    int count = fixed_count;
    if (v.has(CMF_HAS_LOG) && (net->G > 0 ? net->nblk : (v.N * v.T + 63) / 64) > net->partial_blocks)
        return set_error(CMF_ERR_INTERNAL, "advance: partial buffer too small");
    if (v.has(CMF_HAS_ADAPTIVE_TIME)) {
        if ((rc = require({NEED(max_sub_steps), NEED(outer_time_step)}, "advance"))) return rc;
        // the early advance leaves current_step at count - 1 here (one slot)
        const bool early = !use_graph && early_ok(net, v);
        const bool stored = net->G == 0 && v.N > 0 && net->host_counts_dev && cfl_inner_ok(net, v);
        if (stored) {
            // two launches: the CFL maxima, then one block that folds them,
            // sets the time steps and stores the counts in pinned host memory,
            // where the host waits for them to replace its -1s
            for (int64_t t = 0; t < nslots; ++t) net->host_counts[t] = -1;
            const CflTail<REAL> tail{f->max_sub_steps, static_cast<REAL*>(f->time_step),
                                     static_cast<REAL*>(f->outer_time_step), early ? f->current_step : nullptr,
                                     net->host_counts_dev, outer, v.has(CMF_PER_TRIAL_DT) ? 1 : 0};
            if ((rc = launch_cfl(net, v, s, &tail))) return rc;
        } else {
            hipLaunchKernelGGL((k_prepare_counts<REAL>), dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s,
                               static_cast<REAL*>(f->outer_time_step), outer, f->max_sub_steps, nslots);
            if (net->G > 0) {
                hipLaunchKernelGGL((k_cfl_grouped<REAL, STO>),
                                   dim3((unsigned)((net->nblk + kCflRecords - 1) / kCflRecords)), dim3(kGroupRecord),
                                   0, s, v, (int64_t)net->nblk);
            } else if (v.N > 0) {
                if ((rc = launch_cfl(net, v, s))) return rc;
            }
            hipLaunchKernelGGL((k_set_dt<REAL>), dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s,
                               static_cast<REAL*>(f->time_step), f->max_sub_steps, nslots, outer,
                               early ? f->current_step : nullptr);
            CMF_HIP_CHECK(hipGetLastError());
            CMF_HIP_CHECK(hipMemcpyAsync(net->host_counts, f->max_sub_steps, sizeof(int32_t) * nslots,
                                         hipMemcpyDeviceToHost, s));
        }
        if (early) return early_advance(net, v, f, out_count, s, stored);
        if (stored) {
            if ((rc = wait_counts(net, nslots, s))) return rc;
        } else {
            CMF_HIP_CHECK(hipStreamSynchronize(s));
        }
        count = 1;
        for (int64_t t = 0; t < nslots; ++t) count = std::max(count, (int)net->host_counts[t]);
    } else {
        if (fixed_count < 1) return set_error(CMF_ERR_INVALID, "advance: fixed_count must be >= 1");
        if (per_trial) {
            if (!f->max_sub_steps) return set_error(CMF_ERR_INVALID, "advance: max_sub_steps is NULL");
            hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s,
                               f->max_sub_steps, nslots, (int32_t)fixed_count);
        }
        if (f->outer_time_step)
            hipLaunchKernelGGL((k_prepare<REAL>), dim3(1), dim3(64), 0, s, static_cast<REAL*>(f->outer_time_step), outer);
        hipLaunchKernelGGL((k_set_dt_fixed<REAL>), dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s,
                           static_cast<REAL*>(f->time_step), nslots, outer / (double)fixed_count);
        CMF_HIP_CHECK(hipGetLastError());
    }
    if (net->G > 0 && v.has(CMF_HAS_ADAPTIVE_TIME) && !use_graph) {
        if ((rc = build_level_tables(net, f->max_sub_steps, count, s))) return rc;
        v.ordered = true;
    }
    rc = use_graph ? substep_loop_graph(net, v, f, count, s) : substep_loop(net, v, f, count, s);
    if (rc) return rc;
    if (out_count) *out_count = count;
    return CMF_OK;
}

// One fused sub-step that also records the adjoint's tape: the outflow pass
// writes the input state, the unlimited flows and limiters go to the tape
// instead of the network scratch.
template <class REAL, class STO>
int run_substep_taped(cmf_network* net, const cmf_fields* f, int substep, int store_inflow, const cmf_tape* tp,
                      void* stream) {
    View<REAL, STO> v;
    int rc = make_view<REAL, STO>(net, f, v);
    if (rc) return rc;
    if ((rc = need_base(f, "substep_taped")) || (rc = need_features(f, "substep_taped"))) return rc;
    if (!tp) return set_error(CMF_ERR_INVALID, "substep_taped: null tape");
    if (v.T != 1 || net->G > 0)
        return set_error(CMF_ERR_UNSUPPORTED, "substep_taped: one trial, ungrouped");
    void* ptrs[] = {tp->river_outflow, tp->flood_outflow, tp->river_cross_section_depth,
                    tp->flood_cross_section_depth, tp->flood_cross_section_area, tp->river_storage,
                    tp->flood_storage, tp->river_depth, tp->flood_depth, tp->qraw_r, tp->qraw_f, tp->lim};
    for (void* p : ptrs)
        if (!p) return set_error(CMF_ERR_INVALID, "substep_taped: every tape slot is needed");
    v.tape_q_r = static_cast<REAL*>(tp->river_outflow);
    v.tape_q_f = static_cast<REAL*>(tp->flood_outflow);
    v.tape_rcsd = static_cast<REAL*>(tp->river_cross_section_depth);
    v.tape_fcsd = static_cast<REAL*>(tp->flood_cross_section_depth);
    v.tape_fcsa = static_cast<REAL*>(tp->flood_cross_section_area);
    v.tape_rs = static_cast<STO*>(tp->river_storage);
    v.tape_fs = static_cast<STO*>(tp->flood_storage);
    v.tape_rd = static_cast<REAL*>(tp->river_depth);
    v.tape_fd = static_cast<REAL*>(tp->flood_depth);
    v.qraw_r = static_cast<REAL*>(tp->qraw_r);
    v.qraw_f = static_cast<REAL*>(tp->qraw_f);
    v.lim = static_cast<REAL*>(tp->lim);
    if (v.has(CMF_HAS_LEVEE)) {
        if (!tp->protected_storage || !tp->protected_depth)
            return set_error(CMF_ERR_INVALID, "substep_taped: the protected storage and depth slots are needed "
                                              "with levees");
        v.tape_ps = static_cast<STO*>(tp->protected_storage);
        v.tape_pd = static_cast<REAL*>(tp->protected_depth);
    }
    if (v.has(CMF_HAS_RESERVOIR)) {
        if (!tp->reservoir_total_inflow)
            return set_error(CMF_ERR_INVALID, "substep_taped: the reservoir inflow slot is needed with reservoirs");
        v.tape_ri = static_cast<STO*>(tp->reservoir_total_inflow);
    }
    if (v.has(CMF_HAS_BIFURCATION) && v.P > 0) {
        if (!tp->bifurcation_outflow || !tp->bifurcation_cross_section_depth)
            return set_error(CMF_ERR_INVALID, "substep_taped: the path state slots are needed with bifurcation");
        v.tape_bq = static_cast<REAL*>(tp->bifurcation_outflow);
        v.tape_bcsd = static_cast<REAL*>(tp->bifurcation_cross_section_depth);
        if (tp->path_terms) {
            REAL* pt = static_cast<REAL*>(tp->path_terms);
            v.tape_bqlim = pt;
            v.path_raw = pt + v.P * v.LB;
            v.path_pos = v.path_raw + v.P;
            v.path_neg = v.path_pos + v.P;
        }
    }
    return fused_substep(net, v, f, substep, store_inflow, stream);
}

// Reverse sweep of one fused sub-step (base physics, one trial, ungrouped).
template <class REAL, class STO>
int run_adjoint(cmf_network* net, const cmf_fields* f, const cmf_adjoint* adj, void* stream) {
    View<REAL, STO> v;
    int rc = make_view<REAL, STO>(net, f, v);
    if (rc) return rc;
    if (!adj) return set_error(CMF_ERR_INVALID, "substep_adjoint: null adjoint table");
    if (v.T != 1 || net->G > 0 || f->batched)
        return set_error(CMF_ERR_UNSUPPORTED, "substep_adjoint: one trial, ungrouped, no per-trial parameters");
    if ((rc = require({NEED(downstream_idx), NEED(river_manning), NEED(flood_manning), NEED(river_width),
                       NEED(river_length), NEED(river_height), NEED(catchment_elevation), NEED(catchment_area),
                       NEED(downstream_distance), NEED(protected_storage), NEED(runoff), NEED(flood_depth_table)},
                      "substep_adjoint")))
        return rc;
    const void* state[ADJ_COUNT] = {adj->river_outflow, adj->flood_outflow, adj->river_cross_section_depth,
                                    adj->flood_cross_section_depth, adj->flood_cross_section_area,
                                    adj->river_storage, adj->flood_storage, adj->river_depth, adj->flood_depth};
    for (int k = 0; k < ADJ_COUNT; ++k)
        if (!state[k] || !adj->adj_out[k] || !adj->adj_in[k])
            return set_error(CMF_ERR_INVALID, "substep_adjoint: state, adj_out and adj_in need all %d fields", ADJ_COUNT);
    if (!adj->grad_river_manning || !adj->grad_river_height)
        return set_error(CMF_ERR_INVALID, "substep_adjoint: gradient accumulators are NULL");
    if (!(adj->time_step > 0.0) && !(adj->mode & CMF_ADJ_FLUSH))
        return set_error(CMF_ERR_INVALID, "substep_adjoint: time_step must be > 0");
    const int64_t n = v.N;
    const bool bif = v.has(CMF_HAS_BIFURCATION) && v.P > 0;
    if (bif) {
        if ((rc = require({NEED(bifurcation_catchment_idx), NEED(bifurcation_downstream_idx),
                           NEED(bifurcation_manning), NEED(bifurcation_width), NEED(bifurcation_length),
                           NEED(bifurcation_elevation)}, "substep_adjoint")))
            return rc;
        if (v.LB < 1 || v.LB > kAdjMaxLevels)
            return set_error(CMF_ERR_UNSUPPORTED, "substep_adjoint: 1 to %d bifurcation levels", kAdjMaxLevels);
        if (!(adj->mode & CMF_ADJ_FLUSH) &&
            (!adj->bifurcation_outflow || !adj->bifurcation_cross_section_depth || !adj->adj_out_path[0] ||
             !adj->adj_out_path[1] || !adj->adj_in_path[0] || !adj->adj_in_path[1]))
            return set_error(CMF_ERR_INVALID, "substep_adjoint: the path state and its adjoints are needed with "
                                              "bifurcation");
    }
    const bool lev = v.has(CMF_HAS_LEVEE);
    if (lev) {
        if ((rc = require({NEED(is_levee), NEED(levee_index), NEED(levee_base_height), NEED(levee_crown_height),
                           NEED(levee_fraction), NEED(protected_depth)}, "substep_adjoint")))
            return rc;
        if (!(adj->mode & CMF_ADJ_FLUSH) &&
            (!adj->protected_storage || !adj->protected_depth || !adj->adj_out_levee[0] || !adj->adj_out_levee[1] ||
             !adj->adj_in_levee[0] || !adj->adj_in_levee[1]))
            return set_error(CMF_ERR_INVALID, "substep_adjoint: the protected storage and depth and their adjoints "
                                              "are needed with levees");
    }
    const bool resv = v.has(CMF_HAS_RESERVOIR);
    if (resv) {
        if ((rc = require({NEED(reservoir_catchment_idx), NEED(conservation_volume), NEED(emergency_volume),
                           NEED(adjustment_volume), NEED(effective_normal_outflow), NEED(adjustment_outflow),
                           NEED(flood_control_outflow), NEED(is_reservoir), NEED(is_dam_upstream)},
                          "substep_adjoint")))
            return rc;
        if (!(adj->mode & CMF_ADJ_FLUSH) &&
            (!adj->reservoir_total_inflow || !adj->adj_out_reservoir || !adj->adj_in_reservoir || !adj->qraw_r ||
             !adj->qraw_f || !adj->lim))
            return set_error(CMF_ERR_INVALID, "substep_adjoint: with reservoirs the gathered inflow, its adjoints "
                                              "and the taped intermediates are needed");
    }
    if (!net->adj_scratch && (rc = dev_alloc(&net->adj_scratch, (size_t)n * 8 * 14))) return rc;
    if (resv && !net->adj_res_rows && (rc = dev_alloc(&net->adj_res_rows, (size_t)n * 4))) return rc;
    if (bif && !net->adj_path_scratch &&
        (rc = dev_alloc(&net->adj_path_scratch, (size_t)v.P * 8 * (size_t)(v.LB + 9))))
        return rc;
    AdjView<REAL> a{};
    a.q_r = static_cast<const REAL*>(adj->river_outflow);
    a.q_f = static_cast<const REAL*>(adj->flood_outflow);
    a.rcsd = static_cast<const REAL*>(adj->river_cross_section_depth);
    a.fcsd = static_cast<const REAL*>(adj->flood_cross_section_depth);
    a.fcsa = static_cast<const REAL*>(adj->flood_cross_section_area);
    a.rs_raw = adj->river_storage;
    a.fs_raw = adj->flood_storage;
    a.rd = static_cast<const REAL*>(adj->river_depth);
    a.fd = static_cast<const REAL*>(adj->flood_depth);
    for (int k = 0; k < ADJ_COUNT; ++k) {
        a.lout[k] = static_cast<const REAL*>(adj->adj_out[k]);
        a.lin[k] = static_cast<REAL*>(adj->adj_in[k]);
    }
    a.g_n = adj->grad_river_manning;
    a.g_h = adj->grad_river_height;
    REAL* r = static_cast<REAL*>(net->adj_scratch);
    const int64_t st = n * (int64_t)(8 / sizeof(REAL));
    REAL* w[14];
    for (int k = 0; k < 14; ++k) w[k] = r + k * st;
    a.w_qr = w[0]; a.w_qf = w[1]; a.w_lim = w[2]; a.pos = w[3]; a.neg = w[4];
    a.a_rout = w[5]; a.a_fout = w[6]; a.a_ri = w[7]; a.a_fi = w[8]; a.a_qr = w[9]; a.a_qf = w[10];
    a.a_og = w[11]; a.d_wd = w[12]; a.a_gb = w[13];
    a.bif = bif;
    a.lev = lev;
    a.res = resv;
    if (resv) {
        a.ri_raw = adj->reservoir_total_inflow;
        a.lout_ri = static_cast<const REAL*>(adj->adj_out_reservoir);
        a.lin_ri = static_cast<REAL*>(adj->adj_in_reservoir);
        a.res_row = static_cast<int32_t*>(net->adj_res_rows);
    }
    if (lev) {
        a.ps_raw = adj->protected_storage;
        a.pd = static_cast<const REAL*>(adj->protected_depth);
        a.lout_ps = static_cast<const REAL*>(adj->adj_out_levee[0]);
        a.lout_pd = static_cast<const REAL*>(adj->adj_out_levee[1]);
        a.lin_ps = static_cast<REAL*>(adj->adj_in_levee[0]);
        a.lin_pd = static_cast<REAL*>(adj->adj_in_levee[1]);
    }
    if (bif) {
        a.bq = static_cast<const REAL*>(adj->bifurcation_outflow);
        a.bcsd = static_cast<const REAL*>(adj->bifurcation_cross_section_depth);
        a.lout_bq = static_cast<const REAL*>(adj->adj_out_path[0]);
        a.lout_bcsd = static_cast<const REAL*>(adj->adj_out_path[1]);
        a.lin_bq = static_cast<REAL*>(adj->adj_in_path[0]);
        a.lin_bcsd = static_cast<REAL*>(adj->adj_in_path[1]);
        REAL* pr = static_cast<REAL*>(net->adj_path_scratch);
        const int64_t pst = v.P * (int64_t)(8 / sizeof(REAL));
        a.p_qlim = pr;
        REAL* q = pr + v.LB * pst;
        REAL** one[] = {&a.p_raw, &a.p_pos, &a.p_neg, &a.p_w0, &a.p_w1, &a.p_t0, &a.p_t1, &a.p_d0, &a.p_d1};
        for (int k = 0; k < 9; ++k) *one[k] = q + k * pst;
        if (adj->path_terms) {
            // the forward's own limited level flows, level sums and terms
            REAL* pt = const_cast<REAL*>(static_cast<const REAL*>(adj->path_terms));
            a.p_qlim = pt;
            a.p_raw = pt + v.P * v.LB;
            a.p_pos = a.p_raw + v.P;
            a.p_neg = a.p_pos + v.P;
        }
    }
    a.dt = (REAL)adj->time_step;
    const bool taped = adj->qraw_r && adj->qraw_f && adj->lim;
    a.qr = taped ? static_cast<const REAL*>(adj->qraw_r) : a.w_qr;
    a.qf = taped ? static_cast<const REAL*>(adj->qraw_f) : a.w_qf;
    a.lim = taped ? static_cast<const REAL*>(adj->lim) : a.w_lim;
    const bool contig = net->contiguous != 0;
    if (adj->mode & CMF_ADJ_FLUSH) {
        // the deferred gather of the last call only
        if (contig) LAUNCH((k_adj_gather<REAL, STO, true>), n, stream, v, a);
        else LAUNCH((k_adj_gather<REAL, STO, false>), n, stream, v, a);
        return CMF_OK;
    }
    if (resv) LAUNCH((k_adj_res_rows<REAL, STO>), v.R, stream, v, a);
    if (bif && !adj->path_terms) LAUNCH_B((k_adj_path_fwd<REAL, STO>), v.P, 64, stream, v, a);
    if (!taped) {
        if (resv) LAUNCH((k_adj_outflow_fwd<REAL, STO, true>), n, stream, v, a);
        else LAUNCH((k_adj_outflow_fwd<REAL, STO, false>), n, stream, v, a);
        if (contig) LAUNCH((k_adj_limiter_fwd<REAL, STO, true>), n, stream, v, a);
        else LAUNCH((k_adj_limiter_fwd<REAL, STO, false>), n, stream, v, a);
    }
    const bool pending = (adj->mode & CMF_ADJ_PENDING) != 0;
    const bool defer = (adj->mode & CMF_ADJ_DEFER) != 0;
#define CMF_AS(C, LF, PD) LAUNCH((k_adj_stage<REAL, STO, C, LF, PD>), n, stream, v, a)
    if (contig) {
        if (v.L == 10) { if (pending) CMF_AS(true, 10, true); else CMF_AS(true, 10, false); }
        else { if (pending) CMF_AS(true, 0, true); else CMF_AS(true, 0, false); }
        LAUNCH((k_adj_route<REAL, STO, true>), n, stream, v, a);
    } else {
        if (v.L == 10) { if (pending) CMF_AS(false, 10, true); else CMF_AS(false, 10, false); }
        else { if (pending) CMF_AS(false, 0, true); else CMF_AS(false, 0, false); }
        LAUNCH((k_adj_route<REAL, STO, false>), n, stream, v, a);
    }
#undef CMF_AS
    if (bif) LAUNCH((k_adj_path<REAL, STO>), v.P * 8, stream, v, a);   // eight lanes per path
    if (resv) LAUNCH((k_adj_outflow<REAL, STO, true>), n, stream, v, a);
    else LAUNCH((k_adj_outflow<REAL, STO, false>), n, stream, v, a);
    if (!defer) {
        if (contig) LAUNCH((k_adj_gather<REAL, STO, true>), n, stream, v, a);
        else LAUNCH((k_adj_gather<REAL, STO, false>), n, stream, v, a);
    }
    return CMF_OK;
}

}  // namespace

extern "C" {

const char* cmf_last_error(void) { return g_last_error.c_str(); }
int cmf_version(void) { return 10000; }
int64_t cmf_fields_size(void) { return (int64_t)sizeof(cmf_fields); }

int cmf_network_create_grouped(int64_t n, const int32_t* downstream_idx, int64_t num_paths,
                               const int32_t* bif_ci, const int32_t* bif_di, int64_t num_trials,
                               int64_t num_groups, const int64_t* group_ptr, cmf_network** out) {
    if (!out) return set_error(CMF_ERR_INVALID, "cmf_network_create: out is NULL");
    *out = nullptr;
    if (n < 1 || n > INT32_MAX - 1 || !downstream_idx)
        return set_error(CMF_ERR_INVALID, "cmf_network_create: bad catchment count or downstream_idx");
    if (num_trials < 1) return set_error(CMF_ERR_INVALID, "cmf_network_create: num_trials must be >= 1");
    if (num_paths < 0 || (num_paths > 0 && (!bif_ci || !bif_di)))
        return set_error(CMF_ERR_INVALID, "cmf_network_create: bad bifurcation arguments");
    int rc = cmf_check_flow_direction(n, downstream_idx);
    if (rc) return rc;
    for (int64_t p = 0; p < num_paths; ++p)
        if (bif_ci[p] < 0 || bif_ci[p] >= n || bif_di[p] < 0 || bif_di[p] >= n)
            return set_error(CMF_ERR_INVALID, "cmf_network_create: bifurcation index out of range at path %lld",
                             (long long)p);
    // groups: contiguous ranges closed under downstream links and paths
    std::vector<int32_t> cgroup, gblk_ptr;
    std::vector<int4> rec;
    if (num_groups < 0 || (num_groups > 0 && !group_ptr))
        return set_error(CMF_ERR_INVALID, "cmf_network_create_grouped: bad group arguments");
    if (num_groups > 0) {
        if (num_trials != 1)
            return set_error(CMF_ERR_UNSUPPORTED, "cmf_network_create_grouped: grouped networks take one trial");
        if (group_ptr[0] != 0 || group_ptr[num_groups] != n)
            return set_error(CMF_ERR_INVALID, "cmf_network_create_grouped: group_ptr must run from 0 to n");
        cgroup.resize((size_t)n);
        for (int64_t g = 0; g < num_groups; ++g) {
            if (group_ptr[g + 1] <= group_ptr[g])
                return set_error(CMF_ERR_INVALID, "cmf_network_create_grouped: group %lld is empty or reversed",
                                 (long long)g);
            gblk_ptr.push_back((int32_t)rec.size());
            for (int64_t i = group_ptr[g]; i < group_ptr[g + 1]; ++i) cgroup[i] = (int32_t)g;
            for (int64_t b = group_ptr[g]; b < group_ptr[g + 1]; b += kGroupRecord) {
                int4 r;
                r.x = (int32_t)b;
                r.y = (int32_t)std::min<int64_t>(b + kGroupRecord, group_ptr[g + 1]);
                r.z = (int32_t)rec.size();
                r.w = (int32_t)g;
                rec.push_back(r);
            }
        }
        gblk_ptr.push_back((int32_t)rec.size());
        for (int64_t i = 0; i < n; ++i)
            if (cgroup[downstream_idx[i]] != cgroup[i])
                return set_error(CMF_ERR_INVALID, "cmf_network_create_grouped: catchment %lld drains into another group",
                                 (long long)i);
        for (int64_t p = 0; p < num_paths; ++p)
            if (cgroup[bif_ci[p]] != cgroup[bif_di[p]])
                return set_error(CMF_ERR_INVALID, "cmf_network_create_grouped: bifurcation path %lld joins two groups",
                                 (long long)p);
    }
    // upstream lists in ascending network index
    std::vector<int32_t> up_ptr((size_t)n + 1, 0), up_idx;
    for (int64_t i = 0; i < n; ++i) if (downstream_idx[i] != i) up_ptr[downstream_idx[i] + 1]++;
    for (int64_t i = 0; i < n; ++i) up_ptr[i + 1] += up_ptr[i];
    up_idx.resize((size_t)up_ptr[n]);
    {
        std::vector<int32_t> fill(up_ptr.begin(), up_ptr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            if (downstream_idx[i] != i) up_idx[fill[downstream_idx[i]]++] = (int32_t)i;
    }
    // Contiguous layout: the upstream catchments of every catchment are
    // consecutive indices.  Monotone (the level orders): walking the
    // catchments in order, the upstream catchments of each are the next ones,
    // so one array of starts holds every range (k's ends at the start of
    // k + 1); a mouth whose own index comes next claims itself as well (the
    // basin-by-basin level order puts each mouth right after its upstream
    // range) and kernels skip the owner.  Otherwise (the keyed level order,
    // whose ranges are not in the order of their owners) starts and ends.
    std::vector<int32_t> cup((size_t)n + 1, 0), cend;
    bool contiguous = true, monotone = true;
    {
        int64_t pos = 0;
        for (int64_t k = 0; k < n && monotone; ++k) {
            cup[k] = (int32_t)pos;
            for (int32_t q = up_ptr[k]; q < up_ptr[k + 1]; ++q) {
                if (up_idx[q] != pos) { monotone = false; break; }
                ++pos;
            }
            if (downstream_idx[k] == k && pos == k) ++pos;
        }
        cup[n] = (int32_t)pos;
    }
    if (!monotone) {
        cup.assign((size_t)n + 1, 0);
        cend.assign((size_t)n, 0);
        for (int64_t k = 0; k < n && contiguous; ++k) {
            const int32_t b = up_ptr[k], e = up_ptr[k + 1];
            const int32_t first = e > b ? up_idx[b] : 0;
            for (int32_t q = b; q < e; ++q)
                if (up_idx[q] != first + (q - b)) { contiguous = false; break; }
            cup[k] = first;
            cend[k] = first + (e - b);
        }
    }
    // the end of each range on the host, in every representation
    std::vector<int32_t> range_end((size_t)n);
    if (contiguous && !monotone) range_end = cend;
    else for (int64_t k = 0; k < n; ++k) range_end[k] = (contiguous ? cup : up_ptr)[k + 1];
    std::vector<int2> rng;
    if (contiguous && !monotone) {
        rng.resize((size_t)n);
        for (int64_t k = 0; k < n; ++k) rng[(size_t)k] = make_int2(cup[k], cend[k]);
    }
    if (contiguous) up_ptr.swap(cup);
    // bifurcation events per catchment, ascending path, upstream end first
    std::vector<int32_t> ev_ptr((size_t)n + 1, 0), ev_code((size_t)(2 * num_paths)), ev_other((size_t)(2 * num_paths));
    for (int64_t p = 0; p < num_paths; ++p) { ev_ptr[bif_ci[p] + 1]++; ev_ptr[bif_di[p] + 1]++; }
    for (int64_t i = 0; i < n; ++i) ev_ptr[i + 1] += ev_ptr[i];
    std::vector<int2> pev((size_t)std::max<int64_t>(num_paths, 1));
    {
        std::vector<int32_t> fill(ev_ptr.begin(), ev_ptr.end() - 1);
        for (int64_t p = 0; p < num_paths; ++p) {
            pev[(size_t)p].x = fill[bif_ci[p]];
            ev_other[fill[bif_ci[p]]] = bif_di[p];
            ev_code[fill[bif_ci[p]]++] = (int32_t)(2 * p);
            pev[(size_t)p].y = fill[bif_di[p]];
            ev_other[fill[bif_di[p]]] = bif_ci[p];
            ev_code[fill[bif_di[p]]++] = (int32_t)(2 * p + 1);
        }
    }
    std::vector<int2> ev1((size_t)std::max<int64_t>(n, 1));
    for (int64_t i = 0; i < n; ++i) {
        const int32_t b = ev_ptr[i], e = ev_ptr[i + 1];
        if (e == b) {
            ev1[i].x = -1;
            ev1[i].y = 0;
        } else {
            ev1[i].x = ev_code[b];
            ev1[i].y = (e - b > 1) ? -(ev_other[b] + 1) : ev_other[b];
        }
    }
    std::vector<uint8_t> ev_flag((size_t)std::max<int64_t>(n, 1), 0);
    for (int64_t i = 0; i < n; ++i) ev_flag[(size_t)i] = ev_ptr[i + 1] != ev_ptr[i] ? 1 : 0;
    // path-end records of the limiter fix-up pass, ascending catchment
    std::vector<int4> pend;
    for (int64_t i = 0; i < n; ++i)
        if (ev_ptr[i + 1] != ev_ptr[i])
            pend.push_back(make_int4((int32_t)i, up_ptr[i], range_end[i], ev_ptr[i]));
    const int64_t pend_n = (int64_t)pend.size();
    pend.push_back(make_int4(-1, 0, 0, ev_ptr[n]));
    // block-local schedule tables for blocks of kLocalBlock catchments
    std::vector<uint8_t> lflags;
    std::vector<int32_t> lim_list, stage_list;
    const bool local = contiguous && num_groups == 0 && num_trials == 1;
    if (local) local_tables(n, downstream_idx, up_ptr, range_end, ev_ptr, kLocalBlock, lflags, lim_list, stage_list);
    cmf_network* net = new cmf_network();
    net->n = n;
    net->T = num_trials;
    net->P = num_paths;
    net->contiguous = contiguous ? 1 : 0;
    (void)hipGetDevice(&net->device);
    const size_t el = (size_t)n * (size_t)num_trials * 8;
    const size_t pel = (size_t)std::max<int64_t>(num_paths, 1) * (size_t)num_trials * 8;
    net->G = num_groups;
    net->nblk = (int64_t)rec.size();
    // partials for the smallest block size a tunable pass may use (64), and
    // for the block-local pass ahead of its stage list
    net->partial_blocks = std::max<int64_t>((n * num_trials + 63) / 64, net->nblk) + (n + kLocalBlock - 1) / kLocalBlock + 2;
    if (num_trials <= 64 && 64 % num_trials == 0 &&
        (rc = dev_alloc((void**)&net->cfl_part, (size_t)kCflInnerBlocks * (size_t)num_trials * sizeof(int32_t)))) {
        free_network(net);
        return rc;
    }
    if ((rc = upload((void**)&net->up_ptr, up_ptr.data(), up_ptr.size() * 4)) ||
        (contiguous && !monotone && (rc = upload((void**)&net->up_rng, rng.data(), rng.size() * sizeof(int2)))) ||
        (rc = upload((void**)&net->up_idx, up_idx.data(), up_idx.size() * 4)) ||
        (rc = upload((void**)&net->ev_ptr, ev_ptr.data(), ev_ptr.size() * 4)) ||
        (rc = upload((void**)&net->ev_code, ev_code.data(), ev_code.size() * 4)) ||
        (rc = upload((void**)&net->ev_other, ev_other.data(), ev_other.size() * 4)) ||
        (rc = upload((void**)&net->ev1, ev1.data(), (size_t)n * sizeof(int2))) ||
        (rc = upload((void**)&net->ev_flag, ev_flag.data(), ev_flag.size())) ||
        (rc = dev_alloc(&net->pos_term, el)) || (rc = dev_alloc(&net->neg_term, el)) ||
        (rc = dev_alloc(&net->qraw_r, el)) || (rc = dev_alloc(&net->qraw_f, el)) ||
        (rc = dev_alloc(&net->lim, el)) || (rc = dev_alloc(&net->path_pos, pel)) ||
        (rc = dev_alloc(&net->path_neg, pel)) || (rc = dev_alloc(&net->path_net, pel)) ||
        (rc = dev_alloc(&net->path_raw, pel)) ||
        (rc = dev_alloc((void**)&net->partials, (size_t)net->partial_blocks * CMF_LOG_COUNT * 8)) ||
        (rc = dev_alloc((void**)&net->bmask, (((size_t)n * (size_t)num_trials + 31) / 32) * 4)) ||
        (rc = upload((void**)&net->pend_list, pend.data(), pend.size() * sizeof(int4))) ||
        (rc = upload((void**)&net->pev, pev.data(), pev.size() * sizeof(int2))) ||
        (rc = dev_alloc(&net->eterm, 2 * pel))) {
        free_network(net);
        return rc;
    }
    net->pend_count = pend_n;
    {
        net->bmask_words = ((int64_t)n * num_trials + 31) / 32;
        const hipError_t me = hipMemset(net->bmask, 0, (size_t)net->bmask_words * 4);
        if (me == hipSuccess) (void)hipDeviceSynchronize();
        if (me != hipSuccess) {
            free_network(net);
            return set_error(CMF_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(me));
        }
    }
    if (local) {
        net->local_block = kLocalBlock;
        net->lim_count = (int64_t)lim_list.size();
        net->stage_count = (int64_t)stage_list.size();
        if ((rc = upload((void**)&net->lflags, lflags.data(), lflags.size())) ||
            (rc = upload((void**)&net->lim_list, lim_list.data(), lim_list.size() * 4)) ||
            (rc = upload((void**)&net->stage_list, stage_list.data(), stage_list.size() * 4))) {
            free_network(net);
            return rc;
        }
    }
    if (num_groups > 0 &&
        ((rc = upload((void**)&net->cgroup, cgroup.data(), cgroup.size() * 4)) ||
         (rc = upload((void**)&net->blk_rec_id, rec.data(), rec.size() * sizeof(int4))) ||
         (rc = upload((void**)&net->gblk_ptr, gblk_ptr.data(), gblk_ptr.size() * 4)))) {
        free_network(net);
        return rc;
    }
    const size_t slots = (size_t)std::max<int64_t>(num_trials, num_groups);
    hipError_t e = hipHostMalloc((void**)&net->host_counts, sizeof(int32_t) * slots, hipHostMallocDefault);
    if (e != hipSuccess) {
        free_network(net);
        return set_error(CMF_ERR_NOMEM, "hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    // kernels may store the sub-step counts there themselves; without a device
    // address they are copied
    if (hipHostGetDevicePointer((void**)&net->host_counts_dev, net->host_counts, 0) != hipSuccess) {
        net->host_counts_dev = nullptr;
        (void)hipGetLastError();
    }
    if (num_groups > 0) {
        net->gblk_host = gblk_ptr;
        if ((e = hipEventCreateWithFlags(&net->order_copied, hipEventDisableTiming)) != hipSuccess) {
            free_network(net);
            return set_error(CMF_ERR_NOMEM, "hipEventCreateWithFlags: %s", hipGetErrorString(e));
        }
    }
    *out = net;
    return CMF_OK;
}

int cmf_network_create(int64_t n, const int32_t* downstream_idx, int64_t num_paths,
                       const int32_t* bif_ci, const int32_t* bif_di, int64_t num_trials,
                       cmf_network** out) {
    return cmf_network_create_grouped(n, downstream_idx, num_paths, bif_ci, bif_di, num_trials, 0, nullptr, out);
}

int64_t cmf_network_num_groups(const cmf_network* net) { return net ? net->G : 0; }

int cmf_network_set_outflow_accumulator(cmf_network* net, double* acc) {
    if (!net) return set_error(CMF_ERR_INVALID, "cmf_network_set_outflow_accumulator: null network");
    net->acc_outflow = acc;
    for (auto& g : net->graphs) (void)hipGraphExecDestroy(g.second);
    net->graphs.clear();
    return CMF_OK;
}

int cmf_hip_map_runoff(int64_t rows, const int64_t* indptr, const int32_t* indices, const float* weights,
                       const float* grid, double scale, const int32_t* out_index, int32_t precision, void* out,
                       void* stream) {
    if (rows < 0 || (rows > 0 && (!indptr || !indices || !weights || !grid || !out)))
        return set_error(CMF_ERR_INVALID, "cmf_hip_map_runoff: bad arguments");
    if (rows == 0) return CMF_OK;
    if (precision == CMF_PRECISION_F32_F64 || precision == CMF_PRECISION_F32_F32)
        LAUNCH((k_map_runoff<float>), rows, stream, rows, indptr, indices, weights, grid, scale, out_index,
               static_cast<float*>(out));
    else if (precision == CMF_PRECISION_F64_F64)
        LAUNCH((k_map_runoff<double>), rows, stream, rows, indptr, indices, weights, grid, scale, out_index,
               static_cast<double*>(out));
    else
        return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", precision);
    return CMF_OK;
}

int cmf_network_set_block_sizes(cmf_network* net, const int32_t* sizes, int32_t count) {
    if (!net || !sizes || count < 0 || count > CMF_TUNE_COUNT)
        return set_error(CMF_ERR_INVALID, "cmf_network_set_block_sizes: bad arguments");
    for (int32_t k = 0; k < count; ++k)
        if (sizes[k] < 64 || sizes[k] > kMaxBlock || sizes[k] % 64 != 0)
            return set_error(CMF_ERR_INVALID, "cmf_network_set_block_sizes: size %d of kernel %d is not a multiple "
                             "of 64 in [64, %d]", sizes[k], k, kMaxBlock);
    for (int32_t k = 0; k < count; ++k) net->blocks[k] = sizes[k];
    // captured sub-step loops hold the old launch shapes
    for (auto& g : net->graphs) (void)hipGraphExecDestroy(g.second);
    net->graphs.clear();
    return CMF_OK;
}

int cmf_network_get_block_sizes(const cmf_network* net, int32_t* sizes, int32_t count) {
    if (!net || !sizes || count < 0) return set_error(CMF_ERR_INVALID, "cmf_network_get_block_sizes: bad arguments");
    const int32_t n = std::min<int32_t>(count, CMF_TUNE_COUNT);
    for (int32_t k = 0; k < n; ++k) sizes[k] = net->blocks[k];
    return CMF_OK;
}

int cmf_network_destroy(cmf_network* net) {
    free_network(net);
    return CMF_OK;
}

int cmf_network_is_contiguous(const cmf_network* net) { return net ? net->contiguous : 0; }

int cmf_network_set_schedule(cmf_network* net, int32_t schedule) {
    if (!net) return set_error(CMF_ERR_INVALID, "null network");
    if (schedule != CMF_SCHEDULE_THREE_PASS && schedule != CMF_SCHEDULE_BLOCK_LOCAL &&
        schedule != CMF_SCHEDULE_LIMITER_PASS)
        return set_error(CMF_ERR_INVALID, "unknown schedule %d", schedule);
    net->schedule = schedule;
    // captured graphs hold the launches of the other schedule
    for (auto& g : net->graphs) (void)hipGraphExecDestroy(g.second);
    net->graphs.clear();
    net->graph_fields_set = false;
    return CMF_OK;
}

int cmf_network_local_counts(const cmf_network* net, int64_t* block, int64_t* lim_count, int64_t* stage_count) {
    if (!net || !block || !lim_count || !stage_count) return set_error(CMF_ERR_INVALID, "null argument");
    *block = net->local_block;
    *lim_count = net->lim_count;
    *stage_count = net->stage_count;
    return CMF_OK;
}

int cmf_hip_outflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_OUTFLOW, s); }
int cmf_hip_inflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_INFLOW, s); }
int cmf_hip_flood_stage(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_STAGE, s); }
int cmf_hip_flood_stage_log(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_STAGE_LOG, s); }
int cmf_hip_adaptive_time_step(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_CFL, s); }
int cmf_hip_bifurcation_outflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_BIF_OUT, s); }
int cmf_hip_bifurcation_inflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_BIF_IN, s); }
int cmf_hip_levee_stage(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_LEVEE, s); }
int cmf_hip_levee_stage_log(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_LEVEE_LOG, s); }
int cmf_hip_levee_bifurcation_outflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_LEVEE_BIF_OUT, s); }
int cmf_hip_reservoir_outflow(cmf_network* net, const cmf_fields* f, void* s) { return dispatch_op(net, f, OP_RESERVOIR, s); }

int cmf_hip_flood_stage_log_noalloc(const cmf_fields* f, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64) return run_stage_log_noalloc<float, double>(f, s);
    if (f->precision == CMF_PRECISION_F64_F64) return run_stage_log_noalloc<double, double>(f, s);
    if (f->precision == CMF_PRECISION_F32_F32) return run_stage_log_noalloc<float, float>(f, s);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

int cmf_hip_substep(cmf_network* net, const cmf_fields* f, int32_t substep, int32_t store_inflow, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64) return run_substep<float, double>(net, f, substep, store_inflow, s);
    if (f->precision == CMF_PRECISION_F64_F64) return run_substep<double, double>(net, f, substep, store_inflow, s);
    if (f->precision == CMF_PRECISION_F32_F32) return run_substep<float, float>(net, f, substep, store_inflow, s);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

int cmf_hip_substep_timed(cmf_network* net, const cmf_fields* f, int32_t substep, int32_t store_inflow,
                          void* s, double* out_ms, const char** out_names, int32_t capacity,
                          int32_t* out_count) {
    if (!out_ms || !out_count || capacity < 0) return set_error(CMF_ERR_INVALID, "timed substep: bad outputs");
    Timing t;
    for (int k = 0; k < 2 * Timing::kMax; ++k) {
        hipError_t e = hipEventCreate(&t.ev[k]);
        if (e != hipSuccess) {
            for (int j = 0; j < k; ++j) (void)hipEventDestroy(t.ev[j]);
            return set_error(CMF_ERR_HIP, "hipEventCreate failed: %s", hipGetErrorString(e));
        }
    }
    g_timing = &t;
    int rc = cmf_hip_substep(net, f, substep, store_inflow, s);
    g_timing = nullptr;
    if (rc == CMF_OK) {
        hipError_t e = hipStreamSynchronize((hipStream_t)s);
        if (e != hipSuccess) rc = set_error(CMF_ERR_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    }
    int n = 0;
    if (rc == CMF_OK) {
        for (int k = 0; k < t.count && k < capacity; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, t.ev[2 * k], t.ev[2 * k + 1]);
            out_ms[k] = ms;
            if (out_names) out_names[k] = t.name[k];
            ++n;
        }
    }
    for (int k = 0; k < 2 * Timing::kMax; ++k) (void)hipEventDestroy(t.ev[k]);
    *out_count = n;
    return rc;
}

int cmf_hip_advance(cmf_network* net, const cmf_fields* f, double outer_seconds, int32_t fixed_count,
                    int32_t* out_count, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64)
        return run_advance<float, double>(net, f, outer_seconds, fixed_count, out_count, s, false);
    if (f->precision == CMF_PRECISION_F64_F64)
        return run_advance<double, double>(net, f, outer_seconds, fixed_count, out_count, s, false);
    if (f->precision == CMF_PRECISION_F32_F32)
        return run_advance<float, float>(net, f, outer_seconds, fixed_count, out_count, s, false);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

int cmf_hip_substep_taped(cmf_network* net, const cmf_fields* f, int32_t substep, int32_t store_inflow,
                          const cmf_tape* tape, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64) return run_substep_taped<float, double>(net, f, substep, store_inflow, tape, s);
    if (f->precision == CMF_PRECISION_F64_F64) return run_substep_taped<double, double>(net, f, substep, store_inflow, tape, s);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

int cmf_hip_advance_taped(cmf_network* net, const cmf_fields* f, double outer_seconds, int32_t count,
                          const cmf_tape* tapes, void* s) {
    if (!f || !tapes || count < 1 || !(outer_seconds > 0.0))
        return set_error(CMF_ERR_INVALID, "cmf_hip_advance_taped: bad arguments");
    if (!f->time_step) return set_error(CMF_ERR_INVALID, "cmf_hip_advance_taped: time_step is NULL");
    const double dt = outer_seconds / (double)count;
    if (f->precision == CMF_PRECISION_F32_F64)
        hipLaunchKernelGGL((k_set_dt_fixed<float>), dim3(1), dim3(64), 0, (hipStream_t)s,
                           static_cast<float*>(f->time_step), (int64_t)1, dt);
    else
        hipLaunchKernelGGL((k_set_dt_fixed<double>), dim3(1), dim3(64), 0, (hipStream_t)s,
                           static_cast<double*>(f->time_step), (int64_t)1, dt);
    for (int32_t k = 0; k < count; ++k) {
        const int rc = cmf_hip_substep_taped(net, f, k, k == count - 1 ? 1 : 0, tapes + k, s);
        if (rc) return rc;
    }
    return CMF_OK;
}

int cmf_hip_adjoint_steps(cmf_network* net, const cmf_fields* f, const cmf_adjoint* adjoints, int32_t count,
                          void* s) {
    if (!adjoints || count < 1) return set_error(CMF_ERR_INVALID, "cmf_hip_adjoint_steps: bad arguments");
    for (int32_t k = count - 1; k >= 0; --k) {
        const int rc = cmf_hip_substep_adjoint(net, f, adjoints + k, s);
        if (rc) return rc;
    }
    return CMF_OK;
}

int cmf_hip_substep_adjoint(cmf_network* net, const cmf_fields* f, const cmf_adjoint* a, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64) return run_adjoint<float, double>(net, f, a, s);
    if (f->precision == CMF_PRECISION_F64_F64) return run_adjoint<double, double>(net, f, a, s);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

int cmf_hip_advance_graph(cmf_network* net, const cmf_fields* f, double outer_seconds, int32_t fixed_count,
                          int32_t* out_count, void* s) {
    if (!f) return set_error(CMF_ERR_INVALID, "null field table");
    if (f->precision == CMF_PRECISION_F32_F64)
        return run_advance<float, double>(net, f, outer_seconds, fixed_count, out_count, s, true);
    if (f->precision == CMF_PRECISION_F64_F64)
        return run_advance<double, double>(net, f, outer_seconds, fixed_count, out_count, s, true);
    if (f->precision == CMF_PRECISION_F32_F32)
        return run_advance<float, float>(net, f, outer_seconds, fixed_count, out_count, s, true);
    return set_error(CMF_ERR_UNSUPPORTED, "unknown precision %d", f->precision);
}

}  // extern "C"
