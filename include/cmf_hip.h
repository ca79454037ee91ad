/*
 * cmf_hip.h — C ABI of the MI355X (gfx950) CaMa-Flood hydrodynamic hot path.
 *
 * One shared library, libcmfhip.so, built from cama-flood-gpu_amd/csrc.
 * Every entry point returns a cmf_status (0 = success).  On failure
 * cmf_last_error() returns a message for the calling thread.
 *
 * Buffers are borrowed, never owned: the caller allocates every field (the
 * reference's module fields, cmfgpu/modules/base.py:92-455) and passes raw
 * device pointers in a cmf_fields table.  The library owns only the network
 * scratch it allocates in cmf_network_create and frees in cmf_network_destroy.
 *
 * Launches are asynchronous on the caller's HIP stream.  Only
 * cmf_hip_advance synchronises, once per outer step, to read the CFL
 * sub-step count back (as CaMaFlood.step_advance does with .item(),
 * cmfgpu/models/cama_flood_model.py:122).
 *
 * Accumulations that the reference performs with float atomics are
 * deterministic here: each target sums its contributions in a fixed order
 * (upstream catchments in ascending network index, then its own term, then
 * bifurcation paths in ascending path index), so results are bitwise
 * reproducible run to run.
 */
#ifndef CMF_HIP_H
#define CMF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cmf_status {
    CMF_OK = 0,
    CMF_ERR_INVALID = 1,     /* bad argument, shape or network */
    CMF_ERR_HIP = 2,         /* HIP runtime or launch failure */
    CMF_ERR_NOMEM = 3,       /* device allocation failed */
    CMF_ERR_UNSUPPORTED = 4, /* feature or dtype combination not built */
    CMF_ERR_INTERNAL = 5
} cmf_status;

/* Feature flags: the reference's compile-time HAS_* constants
 * (cmfgpu/phys/specs.py:28-39), plus the log and adaptive-time modules. */
enum {
    CMF_HAS_BIFURCATION = 1u << 0,
    CMF_HAS_LEVEE = 1u << 1,
    CMF_HAS_RESERVOIR = 1u << 2,
    CMF_HAS_LOG = 1u << 3,
    CMF_HAS_TOTAL_STORAGE = 1u << 4,
    CMF_HAS_WATER_SURFACE = 1u << 5,
    CMF_HAS_PROTECTED_WATER_SURFACE = 1u << 6,
    CMF_HAS_ADAPTIVE_TIME = 1u << 7,
    /* time_step (and max_sub_steps) hold one entry per trial, so every trial
     * advances with its own CFL sub-step (an extension of the reference, whose
     * trials share one time_step, cmfgpu/phys/triton/outflow.py:454). */
    CMF_PER_TRIAL_DT = 1u << 8,
    /* FP32 profile: water surfaces and their differences (slopes, depths
     * above the bed, the bank and path sills) formed in double from the FP32
     * depth, elevation and height, each difference rounded to FP32 once.  An
     * extension of the reference, whose FP32 surfaces
     * (cmfgpu/phys/triton/outflow.py:111-134, bifurcation.py:49-69) lose the
     * centimetres their differences are made of; discharge then follows the
     * FP64 run about 50 times closer.  No effect in the FP64 profile; the
     * adjoint follows the forward's values with it. */
    CMF_PRECISE_SURFACE = 1u << 9
};

/* Per-trial parameter flags: the reference's batched_* constants
 * (cmfgpu/phys/specs.py:169-176,199-206,471-477,651-658,338-341,414). */
enum {
    CMF_BATCHED_RIVER_MANNING = 1u << 0,
    CMF_BATCHED_FLOOD_MANNING = 1u << 1,
    CMF_BATCHED_RIVER_WIDTH = 1u << 2,
    CMF_BATCHED_RIVER_LENGTH = 1u << 3,
    CMF_BATCHED_RIVER_HEIGHT = 1u << 4,
    CMF_BATCHED_CATCHMENT_ELEVATION = 1u << 5,
    CMF_BATCHED_DOWNSTREAM_DISTANCE = 1u << 6,
    CMF_BATCHED_RUNOFF = 1u << 7,
    CMF_BATCHED_FLOOD_DEPTH_TABLE = 1u << 8,
    CMF_BATCHED_CATCHMENT_AREA = 1u << 9,
    CMF_BATCHED_LEVEE_CROWN_HEIGHT = 1u << 10,
    CMF_BATCHED_LEVEE_FRACTION = 1u << 11,
    CMF_BATCHED_LEVEE_BASE_HEIGHT = 1u << 12,
    CMF_BATCHED_BIFURCATION_MANNING = 1u << 13,
    CMF_BATCHED_BIFURCATION_WIDTH = 1u << 14,
    CMF_BATCHED_BIFURCATION_LENGTH = 1u << 15,
    CMF_BATCHED_BIFURCATION_ELEVATION = 1u << 16
};

/* Precision profiles: REAL is the flux/stage dtype, STO the hpfloat storage
 * dtype (cmfgpu/phys/cuda/outflow.cu:301-366 dispatches the same three pairs;
 * F32_F32 is the reference's precision float32 without mixed precision,
 * cmfgpu/phys/triton/utils.py:18-30).  F32_F32 runs the forward entry points
 * (per-op, sub-step, advance, graph); the taped and adjoint entry points take
 * the other two.  The water-balance sums stay FP64 in every profile. */
enum { CMF_PRECISION_F32_F64 = 0, CMF_PRECISION_F64_F64 = 1, CMF_PRECISION_F32_F32 = 2 };

/* Tunable kernels of the fused schedule (block sizes: multiples of 64 up to
 * 512, see cmf_network_set_block_sizes). */
enum {
    CMF_TUNE_OUTFLOW = 0, CMF_TUNE_RESERVOIR, CMF_TUNE_LIMITER, CMF_TUNE_STAGE, CMF_TUNE_COUNT
};

/* The 11 water-balance sums, in LogModule field order
 * (cmfgpu/modules/log.py:207-326). Held in FP64. */
enum {
    CMF_LOG_STORAGE_PRE = 0, CMF_LOG_STORAGE_NEXT, CMF_LOG_STORAGE_NEW,
    CMF_LOG_INFLOW_ERROR, CMF_LOG_INFLOW, CMF_LOG_OUTFLOW, CMF_LOG_STORAGE_STAGE,
    CMF_LOG_STAGE_ERROR, CMF_LOG_RIVER_STORAGE, CMF_LOG_FLOOD_STORAGE,
    CMF_LOG_FLOOD_AREA, CMF_LOG_COUNT
};

/*
 * Field table.  Names are the reference's module field names (the KernelSpec
 * parameter names without "_ptr", cmfgpu/phys/specs.py:124-803).  void*
 * members hold REAL or STO data according to `precision`; the comment gives
 * which.  Catchment state is laid out [num_trials][num_catchments] (the
 * reference's trial layout) unless trial_stride/catchment_stride say
 * otherwise: element (t, i) lives at t * trial_stride + i * catchment_stride.
 * Per-path bifurcation state uses path_trial_stride/path_stride the same way
 * (per path and level: element * num_bifurcation_levels + level).
 * Unused optional fields may be NULL.
 */
typedef struct cmf_fields {
    /* dimensions */
    int64_t num_catchments;
    int64_t num_trials;              /* >= 1 */
    int64_t num_flood_levels;
    int64_t num_bifurcation_paths;
    int64_t num_bifurcation_levels;
    int64_t num_levees;
    int64_t num_reservoirs;
    int64_t log_buffer_size;
    int64_t trial_stride;            /* 0 selects num_catchments */
    int64_t catchment_stride;        /* 0 selects 1 */
    int64_t path_trial_stride;       /* 0 selects num_bifurcation_paths */
    int64_t path_stride;             /* 0 selects 1 */
    int32_t precision;               /* CMF_PRECISION_* */
    uint32_t flags;                  /* CMF_HAS_* */
    uint32_t batched;                /* CMF_BATCHED_* */
    uint32_t reserved0;
    /* scalars (base.py:76-87, adaptive_time.py:30-35) */
    double gravity;
    double adaptive_time_factor;
    double min_kinematic_slope;
    /* shared-state device scalars (base.py:424-455, adaptive_time.py:37-45) */
    void* time_step;                 /* REAL[1], or REAL[num_trials] with CMF_PER_TRIAL_DT */
    void* outer_time_step;           /* REAL[1] */
    int32_t* current_step;           /* int32[1] */
    int32_t* max_sub_steps;          /* int32[1], or int32[num_trials] with CMF_PER_TRIAL_DT */
    /* topology */
    const int32_t* downstream_idx;   /* [N] */
    /* base module */
    void* river_inflow;              /* STO */
    void* flood_inflow;              /* STO */
    void* river_outflow;             /* REAL */
    void* flood_outflow;             /* REAL */
    void* river_manning;             /* REAL */
    void* flood_manning;             /* REAL */
    void* river_depth;               /* REAL */
    void* flood_depth;               /* REAL */
    void* protected_depth;           /* REAL */
    void* river_width;               /* REAL */
    void* river_length;              /* REAL */
    void* river_height;              /* REAL */
    void* catchment_elevation;       /* REAL */
    void* catchment_area;            /* REAL */
    void* downstream_distance;       /* REAL */
    void* river_storage;             /* STO */
    void* flood_storage;             /* STO */
    void* protected_storage;         /* STO */
    void* river_cross_section_depth; /* REAL */
    void* flood_cross_section_depth; /* REAL */
    void* flood_cross_section_area;  /* REAL */
    void* flood_fraction;            /* REAL */
    void* runoff;                    /* REAL */
    void* flood_depth_table;         /* REAL [N, num_flood_levels] */
    void* outgoing_storage;          /* STO */
    void* total_storage;             /* STO, with CMF_HAS_TOTAL_STORAGE */
    void* water_surface_elevation;   /* REAL, with CMF_HAS_WATER_SURFACE */
    void* protected_water_surface_elevation; /* REAL */
    const uint8_t* is_levee;         /* bool [N], with CMF_HAS_LEVEE */
    /* bifurcation module (bifurcation.py:64-214) */
    const int32_t* bifurcation_catchment_idx;   /* [P] */
    const int32_t* bifurcation_downstream_idx;  /* [P] */
    void* bifurcation_manning;       /* REAL [P, Lb] */
    void* bifurcation_outflow;       /* REAL [P, Lb] per trial */
    void* bifurcation_width;         /* REAL [P, Lb] */
    void* bifurcation_length;        /* REAL [P] */
    void* bifurcation_elevation;     /* REAL [P, Lb] */
    void* bifurcation_cross_section_depth; /* REAL [P, Lb] per trial */
    void* global_bifurcation_outflow;      /* STO [N] per trial */
    void* limit_rate;                /* REAL [N] per trial */
    /* levee module (levee.py:73-244) */
    const int32_t* levee_catchment_idx;     /* [V] */
    void* levee_base_height;         /* REAL [V] */
    void* levee_crown_height;        /* REAL [V] */
    void* levee_fraction;            /* REAL [V] */
    /* reservoir module (reservoir.py:64-265) */
    const int32_t* reservoir_catchment_idx; /* [R] */
    void* conservation_volume;       /* REAL [R] */
    void* emergency_volume;          /* REAL [R] */
    void* adjustment_volume;         /* REAL [R] */
    void* effective_normal_outflow;  /* REAL [R] */
    void* adjustment_outflow;        /* REAL [R] */
    void* flood_control_outflow;     /* REAL [R] */
    void* reservoir_total_inflow;    /* STO [N] per trial */
    const uint8_t* is_reservoir;     /* bool [N] */
    const uint8_t* is_dam_upstream;  /* bool [N] */
    const uint8_t* is_dam_related;   /* bool [N] */
    /* levee row of each catchment, -1 where there is none ([N]; optional:
     * when set, the fused sub-step runs the levee stage inside the flood
     * stage pass) */
    const int32_t* levee_index;
    /* log module: CMF_LOG_COUNT FP64 buffers of log_buffer_size entries each;
     * sum k of sub-step s accumulates into log_sums[k][s].  All trials add
     * into the one set (the reference's LogModule supports one trial,
     * cmfgpu/models/cama_flood_model.py:65-67). */
    double* log_sums[CMF_LOG_COUNT];
} cmf_fields;

/* Opaque network context: upstream lists, bifurcation event lists and
 * device scratch, built once per network (and trial count). */
typedef struct cmf_network cmf_network;

/* ------------------------------------------------------------------ info */
const char* cmf_last_error(void);
int cmf_version(void);
/* Size of cmf_fields as compiled into the library (ABI check). */
int64_t cmf_fields_size(void);

/* -------------------------------------------------------- host indexing */
/* cmfgpu/params/utils.py:20-80 */
int cmf_trace_outlets(int64_t n, const int64_t* catchment_id,
                      const int64_t* downstream_id, int64_t* out_mouth);
/* cmfgpu/params/utils.py:83-138; raises on rings like :135-136 */
int cmf_topological_sort(int64_t n, const int64_t* catchment_id,
                         const int64_t* downstream_id, int32_t* out_idx);
/* cmfgpu/params/utils.py:483-503 (ties between equal basin sizes: ascending label) */
int cmf_reorder_by_basin_size(int64_t n, const int64_t* topo_idx, const int64_t* basin_id,
                              int64_t* out_order, int64_t* out_basin_sizes,
                              int64_t* out_num_basins);
/* hydroforge find_indices_in as used at cmfgpu/params/merit_map.py:450 (-1 = absent) */
int cmf_find_indices_in(int64_t n_query, const int64_t* query, int64_t n_ref,
                        const int64_t* ref, int64_t* out);
/* cmfgpu/params/merit_map.py:1063-1070 */
int cmf_check_flow_direction(int64_t n, const int32_t* downstream_idx);
/* cmfgpu/params/estimate_river_geometry.py:76-104 */
int cmf_accumulate_downstream(int64_t n, const int32_t* downstream_idx,
                              const double* local, double* out);
/* Breadth-first level renumbering used for the coalesced device layout
 * (no reference counterpart; the reference keeps its topological order). */
int cmf_level_order(int64_t n, const int32_t* downstream_idx, int32_t* out_perm,
                    int32_t* out_inv, int32_t* out_up_ptr, int32_t* out_dn,
                    int64_t* out_num_levels, int64_t level_ptr_capacity,
                    int64_t* out_level_ptr);
/* Breadth-first levels laid out basin by basin (each tree draining to one
 * mouth is contiguous, its levels deepest first); same outputs as
 * cmf_level_order plus the basin boundaries.  No reference counterpart: the
 * reference groups basins the same way (cmfgpu/params/utils.py:483-503) but
 * keeps its topological order inside them. */
int cmf_basin_level_order(int64_t n, const int32_t* downstream_idx, int32_t* out_perm,
                          int32_t* out_inv, int32_t* out_up_ptr, int32_t* out_dn,
                          int64_t* out_num_basins, int64_t basin_ptr_capacity,
                          int64_t* out_basin_ptr);
/* Breadth-first levels of the whole forest with each level's sibling groups
 * arranged by work class (`key`, one byte per catchment in network order,
 * NULL for none: the model passes bit 0 for a levee catchment and bit 1 for
 * a bifurcation path end), classes alternating direction level by level, so
 * that waves of consecutive catchments mostly hold one class.  Upstream
 * ranges stay contiguous but are not monotone (no up_ptr output).  No
 * reference counterpart. */
int cmf_keyed_level_order(int64_t n, const int32_t* downstream_idx, const uint8_t* key,
                          int32_t* out_perm, int32_t* out_inv, int32_t* out_dn,
                          int64_t* out_num_levels, int64_t level_ptr_capacity,
                          int64_t* out_level_ptr);

/* -------------------------------------------------------------- network */
/* Build the device network context from host copies of the indices.
 * `downstream_idx` must satisfy downstream_idx[i] > i for non-mouths and
 * == i at mouths.  Bifurcation indices may be NULL when num_paths == 0.
 * `num_trials` sizes the scratch.  Allocates device memory once. */
int cmf_network_create(int64_t num_catchments, const int32_t* downstream_idx,
                       int64_t num_paths, const int32_t* bifurcation_catchment_idx,
                       const int32_t* bifurcation_downstream_idx, int64_t num_trials,
                       cmf_network** out);
/* The same, with the catchments partitioned into num_groups groups that each
 * advance with their own CFL sub-step count and time step (independent basins
 * of uneven sizes in one launch per pass).  Group g is the internal catchments
 * [group_ptr[g], group_ptr[g + 1]) (host array, num_groups + 1 entries, from 0
 * to num_catchments, no empty group); no downstream link or bifurcation path
 * may cross groups.  One trial only.  With a grouped network, time_step and
 * max_sub_steps hold one entry per group, and with the log module
 * log_sums[k] holds num_groups * log_buffer_size entries (group g's sub-step s
 * at g * log_buffer_size + s).  Only cmf_hip_advance and cmf_hip_substep take
 * grouped networks; the per-op entries return CMF_ERR_UNSUPPORTED.  No
 * reference counterpart: the reference's trials share one time step
 * (cmfgpu/phys/triton/outflow.py:454) and its ranks one sub-step count
 * (cmfgpu/models/cama_flood_model.py:119-120). */
int cmf_network_create_grouped(int64_t num_catchments, const int32_t* downstream_idx,
                               int64_t num_paths, const int32_t* bifurcation_catchment_idx,
                               const int32_t* bifurcation_downstream_idx, int64_t num_trials,
                               int64_t num_groups, const int64_t* group_ptr, cmf_network** out);
int64_t cmf_network_num_groups(const cmf_network* net);
/* Block sizes of the fused schedule's kernels, indexed by CMF_TUNE_* (count
 * entries; default 256 each).  Grouped networks keep 256 for the catchment
 * passes.  Setting them drops the graphs cached by cmf_hip_advance_graph.
 * Chosen by cmfgpu_amd.autotune; no reference counterpart (the reference
 * passes a fixed BLOCK_SIZE, cmfgpu/phys/specs.py:10-26). */
int cmf_network_set_block_sizes(cmf_network* net, const int32_t* sizes, int32_t count);
int cmf_network_get_block_sizes(const cmf_network* net, int32_t* sizes, int32_t count);
int cmf_network_destroy(cmf_network* net);
/* 1 when every upstream range is the identity (breadth-first layout). */
int cmf_network_is_contiguous(const cmf_network* net);

/* Schedule of the fused sub-step (all give the same bits; no counterpart in
 * the reference).  THREE_PASS (the default): an outflow pass (+ paths) that
 * also leaves each catchment's storage limiter as its own outgoing term
 * fixes it, a fix-up pass that forms in full the limiters of path ends and
 * of catchments an upstream catchment flows back into (flagged by the
 * outflow pass) and of reservoir catchments (whose release replaces the
 * outflow pass's flows), and the stage pass.  LIMITER_PASS:
 * outflow (+ paths), limiter and stage passes, each over the whole network.
 * BLOCK_LOCAL (one trial,
 * ungrouped, contiguous upstream ranges, no reservoirs, 10 flood levels, 5
 * path levels, not taped; otherwise the three passes run): one pass in which
 * each block of 512 catchments completes the limiters and stages it holds
 * whole, then limiter and stage passes over the catchments left (lists fixed
 * at network creation).  Measured slower on MI355X than THREE_PASS; kept as
 * an option. */
enum { CMF_SCHEDULE_THREE_PASS = 0, CMF_SCHEDULE_BLOCK_LOCAL = 1, CMF_SCHEDULE_LIMITER_PASS = 2 };
int cmf_network_set_schedule(cmf_network* net, int32_t schedule);
/* The block-local schedule's block size (0: not available for this network)
 * and the sizes of its limiter and stage lists. */
int cmf_network_local_counts(const cmf_network* net, int64_t* block, int64_t* lim_count, int64_t* stage_count);

/* ------------------------------------------------- per-op entry points ---
 * Each replaces one reference launcher (the .cu files of cmfgpu/phys/cuda, routed at
 * cmfgpu/phys/cuda/__init__.py:68-155) and one Triton kernel; the fields read
 * and written are those of the matching KernelSpec.  `stream` is a
 * hipStream_t (NULL = default stream). */

/* compute_outflow: cuda/outflow.cu:272-298, triton/outflow.py:15-315 (+:386-717 batched) */
int cmf_hip_outflow(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_inflow: cuda/outflow.cu:389-399, triton/outflow.py:318-383 (+:720-794) */
int cmf_hip_inflow(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_flood_stage: cuda/storage.cu:318-334, triton/storage.py:16-193 */
int cmf_hip_flood_stage(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_flood_stage_log: cuda/storage.cu:348-374, triton/storage.py:196-425 */
int cmf_hip_flood_stage_log(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_adaptive_time_step: cuda/adaptive_time.cu:50-56, triton/adaptive_time.py:13-103.
 * Writes max over catchments of the CFL sub-step count into max_sub_steps
 * (atomic max, as the reference; per trial with CMF_PER_TRIAL_DT). */
int cmf_hip_adaptive_time_step(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_bifurcation_outflow: cuda/bifurcation.cu:158-169, triton/bifurcation.py:13-104 */
int cmf_hip_bifurcation_outflow(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_bifurcation_inflow: cuda/bifurcation.cu:201-208, triton/bifurcation.py:106-151 */
int cmf_hip_bifurcation_inflow(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_levee_stage: cuda/levee.cu:343-353, triton/levee.py:62-304 */
int cmf_hip_levee_stage(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_levee_stage_log: cuda/levee.cu:382-395, triton/levee.py:307-562 */
int cmf_hip_levee_stage_log(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_levee_bifurcation_outflow: cuda/levee.cu:424-437, triton/levee.py:565-684 */
int cmf_hip_levee_bifurcation_outflow(cmf_network* net, const cmf_fields* f, void* stream);
/* compute_reservoir_outflow: cuda/reservoir.cu:96-106, triton/reservoir.py:17-184 */
int cmf_hip_reservoir_outflow(cmf_network* net, const cmf_fields* f, void* stream);

/* No counterpart in the reference: compute_flood_stage_log (the results of
 * cmf_hip_flood_stage_log, bit for bit) for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  It takes no network context and calls
 * no hipMalloc, hipFree or allocating library routine: the FP64 per-block
 * partial sums live in a static device array of the library (5.8 MB per
 * GPU, held until the process ends), which limits num_catchments *
 * num_trials to 2^24 (larger inputs return CMF_ERR_UNSUPPORTED).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (concurrent
 * streams, threads) race and must be serialised by the caller. */
int cmf_hip_flood_stage_log_noalloc(const cmf_fields* f, void* stream);

/* ------------------------------------------------------ fused schedule ---
 * One sub-step of CaMaFlood.step_advance (cmfgpu/models/cama_flood_model.py:
 * 134-162) with the ops fused into the fewest passes the dependencies allow:
 * outflow, [reservoir], [bifurcation outflow], limiter, [bifurcation inflow],
 * inflow + flood stage [+ levee stage] [+ log].  outgoing_storage is never
 * materialised (it is zero after every sub-step in the reference too).
 * `substep` is the sub-step index (log slot); with CMF_PER_TRIAL_DT a trial
 * whose count max_sub_steps[t] is <= substep is left unchanged.  Sub-steps
 * after the first of an outer step follow a stage pass, which leaves zero
 * protected storage in every catchment without a levee: they read
 * protected_storage only for levee catchments.
 * `store_inflow` != 0 marks the last sub-step of the outer step: it writes
 * river_inflow / flood_inflow (the reference leaves the last sub-step's
 * inflow there) and the fields no later pass reads, which the reference
 * rewrites every sub-step and which therefore hold the last sub-step's values
 * after it: water_surface_elevation, protected_water_surface_elevation,
 * total_storage, limit_rate, and flood_fraction / protected_depth of
 * catchments without a levee (every sub-step when a separate levee pass
 * reads them). */
int cmf_hip_substep(cmf_network* net, const cmf_fields* f, int32_t substep,
                    int32_t store_inflow, void* stream);
/* cmf_hip_substep with each kernel launch bracketed by HIP events on
 * `stream`; synchronises, then writes the launch durations (ms) in launch
 * order to out_ms and the kernel names (static strings) to out_names (may be
 * NULL).  Instrumentation for measurement; no reference counterpart. */
int cmf_hip_substep_timed(cmf_network* net, const cmf_fields* f, int32_t substep,
                          int32_t store_inflow, void* stream, double* out_ms,
                          const char** out_names, int32_t capacity, int32_t* out_count);
/* One outer step: CFL (with CMF_HAS_ADAPTIVE_TIME), one host read of the
 * sub-step count, time_step = outer / count, then `count` fused sub-steps.
 * `fixed_count` is used without adaptive time.  The count used is written to
 * *out_count (the maximum over trials with CMF_PER_TRIAL_DT). */
int cmf_hip_advance(cmf_network* net, const cmf_fields* f, double outer_seconds,
                    int32_t fixed_count, int32_t* out_count, void* stream);
/* ----------------------------------------------------- forcing, output */
/* Catchment runoff from a gridded field: out[r] = scale * sum_k weights[k] *
 * grid[indices[k]] over k in [indptr[r], indptr[r + 1]) (sums in double, in
 * index order), out in the profile's REAL.  The matrix is the reference's
 * runoff mapping (catchments x grid cells, area weights; the arrays of the
 * mapping .npz read at cmfgpu/params/export_bin.py:963-975), applied where
 * the reference applies it in hydroforge's shard_forcing
 * (scripts/run_daily_bin.py:114-119).  With `out_index` (device, one entry
 * per row; NULL: out[r]) row r lands in out[out_index[r]]: the rows can stay
 * in the mapping's own (spatially ordered) catchment order while the results
 * go to the catchments' places in the device layout. */
int cmf_hip_map_runoff(int64_t rows, const int64_t* indptr, const int32_t* indices, const float* weights,
                       const float* grid, double scale, const int32_t* out_index, int32_t precision, void* out,
                       void* stream);
/* With a non-NULL accumulator (double, one entry per catchment element) the
 * fused stage pass adds (river_outflow + flood_outflow) * dt every sub-step:
 * the time integral behind the reference's "mean" statistics of
 * total_outflow (cmfgpu/modules/base.py:412-419, scripts/run_daily_bin.py:28).
 * NULL turns it off. */
int cmf_network_set_outflow_accumulator(cmf_network* net, double* accumulator);

/* ------------------------------------------------------------- adjoint */
/* Reverse mode of one fused sub-step of the base physics (outflow, limiter
 * and inflow sums, storage update, flood stage; cmfgpu/phys/triton/outflow.py
 * :15-383 and storage.py:16-193) and of the bifurcation paths
 * (cmfgpu/phys/triton/bifurcation.py:13-151): given the adjoint of the state after the
 * sub-step (adj_out) it writes the adjoint of the state before it (adj_in)
 * and adds d/d(river_manning) and d/d(river_height) into the gradient
 * accumulators, with the time step held at its forward value.  State order:
 * river_outflow, flood_outflow, river/flood cross-section depth, flood
 * cross-section area, river/flood storage, river/flood depth.  The state
 * pointers give the sub-step's saved input state (dtypes of the field table:
 * REAL, storages STO); adj_out/adj_in hold REAL, the gradients double.
 * qraw_r/qraw_f/lim (REAL) may give the unlimited flows and limiters the
 * forward sub-step produced (cmf_hip_substep_taped); when NULL they are
 * recomputed from the state.  The field table supplies parameters, runoff and
 * the flood table.  With the bifurcation module open the path state
 * (bifurcation_outflow, bifurcation_cross_section_depth: REAL [P, levels],
 * at most 8 levels) is part of the state: its saved input values and its
 * adjoints after (adj_out_path) and before (adj_in_path) the sub-step; with
 * the levee module open (cmfgpu/phys/triton/levee.py:62-304, and the path
 * variant :565-684) so are the protected storage and depth, and the field
 * table must hold levee_index; with the reservoir module (the kinematic flow
 * upstream of dams, outflow.py:251-272, and the release,
 * reservoir.py:17-184) the reservoirs' gathered inflow is.  The protected
 * storage of catchments without a levee is taken as zero (what every stage
 * pass leaves there).  With CMF_PRECISE_SURFACE the reverse re-forms the
 * surfaces as the forward did (in double), so its branches follow the
 * forward's.  Trials, groups and per-trial parameters return
 * CMF_ERR_UNSUPPORTED.  No reference counterpart (the reference has no
 * differentiation). */
typedef struct cmf_adjoint {
    const void* river_outflow;
    const void* flood_outflow;
    const void* river_cross_section_depth;
    const void* flood_cross_section_depth;
    const void* flood_cross_section_area;
    const void* river_storage;
    const void* flood_storage;
    const void* river_depth;
    const void* flood_depth;
    const void* qraw_r;
    const void* qraw_f;
    const void* lim;
    const void* adj_out[9];
    void* adj_in[9];
    double* grad_river_manning;
    double* grad_river_height;
    double time_step;
    int32_t mode;        /* CMF_ADJ_* bits, 0 for a standalone sub-step */
    int32_t reserved;
    const void* bifurcation_outflow;
    const void* bifurcation_cross_section_depth;
    const void* adj_out_path[2];
    void* adj_in_path[2];
    const void* path_terms;   /* the forward's path intermediates (cmf_tape), or NULL */
    /* levees: protected_storage (STO) and protected_depth (REAL) [N] join the
     * state, with their adjoints after (adj_out_levee) and before
     * (adj_in_levee) the sub-step */
    const void* protected_storage;
    const void* protected_depth;
    const void* adj_out_levee[2];
    void* adj_in_levee[2];
    /* reservoirs: reservoir_total_inflow (STO [N], the inflow each reservoir
     * gathered in the previous sub-step) joins the state; needs the taped
     * intermediates (qraw_r, qraw_f, lim) */
    const void* reservoir_total_inflow;
    const void* adj_out_reservoir;
    void* adj_in_reservoir;
} cmf_adjoint;
/* mode bits for a sweep over consecutive sub-steps on one network:
 * DEFER leaves the last pass (the gather of downstream water-surface
 * adjoints) pending in the network scratch, PENDING folds the pending gather
 * of the previous call (the later sub-step) into this call, FLUSH runs only
 * the pending gather (into adj_in and the height gradient) after the last
 * call.  Results equal those of standalone calls. */
enum { CMF_ADJ_DEFER = 1, CMF_ADJ_PENDING = 2, CMF_ADJ_FLUSH = 4 };
int cmf_hip_substep_adjoint(cmf_network* net, const cmf_fields* fields, const cmf_adjoint* adjoint,
                            void* stream);

/* One fused sub-step (as cmf_hip_substep) that records what the adjoint needs:
 * its input state (same nine fields and order as cmf_adjoint) and the
 * unlimited flows and limiters it forms; with the bifurcation module open
 * also the input path state [P, levels], with the levee module the input
 * protected storage and depth, with the reservoir module the reservoirs'
 * gathered inflow.  One trial. */
typedef struct cmf_tape {
    void* river_outflow;
    void* flood_outflow;
    void* river_cross_section_depth;
    void* flood_cross_section_depth;
    void* flood_cross_section_area;
    void* river_storage;
    void* flood_storage;
    void* river_depth;
    void* flood_depth;
    void* qraw_r;
    void* qraw_f;
    void* lim;
    void* bifurcation_outflow;
    void* bifurcation_cross_section_depth;
    /* optional, REAL [P (LB + 3)]: the limited level flows [P, LB], their
     * sums, and the outgoing terms at the upstream and downstream ends, as the
     * sub-step formed them (NULL: the adjoint recomputes them) */
    void* path_terms;
    /* with levees: the input protected storage (STO) and depth (REAL) */
    void* protected_storage;
    void* protected_depth;
    /* with reservoirs: the input reservoir_total_inflow (STO) */
    void* reservoir_total_inflow;
} cmf_tape;
int cmf_hip_substep_taped(cmf_network* net, const cmf_fields* fields, int32_t substep, int32_t store_inflow,
                          const cmf_tape* tape, void* stream);
/* One outer step of `count` taped sub-steps (dt = outer_seconds / count,
 * written to time_step; tapes[k] records sub-step k), and the reverse sweep
 * of `count` sub-steps (adjoints[k] for sub-step k, run from the last to the
 * first): the per-outer-step loops of the adjoint driver in one call each. */
int cmf_hip_advance_taped(cmf_network* net, const cmf_fields* fields, double outer_seconds, int32_t count,
                          const cmf_tape* tapes, void* stream);
int cmf_hip_adjoint_steps(cmf_network* net, const cmf_fields* fields, const cmf_adjoint* adjoints,
                          int32_t count, void* stream);

/* cmf_hip_advance with the sub-step loop replayed from a HIP graph: the loop
 * for a given sub-step count is captured once (on a private stream; kernel
 * nodes only) and relaunched on `stream` in later outer steps while the field
 * table holds the same pointers and sizes (a changed table drops the cached
 * graphs).  Results are identical to cmf_hip_advance.  No reference
 * counterpart (the reference launches every op from Python each sub-step,
 * cmfgpu/models/cama_flood_model.py:134-162). */
int cmf_hip_advance_graph(cmf_network* net, const cmf_fields* fields, double outer_seconds,
                          int32_t fixed_count, int32_t* out_count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CMF_HIP_H */
