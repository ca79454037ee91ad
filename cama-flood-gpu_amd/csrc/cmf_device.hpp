// Device-side building blocks: IEEE min/max, cube root / exp / log built from
// basic operations, the element view of a cmf_fields table, and the wave64
// reductions.  Every floating-point function here uses only correctly rounded
// IEEE operations (with contraction off), so a value computed on the GPU is
// the value any IEEE host computes with the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cmf_internal.hpp"

namespace cmf {

// ------------------------------------------------------------ min / max --
// IEEE minNum / maxNum: a NaN operand loses (tl.minimum / tl.maximum).
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return (a < b || b != b) ? a : b; }
template <class T> __device__ __forceinline__ T vmax(T a, T b) { return (a > b || b != b) ? a : b; }
template <class T> __device__ __forceinline__ T vclamp(T x, T lo, T hi) { return vmin(vmax(x, lo), hi); }

// ------------------------------------------------------------ cube root --
// Exponent-thirds guess on the double bit pattern (relative error < 7%),
// then Halley steps y <- y (y^3 + 2x) / (2y^3 + x): two steps leave a
// relative error near 1e-11 before rounding to float, four reach double
// precision.
__device__ __forceinline__ double cbrt_pos_d(double x, int steps) {
    double y = __longlong_as_double(
        (long long)((unsigned long long)__double_as_longlong(x) / 3ull + 0x2AA0000000000000ull));
    for (int k = 0; k < steps; ++k) {
        const double t = y * y * y;
        y = y * ((t + x + x) / (t + t + x));
    }
    return y;
}
__device__ __forceinline__ double cbrt_any_d(double x, int steps) {
    if (x != x || x == 0.0 || __builtin_isinf(x)) return x;
    if (x < 0.0) return -cbrt_pos_d(-x, steps);
    if (x < 1e-290) return cbrt_pos_d(x * 0x1p300, steps) * 0x1p-100;
    return cbrt_pos_d(x, steps);
}
// Float cube root without a division: Newton steps on the reciprocal cube
// root, r <- r (4 - a r^3) / 3, from a bit-pattern guess, then a r^2.  Every
// float is a normal double, and after four steps the result rounds to the
// correctly rounded float cube root for every finite float.
__device__ __forceinline__ float dcbrt(float x) {
    if (x != x || x == 0.0f || __builtin_isinf(x)) return x;
    const double a = x < 0.0f ? -(double)x : (double)x;
    double r = __longlong_as_double(
        (long long)(0x553EF0FF289DD796ull - (unsigned long long)__double_as_longlong(a) / 3ull));
    for (int k = 0; k < 4; ++k) {
        const double r3 = r * r * r;
        r = r * ((4.0 - a * r3) * (1.0 / 3.0));
    }
    const double y = a * r * r;
    return (float)(x < 0.0f ? -y : y);
}
__device__ __forceinline__ double dcbrt(double x) { return cbrt_any_d(x, 4); }
// x^(-1/3) of a positive normal double, without a division, to a few units
// in the last place (four Newton steps from the bit-pattern guess: relative
// error below 1e-15): for derivative terms, where powers of it give
// x^(-4/3) and x^(-7/3) by multiplications.
__device__ __forceinline__ double rcbrt_pos_d(double x) {
    double r = __longlong_as_double(
        (long long)(0x553EF0FF289DD796ull - (unsigned long long)__double_as_longlong(x) / 3ull));
    for (int k = 0; k < 4; ++k) {
        const double r3 = r * r * r;
        r = r * ((4.0 - x * r3) * (1.0 / 3.0));
    }
    return r;
}

// ------------------------------------------------------- log / exp -------
// log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1),
// odd series to s^27.  exp: y = k ln2 + r, Taylor to r^22, times 2^k.
__device__ __forceinline__ double log_core_d(double x) {
    if (x != x || x < 0.0) return __longlong_as_double(0x7ff8000000000000ll);
    if (x == 0.0) return -__builtin_inf();
    if (__builtin_isinf(x)) return x;
    int escale = 0;
    if (x < 1e-300) { x *= 0x1p200; escale = -200; }
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 27.0;
    for (int k = 25; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
    const double lnm = 2.0 * s * p;
    const double LN2 = 0.6931471805599453;
    return (double)(e + escale) * LN2 + lnm;
}
__device__ __forceinline__ double exp_core_d(double y) {
    if (y != y) return y;
    if (y > 709.0) return __builtin_inf();
    if (y < -745.0) return 0.0;
    const double LN2 = 0.6931471805599453;
    const double kf = floor(y / LN2 + 0.5);
    const double r = y - kf * LN2;
    double p = 1.0;
    for (int k = 22; k >= 1; --k) p = 1.0 + p * r / (double)k;
    const int k = (int)kf;
    const int k1 = k / 2, k2 = k - k1;
    const double s1 = __longlong_as_double((long long)((unsigned long long)(k1 + 1023) << 52));
    const double s2 = __longlong_as_double((long long)((unsigned long long)(k2 + 1023) << 52));
    return p * s1 * s2;
}
__device__ __forceinline__ float dlog(float x) { return (float)log_core_d((double)x); }
__device__ __forceinline__ double dlog(double x) { return log_core_d(x); }
__device__ __forceinline__ float dexp(float x) { return (float)exp_core_d((double)x); }
__device__ __forceinline__ double dexp(double x) { return exp_core_d(x); }

__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }

// ------------------------------------------------------- wave64 reduce ---
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum_double(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------ the view --
// Typed device pointers of a cmf_fields table plus the network scratch.
template <class REAL, class STO>
struct View {
    int64_t N, T, L, P, LB, V, R, logsize;
    int64_t ts, cs, pts, pps;     // element strides: trial / catchment, path trial / path
    uint32_t flags, batched;
    REAL gravity, adaptive_factor, min_kin;
    const REAL* time_step;
    const REAL* outer_time_step;
    const int32_t* current_step;
    int32_t* max_sub_steps;
    const int32_t* dn;
    STO *river_inflow, *flood_inflow;
    REAL *river_outflow, *flood_outflow;
    const REAL *river_manning, *flood_manning;
    REAL *river_depth, *flood_depth, *protected_depth;
    const REAL *river_width, *river_length, *river_height;
    const REAL *elev, *catchment_area, *dist;
    STO *river_storage, *flood_storage, *protected_storage;
    REAL *rcsd, *fcsd, *fcsa, *flood_fraction;
    const REAL *runoff, *table;
    STO *outgoing, *total_storage;
    REAL *wse, *pwse;
    const uint8_t* is_levee;
    const int32_t *bif_ci, *bif_di;
    const REAL *bif_manning, *bif_width, *bif_length, *bif_elev;
    REAL *bif_outflow, *bif_csd;
    STO* gbif;
    REAL* limit_rate;
    const int32_t* levee_idx;
    const int32_t* levee_index;         // catchment -> levee row, -1 if none (optional)
    const REAL *levee_base, *levee_crown, *levee_frac;
    const int32_t* res_idx;
    const REAL *res_cons, *res_emer, *res_adj, *res_qn, *res_qa, *res_qf;
    STO* res_inflow;
    const uint8_t *is_reservoir, *is_dam_up, *is_dam_related;
    double* log_sums[CMF_LOG_COUNT];
    // network
    // Upstream catchments of j: the range up_range(j) directly (up_idx ==
    // nullptr: contiguous ranges, where a mouth's range may end with the mouth
    // itself) or through up_idx.  Where the ranges follow their owners (and
    // for up_idx) they are [up_ptr[j], up_ptr[j+1]); otherwise up_rng[j]
    // holds {start, end}.
    const int32_t *up_ptr, *up_idx;
    const int2* up_rng;
    // bifurcation events per catchment: ev_code = 2 path + role (0: the
    // catchment is the path's upstream end, 1: its downstream end), ev_other
    // the path's other end
    const int32_t *ev_ptr, *ev_code, *ev_other;
    // the first event of each catchment in one 8-byte record for the limiter
    // pass: {code, other end}; code -1: no event; other end -(other + 1):
    // more events follow (the ev_ptr list from its second entry)
    const int2* ev1;
    // 1 byte per catchment: 1 where it has an event (ev_ptr[j] != ev_ptr[j + 1])
    const uint8_t* ev_flag;
    // scratch
    REAL *pos_term, *neg_term, *qraw_r, *qraw_f, *lim;
    REAL *path_pos, *path_neg, *path_net, *path_raw;
    double* partials;
    // provisional-limiter schedule: one bit per element, set where the
    // outflow pass could not form the limiter exactly (an upstream catchment
    // flows back into it), cleared by the fix-up pass; and each path's two
    // outgoing terms at its ends' event slots (eterm[evel(t, pev[p].x)]: the
    // upstream end's, .y the downstream end's; eterm == nullptr: not kept)
    uint32_t* bmask;
    REAL* eterm;
    const int2* pev;
    // groups (G > 0): independent sets of basins, each with its own sub-step
    // count and time step (slot = group of the catchment).  Catchment kernels
    // then run one block per record of blk_rec: {first catchment, end, block
    // table entry (the log partial slot), group}, a single group's
    // catchments [first, end).  `ordered`: the records of this launch belong
    // to groups still sub-stepping only (the table ordered by group count).
    int64_t G;
    const int32_t* cgroup;
    const int4* blk_rec;
    bool ordered;
    // taping (cmf_hip_substep_taped): the outflow pass also writes the
    // sub-step's input state here (nullptr: no tape)
    REAL *tape_q_r, *tape_q_f, *tape_rcsd, *tape_fcsd, *tape_fcsa, *tape_rd, *tape_fd;
    STO *tape_rs, *tape_fs;
    REAL *tape_bq, *tape_bcsd;    // the path pass: input path flows and depths
    REAL* tape_bqlim;             // and its limited level flows
    STO* tape_ps;                 // the outflow pass: input protected storage
    REAL* tape_pd;                // and depth (levee module)
    STO* tape_ri;                 // and the reservoirs' gathered inflow (reservoir module)
    // output statistics: sum over sub-steps of (river + flood outflow) * dt
    // per element (nullptr: off)
    double* acc_outflow;

    __device__ __forceinline__ int64_t el(int64_t t, int64_t i) const { return t * ts + i * cs; }
    __device__ __forceinline__ int64_t pel(int64_t t, int64_t p) const { return t * pts + p * pps; }
    // element of event slot k (2 P slots per trial, laid out as the paths)
    __device__ __forceinline__ int64_t evel(int64_t t, int64_t k) const {
        return t * (pts == 1 ? 1 : 2 * P) + k * pps;
    }
    __host__ __device__ __forceinline__ bool has(uint32_t bit) const { return (flags & bit) != 0; }
    __device__ __forceinline__ int2 up_range(int64_t j) const {
        return up_rng ? up_rng[j] : make_int2(up_ptr[j], up_ptr[j + 1]);
    }
    __host__ __device__ __forceinline__ bool bat(uint32_t bit) const { return (batched & bit) != 0; }
    // parameter read: per-trial element when batched, shared otherwise
    template <class X>
    __device__ __forceinline__ X prm(const X* p, uint32_t bit, int64_t e, int64_t i) const {
        return bat(bit) ? p[e] : p[i];
    }
    // time-step slot of (trial t, catchment c): its group, its trial with
    // per-trial dt, else the shared slot 0
    __device__ __forceinline__ int64_t slot(int64_t t, int64_t c) const {
        return G > 0 ? (int64_t)cgroup[c] : (has(CMF_PER_TRIAL_DT) ? t : 0);
    }
    __device__ __forceinline__ bool per_slot() const { return G > 0 || has(CMF_PER_TRIAL_DT); }
    __device__ __forceinline__ REAL dt(int64_t t, int64_t c) const { return time_step[slot(t, c)]; }
    // With per-trial or per-group sub-step counts a slot past its count sits out.
    __device__ __forceinline__ bool active(int64_t t, int64_t c, int substep) const {
        if (substep < 0 || !per_slot()) return true;
        return substep < max_sub_steps[slot(t, c)];
    }
    // the same in a catchment kernel, whose ordered launches hold active
    // groups only
    __device__ __forceinline__ bool active_in_block(int64_t t, int64_t c, int substep) const {
        if (G > 0 && ordered) return true;
        return active(t, c, substep);
    }
    // the last sub-step of (t, c)'s slot, or `shared_last` with one count for all
    __device__ __forceinline__ bool last_substep(int64_t t, int64_t c, int substep, int shared_last) const {
        return (per_slot() && substep >= 0) ? (substep == max_sub_steps[slot(t, c)] - 1) : (shared_last != 0);
    }
    // linear thread index -> (trial, item); `inner` when trials are the
    // fastest-varying dimension of the layout
    __device__ __forceinline__ void split(int64_t g, int64_t n, bool inner, int64_t& t, int64_t& i) const {
        if (T == 1) { t = 0; i = g; }
        else if (inner) { i = g / T; t = g - i * T; }
        else { t = g / n; i = g - t * n; }
    }
    __device__ __forceinline__ void split_c(int64_t g, int64_t& t, int64_t& i) const {
        split(g, N, ts == 1, t, i);
    }
    // log partial slot of this block (its block table entry when grouped)
    __device__ __forceinline__ int64_t partial_slot() const {
        return G > 0 ? (int64_t)blk_rec[blockIdx.x].z : (int64_t)blockIdx.x;
    }
    // catchment of this thread in a catchment kernel (linear index `g` of the
    // plain grid, or the group block table); false past the end
    __device__ __forceinline__ bool catchment_of(int64_t g, int64_t& t, int64_t& i) const {
        if (G > 0) {
            const int4 r = blk_rec[blockIdx.x];
            const int64_t j = (int64_t)r.x + threadIdx.x;
            t = 0;
            i = j;
            return j < r.y;
        }
        if (g >= N * T) return false;
        split_c(g, t, i);
        return true;
    }
    // the same for block `blk` of the catchment grid
    __device__ __forceinline__ bool catchment_in(int64_t blk, int64_t& t, int64_t& i) const {
        if (G > 0) {
            const int4 r = blk_rec[blk];
            const int64_t j = (int64_t)r.x + threadIdx.x;
            t = 0;
            i = j;
            return j < r.y;
        }
        return catchment_of(blk * blockDim.x + threadIdx.x, t, i);
    }
    __device__ __forceinline__ void split_p(int64_t g, int64_t& t, int64_t& p) const {
        split(g, P, pts == 1, t, p);
    }
    // element index -> (trial, catchment) (one trial: the element is the
    // catchment, unit stride)
    __device__ __forceinline__ void elem_split(int64_t e, int64_t& t, int64_t& i) const {
        if (T == 1) { t = 0; i = e; }
        else if (ts == 1) { i = e / cs; t = e - i * cs; }
        else { t = e / ts; i = e - t * ts; }
    }
    // entry g of a per-catchment list run once per trial -> (trial, entry)
    __device__ __forceinline__ void list_split(int64_t g, int64_t& t, int64_t& k) const {
        if (T == 1) { t = 0; k = g; }
        else { k = g / T; t = g - k * T; }
    }
    // GroupView: the group of a launch's catchments is that of its block record
    static constexpr bool kGrouped = false;
    // whether a catchment's levee and event tests read the byte flags
    // (is_levee, ev_flag) before any index: per-lane catchments, where a
    // byte is a quarter of the index; a view whose catchment index is the
    // same in every lane reads the index itself (one scalar load)
    static constexpr bool kByteFlags = true;
};

// The view of a one-trial, ungrouped run without per-trial parameters or
// time steps: every per-trial and per-group index folded away at compile
// time (element = catchment, shared parameters, the one time-step slot).
// Same layout as View; kernels templated on the view type take either.
// DEVLAST: whether the sub-step is the outer step's last is read on the
// device (max_sub_steps[0]), for the first sub-step, which is enqueued before
// the host has read the count (every outer step has at least one sub-step).
template <class REAL, class STO, bool DEVLAST = false>
struct PlainView : View<REAL, STO> {
    __device__ __forceinline__ int64_t el(int64_t, int64_t i) const { return i; }
    __device__ __forceinline__ int64_t pel(int64_t, int64_t p) const { return p; }
    __device__ __forceinline__ int64_t evel(int64_t, int64_t k) const { return k; }
    __host__ __device__ __forceinline__ bool bat(uint32_t) const { return false; }
    template <class X>
    __device__ __forceinline__ X prm(const X* p, uint32_t, int64_t, int64_t i) const { return p[i]; }
    __device__ __forceinline__ int64_t slot(int64_t, int64_t) const { return 0; }
    __device__ __forceinline__ bool per_slot() const { return false; }
    __device__ __forceinline__ REAL dt(int64_t, int64_t) const { return this->time_step[0]; }
    __device__ __forceinline__ bool active(int64_t, int64_t, int) const { return true; }
    __device__ __forceinline__ bool active_in_block(int64_t, int64_t, int) const { return true; }
    __device__ __forceinline__ bool last_substep(int64_t, int64_t, int substep, int shared_last) const {
        return (DEVLAST && substep >= 0) ? substep == this->max_sub_steps[0] - 1 : shared_last != 0;
    }
    __device__ __forceinline__ void split_c(int64_t g, int64_t& t, int64_t& i) const { t = 0; i = g; }
    __device__ __forceinline__ void split_p(int64_t g, int64_t& t, int64_t& p) const { t = 0; p = g; }
    __device__ __forceinline__ void elem_split(int64_t e, int64_t& t, int64_t& i) const { t = 0; i = e; }
    __device__ __forceinline__ void list_split(int64_t g, int64_t& t, int64_t& k) const { t = 0; k = g; }
    __device__ __forceinline__ int64_t partial_slot() const { return (int64_t)blockIdx.x; }
    __device__ __forceinline__ bool catchment_of(int64_t g, int64_t& t, int64_t& i) const {
        t = 0;
        i = g;
        return g < this->N;
    }
    __device__ __forceinline__ bool catchment_in(int64_t blk, int64_t& t, int64_t& i) const {
        return catchment_of(blk * blockDim.x + threadIdx.x, t, i);
    }
};

// The view of a trial-inner run (element = catchment * T + member, path
// element = path * T + member) whose member count T is a power of two,
// ungrouped: element indices are a shift and an add, and splitting a thread
// index into (member, catchment) a shift and a mask instead of a 64-bit
// division.  WAVE (T a multiple of 64): a wave's 64 lanes are members of one
// catchment (or path), so its index is the same in every lane; it is formed
// once per wave (readfirstlane, exact because blocks are whole waves and T
// is a multiple of the wave), and every catchment-indexed address, shared
// parameter and upstream range becomes a scalar.
template <class REAL, class STO, bool WAVE = true>
struct MemberView : View<REAL, STO> {
    static constexpr bool kByteFlags = !WAVE;
    int log2T;
    __device__ __forceinline__ int64_t el(int64_t t, int64_t i) const { return (i << log2T) + t; }
    __device__ __forceinline__ int64_t pel(int64_t t, int64_t p) const { return (p << log2T) + t; }
    __device__ __forceinline__ int64_t evel(int64_t t, int64_t k) const { return (k << log2T) + t; }
    __device__ __forceinline__ int64_t item_of(int64_t g) const {
        return WAVE ? (int64_t)__builtin_amdgcn_readfirstlane((int)(g >> log2T)) : (g >> log2T);
    }
    __device__ __forceinline__ void split_c(int64_t g, int64_t& t, int64_t& i) const {
        i = item_of(g);
        t = g - (i << log2T);
    }
    __device__ __forceinline__ void split_p(int64_t g, int64_t& t, int64_t& p) const {
        p = item_of(g);
        t = g - (p << log2T);
    }
    // element / list entry -> (member, catchment / entry), per lane (a
    // lane's element need not share the wave's catchment)
    __device__ __forceinline__ void elem_split(int64_t e, int64_t& t, int64_t& i) const {
        i = e >> log2T;
        t = e - (i << log2T);
    }
    __device__ __forceinline__ void list_split(int64_t g, int64_t& t, int64_t& k) const {
        k = g >> log2T;
        t = g - (k << log2T);
    }
    __device__ __forceinline__ int64_t slot(int64_t t, int64_t) const {
        return this->has(CMF_PER_TRIAL_DT) ? t : 0;
    }
    __device__ __forceinline__ bool per_slot() const { return this->has(CMF_PER_TRIAL_DT); }
    __device__ __forceinline__ REAL dt(int64_t t, int64_t c) const { return this->time_step[slot(t, c)]; }
    __device__ __forceinline__ bool active(int64_t t, int64_t c, int substep) const {
        if (substep < 0 || !per_slot()) return true;
        return substep < this->max_sub_steps[slot(t, c)];
    }
    __device__ __forceinline__ bool active_in_block(int64_t t, int64_t c, int substep) const {
        return active(t, c, substep);
    }
    __device__ __forceinline__ bool last_substep(int64_t t, int64_t c, int substep, int shared_last) const {
        return (per_slot() && substep >= 0) ? (substep == this->max_sub_steps[slot(t, c)] - 1) : (shared_last != 0);
    }
    __device__ __forceinline__ int64_t partial_slot() const { return (int64_t)blockIdx.x; }
    __device__ __forceinline__ bool catchment_of(int64_t g, int64_t& t, int64_t& i) const {
        if (g >= this->N * this->T) return false;
        split_c(g, t, i);
        return true;
    }
    __device__ __forceinline__ bool catchment_in(int64_t blk, int64_t& t, int64_t& i) const {
        return catchment_of(blk * blockDim.x + threadIdx.x, t, i);
    }
};

// The view of a grouped network with one trial, shared parameters and no
// bifurcation paths, for the catchment kernels, which run one block per
// record of blk_rec.  A group is a union of whole basins, so every catchment
// a thread reaches through the river network (itself, its downstream and
// upstream catchments) is in the group of the block's record: the time-step
// slot, dt and sub-step count are the record's, read once per block instead
// of per catchment, and the per-trial indexing folds away as in PlainView.
template <class REAL, class STO>
struct GroupView : View<REAL, STO> {
    static constexpr bool kGrouped = true;
    __device__ __forceinline__ int4 rec() const { return this->blk_rec[blockIdx.x]; }
    __device__ __forceinline__ int64_t el(int64_t, int64_t i) const { return i; }
    __device__ __forceinline__ int64_t pel(int64_t, int64_t p) const { return p; }
    __host__ __device__ __forceinline__ bool bat(uint32_t) const { return false; }
    template <class X>
    __device__ __forceinline__ X prm(const X* p, uint32_t, int64_t, int64_t i) const { return p[i]; }
    __device__ __forceinline__ int64_t slot(int64_t, int64_t) const { return (int64_t)rec().w; }
    __device__ __forceinline__ bool per_slot() const { return true; }
    __device__ __forceinline__ REAL dt(int64_t, int64_t) const { return this->time_step[rec().w]; }
    __device__ __forceinline__ bool active(int64_t, int64_t, int substep) const {
        return substep < 0 || substep < this->max_sub_steps[rec().w];
    }
    __device__ __forceinline__ bool active_in_block(int64_t t, int64_t c, int substep) const {
        return this->ordered || active(t, c, substep);
    }
    __device__ __forceinline__ bool last_substep(int64_t, int64_t, int substep, int shared_last) const {
        return substep >= 0 ? substep == this->max_sub_steps[rec().w] - 1 : shared_last != 0;
    }
    __device__ __forceinline__ void split_c(int64_t g, int64_t& t, int64_t& i) const { t = 0; i = g; }
    __device__ __forceinline__ void split_p(int64_t g, int64_t& t, int64_t& p) const { t = 0; p = g; }
    __device__ __forceinline__ int64_t partial_slot() const { return (int64_t)rec().z; }
    __device__ __forceinline__ bool catchment_of(int64_t, int64_t& t, int64_t& i) const {
        const int4 r = rec();
        t = 0;
        i = (int64_t)r.x + threadIdx.x;
        return i < r.y;
    }
    // block b of a launch is record b, whatever renumbering the caller applied
    // (the record must be the one slot() reads)
    __device__ __forceinline__ bool catchment_in(int64_t, int64_t& t, int64_t& i) const {
        return catchment_of(0, t, i);
    }
};

// CMF_PRECISE_SURFACE in the FP32 profile: water surfaces formed in double
// (the FP64 profile forms them in double anyway)
template <class REAL, class STO>
__device__ __forceinline__ bool precise_surfaces(const View<REAL, STO>& v) {
    return sizeof(REAL) < sizeof(double) && v.has(CMF_PRECISE_SURFACE);
}

}  // namespace cmf
