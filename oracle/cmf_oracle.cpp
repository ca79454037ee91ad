// CPU oracle for the CaMa-Flood per-sub-step hydrodynamic path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
// check and the CPU-baseline leg of the benchmark load it.  Nothing in the
// product path may link, load or call it.
//
// It restates, in plain sequential C++, the arithmetic of the reference's
// Triton kernels (the primary implementation) in their exact operation order:
//   cmfgpu/phys/triton/outflow.py      compute_outflow_kernel   :15-315
//                                      compute_inflow_kernel    :318-383
//   cmfgpu/phys/triton/storage.py      compute_flood_stage_kernel     :16-193
//                                      compute_flood_stage_log_kernel :196-425
//   cmfgpu/phys/triton/adaptive_time.py compute_adaptive_time_step_kernel :13-50
//   cmfgpu/phys/triton/bifurcation.py  outflow :13-104, inflow :106-151
//   cmfgpu/phys/triton/levee.py        partition :18-59, stage :62-304,
//                                      stage_log :307-562, bifurcation :565-684
//   cmfgpu/phys/triton/reservoir.py    compute_reservoir_outflow_kernel :17-184
// and the sub-step sequence of cmfgpu/models/cama_flood_model.py:104-173.
//
// Precision is a template pair (REAL, STO): (float, double) is the FP32
// profile, (double, double) the FP64 oracle run.  With the PRECISE_SURFACE
// flag the FP32 profile forms water surfaces in double (see outflow()); the
// flag changes nothing in the FP64 profile.
//
// Accumulation order.  The reference accumulates with float atomics, whose
// order is unspecified.  The oracle fixes a canonical order, which is the order
// of a sequential loop over the launching index:
//   * outgoing_storage[j]  = sum over upstream u of j (ascending u) of the
//     negative-flow terms, then j's own positive term (outflow.py:306-314),
//     then bifurcation terms in ascending path order (bifurcation.py:101-104).
//     A reservoir replaces its catchment's outflow terms and the two affected
//     totals are re-summed in the same order (the reference's undo/redo pair,
//     reservoir.py:61-77,172-184, is the same replacement in exact arithmetic).
//   * river_inflow / flood_inflow / reservoir_total_inflow[j] = sum over
//     upstream u of j in ascending u (outflow.py:374-383).
//   * global_bifurcation_outflow[j] = path terms in ascending path order
//     (bifurcation.py:150-151).
//   * water-balance log sums: FP64, sequential.
// Upstream means downstream_idx[u] == j; the network order satisfies
// downstream_idx[i] > i for every non-mouth (merit_map.py:1063-1070).
//
// min / max follow IEEE minNum / maxNum (a NaN operand loses), which is what
// tl.minimum / tl.maximum do.  cbrt, exp and log are evaluated in double from
// basic operations only, so the result is the same on every IEEE machine; the
// reference's HIP path evaluates cbrt as pow(x, 1/3) (triton/utils.py:39-44).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#include <omp.h>

namespace {

// ---------------------------------------------------------------- math ----
template <class T> inline T fmin_(T a, T b) { return (a < b || b != b) ? a : b; }
template <class T> inline T fmax_(T a, T b) { return (a > b || b != b) ? a : b; }
template <class T> inline T clamp_(T x, T lo, T hi) { return fmin_(fmax_(x, lo), hi); }

inline uint64_t dbits(double x) { uint64_t b; std::memcpy(&b, &x, 8); return b; }
inline double dfrom(uint64_t b) { double x; std::memcpy(&x, &b, 8); return x; }

// Cube root of a positive finite double: exponent-thirds initial guess on the
// bit pattern (error below 7%), then Halley steps y <- y (y^3 + 2x) / (2y^3 + x).
inline double cbrt_pos(double x, int halley_steps) {
    double y = dfrom(dbits(x) / 3u + 0x2AA0000000000000ull);
    for (int k = 0; k < halley_steps; ++k) {
        double t = y * y * y;
        y = y * ((t + x + x) / (t + t + x));
    }
    return y;
}
inline double cbrt_any(double x, int steps) {
    if (x != x || x == 0.0 || x == INFINITY || x == -INFINITY) return x;
    if (x < 0.0) return -cbrt_pos(-x, steps);
    if (x < 1e-290) return cbrt_pos(x * 0x1p300, steps) * 0x1p-100;  // keep the guess normal
    return cbrt_pos(x, steps);
}
// Float cube root: Newton steps on the reciprocal cube root,
// r <- r (4 - a r^3) / 3, from a bit-pattern guess, then a r^2 (no division).
// Four steps round to the correctly rounded float cube root for every finite
// float (checked exhaustively against the C library's double cbrt).
inline float cbrt_real(float x) {
    if (x != x || x == 0.0f || x == INFINITY || x == -INFINITY) return x;
    const double a = x < 0.0f ? -(double)x : (double)x;
    double r = dfrom(0x553EF0FF289DD796ull - dbits(a) / 3u);
    for (int k = 0; k < 4; ++k) {
        const double r3 = r * r * r;
        r = r * ((4.0 - a * r3) * (1.0 / 3.0));
    }
    const double y = a * r * r;
    return (float)(x < 0.0f ? -y : y);
}
inline double cbrt_real(double x) { return cbrt_any(x, 4); }

// Natural logarithm: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| < 0.172, series to s^27.
inline double log_core(double x) {
    if (x != x || x < 0.0) return std::numeric_limits<double>::quiet_NaN();
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    int escale = 0;
    if (x < 1e-300) { x *= 0x1p200; escale = -200; }
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = dfrom((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 27.0;
    for (int k = 25; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
    double lnm = 2.0 * s * p;
    const double LN2 = 0.6931471805599453;
    return (double)(e + escale) * LN2 + lnm;
}
// Exponential: y = k ln2 + r, |r| <= ln2/2, Taylor to r^22, scaled by 2^k.
inline double exp_core(double y) {
    if (y != y) return y;
    if (y > 709.0) return INFINITY;
    if (y < -745.0) return 0.0;
    const double LN2 = 0.6931471805599453;
    double kf = std::floor(y / LN2 + 0.5);
    double r = y - kf * LN2;
    double p = 1.0;
    for (int k = 22; k >= 1; --k) p = 1.0 + p * r / (double)k;
    int k = (int)kf;
    // scale in two halves so that 2^k stays representable
    int k1 = k / 2, k2 = k - k1;
    double s1 = dfrom((uint64_t)(k1 + 1023) << 52);
    double s2 = dfrom((uint64_t)(k2 + 1023) << 52);
    return p * s1 * s2;
}
inline float log_real(float x) { return (float)log_core((double)x); }
inline double log_real(double x) { return log_core(x); }
inline float exp_real(float x) { return (float)exp_core((double)x); }
inline double exp_real(double x) { return exp_core(x); }

// ------------------------------------------------------------ field table --
// The ctypes wrapper (oracle/oracle.py) passes a table of pointers indexed by
// these ids, a table of dimensions and a table of scalars.
enum Field {
    F_DOWNSTREAM_IDX = 0,
    F_RIVER_INFLOW, F_FLOOD_INFLOW, F_RIVER_OUTFLOW, F_FLOOD_OUTFLOW,
    F_RIVER_MANNING, F_FLOOD_MANNING, F_RIVER_DEPTH, F_FLOOD_DEPTH,
    F_PROTECTED_DEPTH, F_RIVER_WIDTH, F_RIVER_LENGTH, F_RIVER_HEIGHT,
    F_CATCHMENT_ELEVATION, F_CATCHMENT_AREA, F_DOWNSTREAM_DISTANCE,
    F_RIVER_STORAGE, F_FLOOD_STORAGE, F_PROTECTED_STORAGE,
    F_RIVER_CSD, F_FLOOD_CSD, F_FLOOD_CSA, F_FLOOD_FRACTION, F_RUNOFF,
    F_FLOOD_DEPTH_TABLE, F_OUTGOING_STORAGE, F_TOTAL_STORAGE,
    F_WSE, F_PWSE, F_IS_LEVEE,
    // bifurcation
    F_BIF_CATCHMENT_IDX, F_BIF_DOWNSTREAM_IDX, F_BIF_MANNING, F_BIF_OUTFLOW,
    F_BIF_WIDTH, F_BIF_LENGTH, F_BIF_ELEVATION, F_BIF_CSD,
    F_GLOBAL_BIF_OUTFLOW, F_LIMIT_RATE,
    // levee
    F_LEVEE_CATCHMENT_IDX, F_LEVEE_BASE_HEIGHT, F_LEVEE_CROWN_HEIGHT, F_LEVEE_FRACTION,
    // reservoir
    F_RES_CATCHMENT_IDX, F_CONSERVATION_VOLUME, F_EMERGENCY_VOLUME,
    F_ADJUSTMENT_VOLUME, F_EFFECTIVE_NORMAL_OUTFLOW, F_ADJUSTMENT_OUTFLOW,
    F_FLOOD_CONTROL_OUTFLOW, F_RES_TOTAL_INFLOW, F_IS_RESERVOIR,
    F_IS_DAM_UPSTREAM, F_IS_DAM_RELATED,
    // log: 11 FP64 buffers of log_buffer_size entries
    F_LOG_SUMS,
    F_COUNT
};
enum Dim { D_N = 0, D_L, D_P, D_LB, D_V, D_R, D_LOGSIZE, D_COUNT };
enum Scalar { S_GRAVITY = 0, S_ADAPTIVE_FACTOR, S_MIN_KIN_SLOPE, S_COUNT };
enum Flag {
    HAS_BIFURCATION = 1, HAS_LEVEE = 2, HAS_RESERVOIR = 4, HAS_LOG = 8,
    HAS_TOTAL_STORAGE = 16, HAS_WATER_SURFACE = 32, HAS_PROTECTED_WATER_SURFACE = 64,
    HAS_ADAPTIVE = 128, PRECISE_SURFACE = 512
};
// order of the 11 water-balance sums in F_LOG_SUMS (LogModule field order,
// cmfgpu/modules/log.py:207-326)
enum LogSum {
    L_STORAGE_PRE = 0, L_STORAGE_NEXT, L_STORAGE_NEW, L_INFLOW_ERROR, L_INFLOW,
    L_OUTFLOW, L_STORAGE_STAGE, L_STAGE_ERROR, L_RIVER_STORAGE, L_FLOOD_STORAGE,
    L_FLOOD_AREA, L_COUNT
};

template <class REAL, class STO>
struct Model {
    void* const* p;
    const int64_t* d;
    const double* s;
    int flags;
    // oracle-side bookkeeping of the outgoing-storage terms (canonical order)
    std::vector<REAL> pos_term, neg_term;
    std::vector<int64_t> up_ptr, up_idx;       // upstream lists, ascending u

    template <class X> X* f(int id) const { return static_cast<X*>(p[id]); }
    int64_t N() const { return d[D_N]; }

    void build_upstream() {
        int64_t n = N();
        const int32_t* dn = f<int32_t>(F_DOWNSTREAM_IDX);
        up_ptr.assign(n + 1, 0);
        for (int64_t i = 0; i < n; ++i) if (dn[i] != i) up_ptr[dn[i] + 1]++;
        for (int64_t i = 0; i < n; ++i) up_ptr[i + 1] += up_ptr[i];
        up_idx.assign(up_ptr[n], 0);
        std::vector<int64_t> fill(up_ptr.begin(), up_ptr.end() - 1);
        for (int64_t i = 0; i < n; ++i) if (dn[i] != i) up_idx[fill[dn[i]]++] = i;
        pos_term.assign(n, (REAL)0);
        neg_term.assign(n, (REAL)0);
    }

    // Canonical outgoing total of catchment j from the catchment terms.
    STO outgoing_sum(int64_t j) const {
        STO acc = (STO)0;
        for (int64_t k = up_ptr[j]; k < up_ptr[j + 1]; ++k) acc += (STO)neg_term[up_idx[k]];
        acc += (STO)pos_term[j];
        return acc;
    }

    // ------------------------------------------------- compute_outflow ----
    // triton/outflow.py:15-315
    void outflow(REAL time_step) {
        const int64_t n = N();
        const REAL gravity = (REAL)s[S_GRAVITY];
        const REAL min_kin = (REAL)s[S_MIN_KIN_SLOPE];
        const int32_t* dn_idx = f<int32_t>(F_DOWNSTREAM_IDX);
        STO* river_inflow = f<STO>(F_RIVER_INFLOW);
        STO* flood_inflow = f<STO>(F_FLOOD_INFLOW);
        REAL* river_outflow = f<REAL>(F_RIVER_OUTFLOW);
        REAL* flood_outflow = f<REAL>(F_FLOOD_OUTFLOW);
        const REAL* river_manning = f<REAL>(F_RIVER_MANNING);
        const REAL* flood_manning = f<REAL>(F_FLOOD_MANNING);
        const REAL* river_depth = f<REAL>(F_RIVER_DEPTH);
        const REAL* flood_depth = f<REAL>(F_FLOOD_DEPTH);
        const REAL* protected_depth = f<REAL>(F_PROTECTED_DEPTH);
        const REAL* river_width = f<REAL>(F_RIVER_WIDTH);
        const REAL* river_length = f<REAL>(F_RIVER_LENGTH);
        const REAL* river_height = f<REAL>(F_RIVER_HEIGHT);
        const REAL* elev = f<REAL>(F_CATCHMENT_ELEVATION);
        const REAL* dist = f<REAL>(F_DOWNSTREAM_DISTANCE);
        const STO* river_storage = f<STO>(F_RIVER_STORAGE);
        const STO* flood_storage = f<STO>(F_FLOOD_STORAGE);
        const STO* protected_storage = f<STO>(F_PROTECTED_STORAGE);
        REAL* rcsd = f<REAL>(F_RIVER_CSD);
        REAL* fcsd = f<REAL>(F_FLOOD_CSD);
        REAL* fcsa = f<REAL>(F_FLOOD_CSA);
        STO* gbif = f<STO>(F_GLOBAL_BIF_OUTFLOW);
        STO* total_storage_out = f<STO>(F_TOTAL_STORAGE);
        STO* outgoing = f<STO>(F_OUTGOING_STORAGE);
        REAL* wse_out = f<REAL>(F_WSE);
        REAL* pwse_out = f<REAL>(F_PWSE);
        const uint8_t* is_levee = f<uint8_t>(F_IS_LEVEE);
        const uint8_t* is_dam_up = f<uint8_t>(F_IS_DAM_UPSTREAM);
        const bool has_levee = flags & HAS_LEVEE, has_res = flags & HAS_RESERVOIR;
        const bool has_bif = flags & HAS_BIFURCATION;
        const bool precise = (sizeof(REAL) < sizeof(double)) && (flags & PRECISE_SURFACE);

#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const int64_t dn = dn_idx[i];
            const bool mouth = (dn == i);
            const REAL r_out = river_outflow[i], r_man = river_manning[i];
            const REAL r_dep = river_depth[i], r_wid = river_width[i];
            const REAL r_len = river_length[i], r_hgt = river_height[i];
            const REAL f_out = flood_outflow[i], f_man = flood_manning[i];
            const REAL f_dep = flood_depth[i], p_dep = protected_depth[i];
            const REAL c_elv = elev[i], dn_dist = dist[i];
            const REAL rs = (REAL)river_storage[i];
            const REAL fs = (REAL)flood_storage[i];
            const REAL ps = (REAL)protected_storage[i];
            const REAL r_csd = rcsd[i], f_csd = fcsd[i], f_csa = fcsa[i];

            const REAL total_storage = rs + fs + ps;
            const REAL r_dep_dn = river_depth[dn], r_hgt_dn = river_height[dn], c_elv_dn = elev[dn];
            // wse: water surface; dwse: its drop to the downstream surface;
            // upd_rcsd / fsd: the higher surface above the river bed / the bank
            REAL wse, pwse, dwse, upd_rcsd, fsd;
            if (precise) {
                // in double from the REAL depth, elevation and height (the
                // sums of two or three floats are exact), each difference
                // rounded once: the REAL surfaces (hundreds of metres) lose
                // the centimetres their differences are made of
                const double rel = (double)c_elv - (double)r_hgt;
                const double w = (double)r_dep + rel;
                const double w_dn = mouth ? (double)c_elv : (double)r_dep_dn + ((double)c_elv_dn - (double)r_hgt_dn);
                const double mw = fmax_(w, w_dn);
                wse = (REAL)w;
                pwse = (has_levee && is_levee[i]) ? (REAL)fmin_((double)c_elv + (double)p_dep, w) : wse;
                dwse = (REAL)(w - w_dn);
                upd_rcsd = mouth ? r_dep : (REAL)(mw - rel);
                fsd = (REAL)((mouth ? w : mw) - (double)c_elv);
            } else {
                const REAL river_elevation = c_elv - r_hgt;
                wse = r_dep + river_elevation;
                pwse = wse;
                if (has_levee && is_levee[i]) pwse = fmin_(c_elv + p_dep, wse);
                const REAL river_elevation_dn = c_elv_dn - r_hgt_dn;
                REAL wse_dn = r_dep_dn + river_elevation_dn;
                if (mouth) wse_dn = c_elv;
                const REAL max_wse = fmax_(wse, wse_dn);
                dwse = wse - wse_dn;
                upd_rcsd = mouth ? r_dep : (max_wse - river_elevation);
                const REAL flood_surface = mouth ? wse : max_wse;
                fsd = flood_surface - c_elv;
            }

            const REAL river_slope = dwse / dn_dist;
            const REAL flood_slope = clamp_(river_slope, (REAL)-0.005, (REAL)0.005);

            const REAL r_sifd = fmax_((REAL)std::sqrt(upd_rcsd * r_csd), (REAL)1e-6);
            const REAL upd_fcsd = fmax_(fsd, (REAL)0.0);
            const REAL f_sifd = fmax_((REAL)std::sqrt(upd_fcsd * f_csd), (REAL)1e-6);

            const REAL upd_fcsa = fmax_(fs / r_len - f_dep * r_wid, (REAL)0.0);
            const REAL f_imp_area = fmax_((REAL)std::sqrt(upd_fcsa * fmax_(f_csa, (REAL)1e-6)), (REAL)1e-6);

            const REAL r_cs_area = upd_rcsd * r_wid;
            const bool river_cond = (r_sifd > (REAL)1e-5) && (r_cs_area > (REAL)1e-5);
            const REAL unit_r = r_out / r_wid;
            const REAL num_r = r_wid * (unit_r + gravity * time_step * r_sifd * river_slope);
            const REAL den_r = (REAL)1.0 + gravity * time_step * (r_man * r_man) * std::fabs(unit_r)
                             * ((REAL)1.0 / (r_sifd * r_sifd * cbrt_real(r_sifd)));
            REAL q_r = num_r / den_r;
            q_r = river_cond ? q_r : (REAL)0.0;

            const bool flood_cond = (f_sifd > (REAL)1e-5) && (upd_fcsa > (REAL)1e-5);
            const REAL num_f = f_out + gravity * time_step * f_imp_area * flood_slope;
            const REAL den_f = (REAL)1.0 + gravity * time_step * (f_man * f_man) * std::fabs(f_out)
                             * ((REAL)1.0 / (f_sifd * cbrt_real(f_sifd))) / f_imp_area;
            REAL q_f = num_f / den_f;
            q_f = flood_cond ? q_f : (REAL)0.0;

            if ((q_r * q_f) < (REAL)0.0) q_f = (REAL)0.0;
            const bool negative = (q_r < (REAL)0.0) && !mouth;
            const REAL total_neg = negative ? (-q_r - q_f) * time_step : (REAL)1.0;
            const REAL limit = fmin_(negative ? (REAL)0.05 * total_storage / total_neg : (REAL)1.0, (REAL)1.0);
            if (negative) { q_r = q_r * limit; q_f = q_f * limit; }

            if (has_res && is_dam_up[i]) {
                // kinematic override upstream of dams, triton/outflow.py:251-272
                REAL bed_slope = (c_elv - elev[dn]) / dn_dist;
                bed_slope = fmax_(bed_slope, min_kin);
                const REAL kin_riv_vel = ((REAL)1.0 / r_man) * (REAL)std::sqrt(bed_slope) * cbrt_real(r_dep * r_dep);
                REAL kin_riv = r_wid * r_dep * kin_riv_vel;
                kin_riv = fmin_(kin_riv, rs / time_step);
                kin_riv = fmax_(kin_riv, (REAL)0.0);
                const REAL bed_slope_f = fmin_(bed_slope, (REAL)0.005);
                const REAL kin_fld_vel = ((REAL)1.0 / f_man) * (REAL)std::sqrt(bed_slope_f) * cbrt_real(f_dep * f_dep);
                const REAL kin_fld_area = fmax_(fs / r_len - f_dep * r_wid, (REAL)0.0);
                REAL kin_fld = kin_fld_area * kin_fld_vel;
                kin_fld = fmin_(kin_fld, fs / time_step);
                kin_fld = fmax_(kin_fld, (REAL)0.0);
                q_r = kin_riv;
                q_f = kin_fld;
            }

            river_outflow[i] = q_r;
            flood_outflow[i] = q_f;
            if (flags & HAS_WATER_SURFACE) wse_out[i] = wse;
            if (flags & HAS_PROTECTED_WATER_SURFACE) pwse_out[i] = pwse;
            rcsd[i] = upd_rcsd;
            fcsd[i] = upd_fcsd;
            fcsa[i] = upd_fcsa;
            if (flags & HAS_TOTAL_STORAGE) total_storage_out[i] = (STO)total_storage;
            river_inflow[i] = (STO)0;
            flood_inflow[i] = (STO)0;
            if (has_bif) gbif[i] = (STO)0;

            const REAL pos = fmax_(q_r, (REAL)0.0) + fmax_(q_f, (REAL)0.0);
            const REAL neg = fmin_(q_r, (REAL)0.0) + fmin_(q_f, (REAL)0.0);
            pos_term[i] = pos * time_step;
            neg_term[i] = mouth ? (REAL)0.0 : -neg * time_step;
        }
        // outgoing_storage += contributions, canonical order (see header)
        for (int64_t j = 0; j < n; ++j) outgoing[j] = outgoing[j] + outgoing_sum(j);
    }

    // --------------------------------------- compute_reservoir_outflow ----
    // triton/reservoir.py:17-184
    void reservoir(REAL time_step) {
        const int64_t R = d[D_R];
        const int32_t* res_idx = f<int32_t>(F_RES_CATCHMENT_IDX);
        const int32_t* dn_idx = f<int32_t>(F_DOWNSTREAM_IDX);
        STO* res_inflow = f<STO>(F_RES_TOTAL_INFLOW);
        REAL* river_outflow = f<REAL>(F_RIVER_OUTFLOW);
        REAL* flood_outflow = f<REAL>(F_FLOOD_OUTFLOW);
        const STO* river_storage = f<STO>(F_RIVER_STORAGE);
        const STO* flood_storage = f<STO>(F_FLOOD_STORAGE);
        const REAL* cons = f<REAL>(F_CONSERVATION_VOLUME);
        const REAL* emer = f<REAL>(F_EMERGENCY_VOLUME);
        const REAL* adjv = f<REAL>(F_ADJUSTMENT_VOLUME);
        const REAL* qn_p = f<REAL>(F_EFFECTIVE_NORMAL_OUTFLOW);
        const REAL* qa_p = f<REAL>(F_ADJUSTMENT_OUTFLOW);
        const REAL* qf_p = f<REAL>(F_FLOOD_CONTROL_OUTFLOW);
        const REAL* runoff = f<REAL>(F_RUNOFF);
        STO* total_storage_out = f<STO>(F_TOTAL_STORAGE);
        STO* outgoing = f<STO>(F_OUTGOING_STORAGE);
        for (int64_t r = 0; r < R; ++r) {
            const int64_t c = res_idx[r];
            const int64_t dn = dn_idx[c];
            const bool mouth = (dn == c);
            const REAL rs = (REAL)river_storage[c];
            const REAL fs = (REAL)flood_storage[c];
            const REAL total = rs + fs;
            const REAL total_inflow = (REAL)res_inflow[c];
            res_inflow[c] = (STO)0;
            const REAL inflow = total_inflow + runoff[c];
            const REAL cv = cons[r], ev = emer[r], av = adjv[r];
            const REAL qn = qn_p[r], qa = qa_p[r], qf = qf_p[r];
            REAL out = (REAL)0;
            if (total <= cv) {
                out = qn * (REAL)std::sqrt(total / cv);
            }
            if ((total > cv) && (total <= av)) {
                const REAL frac2 = (total - cv) / (av - cv);
                out = qn + exp_real((REAL)3.0 * log_real(frac2)) * (qa - qn);
            }
            if ((total > av) && (total <= ev)) {
                const bool flood_period = inflow >= qf;
                const REAL frac3 = (total - av) / (ev - av);
                const REAL tmp = qa + exp_real((REAL)0.1 * log_real(frac3)) * (qf - qa);
                if (flood_period) {
                    const REAL o_flood = qn + ((total - cv) / (ev - cv)) * (inflow - qn);
                    out = fmax_(o_flood, tmp);
                } else {
                    out = tmp;
                }
            }
            if (total > ev) out = (inflow >= qf) ? inflow : qf;
            out = clamp_(out, (REAL)0.0, total / time_step);

            river_outflow[c] = out;
            flood_outflow[c] = (REAL)0.0;
            total_storage_out[c] = (STO)total;
            const REAL new_pos = fmax_(out, (REAL)0.0);
            const REAL new_neg = fmin_(out, (REAL)0.0);
            pos_term[c] = new_pos * time_step;
            neg_term[c] = mouth ? (REAL)0.0 : -(new_neg * time_step);
            outgoing[c] = outgoing_sum(c);
            if (!mouth) outgoing[dn] = outgoing_sum(dn);
        }
    }

    // ------------------------------------- compute_bifurcation_outflow ----
    // triton/bifurcation.py:13-104 and, with levee, triton/levee.py:565-684
    void bifurcation_outflow(REAL time_step, bool levee_variant) {
        const int64_t P = d[D_P], LB = d[D_LB];
        const REAL gravity = (REAL)s[S_GRAVITY];
        const int32_t* ci_p = f<int32_t>(F_BIF_CATCHMENT_IDX);
        const int32_t* di_p = f<int32_t>(F_BIF_DOWNSTREAM_IDX);
        const REAL* man = f<REAL>(F_BIF_MANNING);
        REAL* bout = f<REAL>(F_BIF_OUTFLOW);
        const REAL* wid = f<REAL>(F_BIF_WIDTH);
        const REAL* len = f<REAL>(F_BIF_LENGTH);
        const REAL* bel = f<REAL>(F_BIF_ELEVATION);
        REAL* bcsd = f<REAL>(F_BIF_CSD);
        const REAL* wse = f<REAL>(F_WSE);
        const REAL* pwse = f<REAL>(F_PWSE);
        const STO* total_storage = f<STO>(F_TOTAL_STORAGE);
        STO* outgoing = f<STO>(F_OUTGOING_STORAGE);
        // with PRECISE_SURFACE the end surfaces are formed in double from the
        // state, as the outflow pass forms them; the river depth has not
        // changed since that pass
        const bool precise = (sizeof(REAL) < sizeof(double)) && (flags & PRECISE_SURFACE);
        const REAL* rdep = f<REAL>(F_RIVER_DEPTH);
        const REAL* rhgt = f<REAL>(F_RIVER_HEIGHT);
        const REAL* celv = f<REAL>(F_CATCHMENT_ELEVATION);
        const REAL* pdep = f<REAL>(F_PROTECTED_DEPTH);
        const uint8_t* is_levee = f<uint8_t>(F_IS_LEVEE);
        auto surf = [&](int64_t c) { return (double)rdep[c] + ((double)celv[c] - (double)rhgt[c]); };
        auto psurf = [&](int64_t c, double w) {
            return ((flags & HAS_LEVEE) && is_levee[c]) ? fmin_((double)celv[c] + (double)pdep[c], w) : w;
        };
        for (int64_t p = 0; p < P; ++p) {
            const int64_t ci = ci_p[p], di = di_p[p];
            const REAL length = len[p];
            double w0d = 0.0, w1d = 0.0, max_wd = 0.0, max_pwd = 0.0;
            if (precise) {
                w0d = surf(ci);
                w1d = surf(di);
                max_wd = fmax_(w0d, w1d);
                if (levee_variant) max_pwd = fmax_(psurf(ci, w0d), psurf(di, w1d));
            }
            const REAL w0 = wse[ci], w1 = wse[di];
            const REAL max_w = fmax_(w0, w1);
            REAL max_pw = (REAL)0;
            if (levee_variant) max_pw = fmax_(pwse[ci], pwse[di]);
            REAL slope = (precise ? (REAL)(w0d - w1d) : (w0 - w1)) / length;
            slope = clamp_(slope, (REAL)-0.005, (REAL)0.005);
            const REAL ts0 = (REAL)total_storage[ci];
            const REAL ts1 = (REAL)total_storage[di];
            REAL sum = (REAL)0;
            for (int64_t l = 0; l < LB; ++l) {
                const int64_t k = p * LB + l;
                const REAL csd_old = bcsd[k];
                REAL upd_csd, sifd;
                const REAL above = precise ? (REAL)(max_wd - (double)bel[k]) : max_w - bel[k];
                if (!levee_variant) {
                    upd_csd = fmax_(above, (REAL)0.0);
                    sifd = fmax_((REAL)std::sqrt(upd_csd * csd_old), (REAL)std::sqrt(upd_csd * (REAL)0.01));
                } else if (l == 0) {
                    upd_csd = fmax_(above, (REAL)0.0);
                    const REAL semi = (REAL)std::sqrt(upd_csd * csd_old);
                    sifd = (semi <= (REAL)0.0) ? upd_csd : semi;
                } else {
                    upd_csd = fmax_(precise ? (REAL)(max_pwd - (double)bel[k]) : max_pw - bel[k], (REAL)0.0);
                    sifd = upd_csd;
                }
                const REAL w = wid[k];
                const REAL unit = bout[k] / w;
                const REAL num = w * (unit + gravity * time_step * sifd * slope);
                const REAL den = (REAL)1.0 + gravity * time_step * (man[k] * man[k]) * std::fabs(unit)
                               * ((REAL)1.0 / (sifd * sifd * cbrt_real(sifd)));
                REAL q = num / den;
                q = (sifd > (REAL)1e-5) ? q : (REAL)0.0;
                sum += q;
                bcsd[k] = upd_csd;
                bout[k] = q;
            }
            const REAL limit = fmin_((REAL)0.05 * fmin_(ts0, ts1) / (std::fabs(sum) * time_step), (REAL)1.0);
            sum *= limit;
            for (int64_t l = 0; l < LB; ++l) bout[p * LB + l] *= limit;
            const REAL pos = fmax_(sum, (REAL)0.0);
            const REAL neg = fmin_(sum, (REAL)0.0);
            outgoing[ci] += (STO)(pos * time_step);
            outgoing[di] += (STO)(-neg * time_step);
        }
    }

    // ------------------------------------------------- compute_inflow -----
    // triton/outflow.py:318-383
    void inflow() {
        const int64_t n = N();
        const int32_t* dn_idx = f<int32_t>(F_DOWNSTREAM_IDX);
        REAL* river_outflow = f<REAL>(F_RIVER_OUTFLOW);
        REAL* flood_outflow = f<REAL>(F_FLOOD_OUTFLOW);
        const STO* river_storage = f<STO>(F_RIVER_STORAGE);
        const STO* flood_storage = f<STO>(F_FLOOD_STORAGE);
        const STO* outgoing = f<STO>(F_OUTGOING_STORAGE);
        STO* river_inflow = f<STO>(F_RIVER_INFLOW);
        STO* flood_inflow = f<STO>(F_FLOOD_INFLOW);
        REAL* limit_rate = f<REAL>(F_LIMIT_RATE);
        STO* res_inflow = f<STO>(F_RES_TOTAL_INFLOW);
        const uint8_t* is_res = f<uint8_t>(F_IS_RESERVOIR);
        const bool has_bif = flags & HAS_BIFURCATION, has_res = flags & HAS_RESERVOIR;
        std::vector<REAL> lim(n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const REAL og = (REAL)outgoing[i];
            const REAL rate = (REAL)(river_storage[i] + flood_storage[i]);
            lim[i] = (og > (REAL)1e-8) ? fmin_(rate / og, (REAL)1.0) : (REAL)1.0;
        }
        for (int64_t i = 0; i < n; ++i) {
            const int64_t dn = dn_idx[i];
            const REAL l_own = lim[i], l_dn = lim[dn];
            const REAL qr = river_outflow[i], qf = flood_outflow[i];
            const REAL ur = (qr >= (REAL)0.0) ? qr * l_own : qr * l_dn;
            const REAL uf = (qf >= (REAL)0.0) ? qf * l_own : qf * l_dn;
            river_outflow[i] = ur;
            flood_outflow[i] = uf;
            if (has_bif) limit_rate[i] = l_own;
            if (dn != i) {
                river_inflow[dn] += (STO)ur;
                flood_inflow[dn] += (STO)uf;
                if (has_res && is_res[dn]) res_inflow[dn] += (STO)(ur + uf);
            }
        }
    }

    // -------------------------------------- compute_bifurcation_inflow ----
    // triton/bifurcation.py:106-151
    void bifurcation_inflow() {
        const int64_t P = d[D_P], LB = d[D_LB];
        const int32_t* ci_p = f<int32_t>(F_BIF_CATCHMENT_IDX);
        const int32_t* di_p = f<int32_t>(F_BIF_DOWNSTREAM_IDX);
        const REAL* limit_rate = f<REAL>(F_LIMIT_RATE);
        REAL* bout = f<REAL>(F_BIF_OUTFLOW);
        STO* gbif = f<STO>(F_GLOBAL_BIF_OUTFLOW);
        for (int64_t p = 0; p < P; ++p) {
            const int64_t ci = ci_p[p], di = di_p[p];
            const REAL l0 = limit_rate[ci], l1 = limit_rate[di];
            REAL raw_sum = (REAL)0;
            for (int64_t l = 0; l < LB; ++l) {
                const int64_t k = p * LB + l;
                const REAL raw = bout[k];
                raw_sum += raw;
                bout[k] = (raw >= (REAL)0.0) ? raw * l0 : raw * l1;
            }
            const REAL net = (raw_sum >= (REAL)0.0) ? raw_sum * l0 : raw_sum * l1;
            gbif[ci] += (STO)net;
            gbif[di] += (STO)(-net);
        }
    }

    // ----------------------------------------- compute_flood_stage(_log) --
    // triton/storage.py:16-193 (and :196-425 for the log sums)
    void flood_stage(REAL time_step, int current_step) {
        const int64_t n = N();
        const int L = (int)d[D_L];
        const STO* river_inflow_p = f<STO>(F_RIVER_INFLOW);
        const STO* flood_inflow_p = f<STO>(F_FLOOD_INFLOW);
        const REAL* river_outflow = f<REAL>(F_RIVER_OUTFLOW);
        const REAL* flood_outflow = f<REAL>(F_FLOOD_OUTFLOW);
        const STO* gbif_p = f<STO>(F_GLOBAL_BIF_OUTFLOW);
        const REAL* runoff = f<REAL>(F_RUNOFF);
        STO* outgoing = f<STO>(F_OUTGOING_STORAGE);
        STO* river_storage = f<STO>(F_RIVER_STORAGE);
        STO* flood_storage = f<STO>(F_FLOOD_STORAGE);
        STO* protected_storage = f<STO>(F_PROTECTED_STORAGE);
        REAL* river_depth = f<REAL>(F_RIVER_DEPTH);
        REAL* flood_depth = f<REAL>(F_FLOOD_DEPTH);
        REAL* protected_depth = f<REAL>(F_PROTECTED_DEPTH);
        REAL* flood_fraction = f<REAL>(F_FLOOD_FRACTION);
        const REAL* river_height = f<REAL>(F_RIVER_HEIGHT);
        const REAL* table = f<REAL>(F_FLOOD_DEPTH_TABLE);
        const REAL* catchment_area = f<REAL>(F_CATCHMENT_AREA);
        const REAL* river_width = f<REAL>(F_RIVER_WIDTH);
        const REAL* river_length = f<REAL>(F_RIVER_LENGTH);
        const uint8_t* is_levee = f<uint8_t>(F_IS_LEVEE);
        const bool has_bif = flags & HAS_BIFURCATION;
        const bool log = flags & HAS_LOG;
        const bool has_levee = flags & HAS_LEVEE;
        std::vector<double> part;
        if (log) part.assign((size_t)n * L_COUNT, 0.0);

#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const STO rs = river_storage[i], fs = flood_storage[i], ps = protected_storage[i];
            const REAL rin = (REAL)river_inflow_p[i];
            const REAL fin = (REAL)flood_inflow_p[i];
            const REAL rout = river_outflow[i], fout = flood_outflow[i];
            const REAL gbif = has_bif ? (REAL)gbif_p[i] : (REAL)0.0;
            const REAL ro = runoff[i];
            const REAL inflow_extra = (REAL)0.0;   // prescribed inflow module is off

            STO rsu = rs + (STO)((rin - rout) * time_step);
            STO fsu = fs + ((rsu < (STO)0.0) ? rsu : (STO)0.0) + (STO)((fin - fout - gbif) * time_step);
            rsu = fmax_(rsu, (STO)0.0);
            rsu = (fsu < (STO)0.0) ? fmax_(rsu + fsu, (STO)0.0) : rsu;
            fsu = fmax_(fsu, (STO)0.0);
            const STO total_next = rsu + fsu + ps + (STO)(ro * time_step) + (STO)(inflow_extra * time_step);
            const STO total_hp = fmax_(total_next, (STO)0.0);
            const REAL total = (REAL)total_hp;

            const REAL rh = river_height[i], ca = catchment_area[i];
            const REAL rw = river_width[i], rl = river_length[i];
            const REAL river_max_storage = rl * rw * rh;
            const REAL catchment_width = ca / rl;
            const REAL width_increment = catchment_width / (REAL)L;

            int level = (total > river_max_storage) ? 0 : -1;
            REAL S_accum = river_max_storage;
            REAL prev_H = (REAL)0.0;
            REAL prev_W = rw;
            REAL prev_total_storage = river_max_storage;
            REAL prev_flood_depth = (REAL)0.0;
            REAL next_flood_depth = (REAL)0.0;
            for (int k = 0; k < L; ++k) {
                const REAL H = table[(size_t)i * L + k];
                const REAL W = rw + (REAL)(k + 1) * width_increment;
                const REAL dS = rl * (REAL)0.5 * (prev_W + W) * (H - prev_H);
                const REAL S = S_accum + dS;
                if (level == k) next_flood_depth = H;
                const bool above = total > S;
                level += above ? 1 : 0;
                if (above) { prev_total_storage = S; prev_flood_depth = H; }
                S_accum = S;
                prev_H = H;
                prev_W = W;
            }
            const bool no_flood = level < 0;
            if (level < 0) level = 0;
            const REAL prev_total_width = rw + (REAL)level * width_increment;
            const REAL flood_grad = (level == L) ? (REAL)0.0 : (next_flood_depth - prev_flood_depth) / width_increment;
            const REAL diff_width = (REAL)std::sqrt(prev_total_width * prev_total_width
                + (REAL)2.0 * (total - prev_total_storage) / (flood_grad * rl)) - prev_total_width;
            const REAL fd_mid = prev_flood_depth + diff_width * flood_grad;
            const REAL fd_top = prev_flood_depth + (total - prev_total_storage) / (prev_total_width * rl);
            const REAL fdep = no_flood ? (REAL)0.0 : ((level == L) ? fd_top : fd_mid);

            const REAL cap = river_max_storage + rl * rw * fdep;
            const STO rsf_hp = no_flood ? total_hp : fmin_((STO)cap, total_hp);
            const REAL rsf = (REAL)rsf_hp;
            const REAL rdep = rsf / (rl * rw);
            const REAL ff_mid = clamp_((prev_total_width + diff_width - rw) * rl / ca, (REAL)0.0, (REAL)1.0);
            const REAL ffr = no_flood ? (REAL)0.0 : ((level == L) ? (REAL)1.0 : ff_mid);
            const STO fsf_hp = fmax_(total_hp - rsf_hp, (STO)0.0);

            if (log) {
                double* q = &part[(size_t)i * L_COUNT];
                const bool non_levee = !(has_levee && is_levee[i]);
                const STO total_pre = rs + fs + ps;
                q[L_STORAGE_PRE] = (double)total_pre;
                q[L_STORAGE_NEXT] = (double)total_next;
                q[L_STORAGE_NEW] = (double)total_hp;
                q[L_INFLOW] = (double)((rin + fin + inflow_extra) * time_step);
                q[L_OUTFLOW] = (double)((rout + fout) * time_step);
                q[L_INFLOW_ERROR] = (double)(total_pre - total_next
                    + (STO)((rin + fin + ro + inflow_extra - rout - fout - gbif) * time_step));
                if (non_levee) {
                    const STO stage_new = rsf_hp + fsf_hp;
                    q[L_STORAGE_STAGE] = (double)stage_new;
                    q[L_RIVER_STORAGE] = (double)rsf;
                    q[L_FLOOD_STORAGE] = (double)(REAL)fsf_hp;
                    q[L_FLOOD_AREA] = (double)(ffr * ca);
                    q[L_STAGE_ERROR] = (double)(stage_new - total_hp);
                }
            }

            outgoing[i] = (STO)0.0;
            river_storage[i] = rsf_hp;
            flood_storage[i] = fsf_hp;
            protected_storage[i] = (STO)0.0;
            river_depth[i] = rdep;
            flood_depth[i] = fdep;
            protected_depth[i] = fdep;
            flood_fraction[i] = ffr;
        }
        if (log) accumulate_log(part, n, current_step);
    }

    void accumulate_log(const std::vector<double>& part, int64_t count, int current_step) {
        double* sums = f<double>(F_LOG_SUMS);
        const int64_t stride = d[D_LOGSIZE];
        for (int k = 0; k < L_COUNT; ++k) {
            double acc = 0.0;
            for (int64_t i = 0; i < count; ++i) acc += part[(size_t)i * L_COUNT + k];
            sums[k * stride + current_step] += acc * 1e-9;
        }
    }

    // -------------------------------------- compute_levee_stage(_log) -----
    // triton/levee.py:18-59 (partition), :62-304 (stage), :307-562 (log)
    void levee_stage(int current_step) {
        const int64_t V = d[D_V];
        const int L = (int)d[D_L];
        const int32_t* lidx = f<int32_t>(F_LEVEE_CATCHMENT_IDX);
        STO* river_storage = f<STO>(F_RIVER_STORAGE);
        STO* flood_storage = f<STO>(F_FLOOD_STORAGE);
        STO* protected_storage = f<STO>(F_PROTECTED_STORAGE);
        REAL* river_depth = f<REAL>(F_RIVER_DEPTH);
        REAL* flood_depth = f<REAL>(F_FLOOD_DEPTH);
        REAL* protected_depth = f<REAL>(F_PROTECTED_DEPTH);
        const REAL* river_height = f<REAL>(F_RIVER_HEIGHT);
        const REAL* table = f<REAL>(F_FLOOD_DEPTH_TABLE);
        const REAL* catchment_area = f<REAL>(F_CATCHMENT_AREA);
        const REAL* river_width = f<REAL>(F_RIVER_WIDTH);
        const REAL* river_length = f<REAL>(F_RIVER_LENGTH);
        const REAL* base_h = f<REAL>(F_LEVEE_BASE_HEIGHT);
        const REAL* crown_h = f<REAL>(F_LEVEE_CROWN_HEIGHT);
        const REAL* frac_p = f<REAL>(F_LEVEE_FRACTION);
        REAL* flood_fraction = f<REAL>(F_FLOOD_FRACTION);
        const bool log = flags & HAS_LOG;
        std::vector<double> part;
        if (log) part.assign((size_t)V * L_COUNT, 0.0);

#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < V; ++v) {
            const int64_t c = lidx[v];
            const REAL rl = river_length[c], rw = river_width[c], rh = river_height[c];
            const REAL ca = catchment_area[c];
            const REAL crown = crown_h[v], lfrac = frac_p[v], base = base_h[v];
            const STO rs_hp = river_storage[c], fs_hp = flood_storage[c];
            const REAL fdep_curr = flood_depth[c];
            const STO total_hp = rs_hp + fs_hp;
            const REAL rs_c = (REAL)rs_hp, fs_c = (REAL)fs_hp, total = (REAL)total_hp;

            const REAL river_max_storage = rl * rw * rh;
            const REAL dwth_inc = (ca / rl) / (REAL)L;
            const REAL levee_distance = lfrac * (ca / rl);

            REAL s_curr = river_max_storage;
            REAL dhgt_pre = (REAL)0.0;
            REAL dwth_pre = rw;
            REAL base_sto = river_max_storage, fill_sto = river_max_storage;
            bool found_base = false, found_fill = false;
            const int ilev = (int)(lfrac * (REAL)L);
            REAL dsto_fil_B = (REAL)0.0, dwth_fil_B = (REAL)0.0, ddph_fil_B = (REAL)0.0, gradient_B = (REAL)0.0;
            bool found_B = false;
            for (int k = 0; k < L; ++k) {
                const REAL depth_val = table[(size_t)c * L + k];
                REAL dhgt_seg = depth_val - dhgt_pre;
                dhgt_seg = fmax_(dhgt_seg, (REAL)1e-6);
                const REAL dwth_mid = dwth_pre + (REAL)0.5 * dwth_inc;
                const REAL dsto_seg = rl * dwth_mid * dhgt_seg;
                const REAL s_next = s_curr + dsto_seg;
                const REAL gradient = dhgt_seg / dwth_inc;

                const bool cond_base = (base > dhgt_pre) && (base <= depth_val);
                const REAL ratio_base = (base - dhgt_pre) / dhgt_seg;
                const REAL s_base_cand = s_curr + rl * (dwth_pre + (REAL)0.5 * ratio_base * dwth_inc) * (ratio_base * dhgt_seg);
                if (cond_base) base_sto = s_base_cand;
                found_base = found_base || cond_base;

                const bool cond_fill = (crown > dhgt_pre) && (crown <= depth_val);
                const REAL ratio_fill = (crown - dhgt_pre) / dhgt_seg;
                const REAL s_fill_cand = s_curr + rl * (dwth_pre + (REAL)0.5 * ratio_fill * dwth_inc) * (ratio_fill * dhgt_seg);
                if (cond_fill) fill_sto = s_fill_cand;
                found_fill = found_fill || cond_fill;

                const REAL dhgt_dif_loop = crown - base;
                const REAL s_top_loop = base_sto + (levee_distance + rw) * dhgt_dif_loop * rl;
                const REAL wedge = (levee_distance + rw) * (crown - depth_val) * rl;
                const REAL threshold = s_next + wedge;
                const bool cond_check = (k >= ilev) && !found_B;
                const bool cond_found = cond_check && (total < threshold);
                const REAL lower = (k == ilev) ? s_top_loop : dsto_fil_B;
                const bool advance = cond_check && !cond_found;
                dsto_fil_B = advance ? threshold : lower;
                if (advance) {
                    dwth_fil_B = dwth_inc * (REAL)(k + 1) - levee_distance;
                    ddph_fil_B = depth_val - base;
                }
                if (cond_found) gradient_B = gradient;
                found_B = found_B || cond_found;

                s_curr = s_next;
                dhgt_pre = depth_val;
                dwth_pre += dwth_inc;
            }
            const REAL s_base_extra = s_curr + rl * dwth_pre * (base - dhgt_pre);
            base_sto = found_base ? base_sto : ((base > dhgt_pre) ? s_base_extra : river_max_storage);
            const REAL s_fill_extra = s_curr + rl * dwth_pre * (crown - dhgt_pre);
            fill_sto = found_fill ? fill_sto : ((crown > dhgt_pre) ? s_fill_extra : river_max_storage);

            const REAL dhgt_dif = crown - base;
            const REAL s_top = base_sto + (levee_distance + rw) * dhgt_dif * rl;
            const bool case4 = total >= fill_sto;
            const bool case3 = !case4 && (total >= s_top);
            const bool case2 = !case4 && !case3 && (total >= base_sto);

            // case 2
            const REAL dsto_add_c2 = total - base_sto;
            const REAL dwth_add_c2 = levee_distance + rw;
            const REAL f_dph_c2 = base + dsto_add_c2 / dwth_add_c2 / rl;
            const REAL r_sto_c2 = river_max_storage + rl * rw * f_dph_c2;
            const REAL r_dph_c2 = r_sto_c2 / rl / rw;
            const REAL f_sto_c2 = fmax_(total - r_sto_c2, (REAL)0.0);
            const REAL f_frc_c2 = lfrac;
            // case 3
            const REAL dsto_add_B = total - dsto_fil_B;
            const REAL term_B = dwth_fil_B * dwth_fil_B + (REAL)2.0 * dsto_add_B / rl / (gradient_B + (REAL)1e-9);
            const REAL dwth_add_B = -dwth_fil_B + (REAL)std::sqrt(fmax_(term_B, (REAL)0.0));
            const REAL ddph_add_B = dwth_add_B * gradient_B;
            const REAL p_dph_B_found = base + ddph_fil_B + ddph_add_B;
            const REAL f_frc_B_found = (dwth_fil_B + levee_distance) / (dwth_inc * (REAL)L);
            const REAL ddph_add_B_extra = dsto_add_B / (dwth_fil_B * rl + (REAL)1e-9);
            const REAL p_dph_B_extra = base + ddph_fil_B + ddph_add_B_extra;
            const REAL p_dph_B = found_B ? p_dph_B_found : p_dph_B_extra;
            const REAL f_frc_B = found_B ? f_frc_B_found : (REAL)1.0;
            const REAL f_dph_c3 = crown;
            const REAL r_sto_c3 = river_max_storage + rl * rw * f_dph_c3;
            const REAL r_dph_c3 = r_sto_c3 / rl / rw;
            const REAL f_sto_c3 = fmax_(s_top - r_sto_c3, (REAL)0.0);
            const REAL p_dph_c3 = p_dph_B;
            const REAL f_frc_c3 = clamp_(f_frc_B, (REAL)0.0, (REAL)1.0);
            // case 4
            const REAL f_dph_c4 = fdep_curr;
            const REAL r_sto_c4 = rs_c;
            const REAL dsto_add_c4 = (f_dph_c4 - crown) * (levee_distance + rw) * rl;
            const REAL f_sto_c4 = fmax_(s_top + dsto_add_c4 - r_sto_c4, (REAL)0.0);
            const REAL p_dph_c4 = f_dph_c4;

            const REAL r_dph_curr = river_depth[c];
            const REAL r_cand = case2 ? r_sto_c2 : (case3 ? r_sto_c3 : (case4 ? r_sto_c4 : rs_c));
            const REAL f_cand = case2 ? f_sto_c2 : (case3 ? f_sto_c3 : (case4 ? f_sto_c4 : fs_c));
            // storage partition, triton/levee.py:18-59
            const bool none = !(case2 || case3 || case4);
            STO r_hp = fmin_((STO)r_cand, total_hp);
            r_hp = (case4 || none) ? rs_hp : r_hp;
            const STO remaining = fmax_(total_hp - r_hp, (STO)0.0);
            STO f_hp = fmin_(fmax_((STO)f_cand, (STO)0.0), remaining);
            f_hp = case2 ? remaining : (none ? fs_hp : f_hp);
            const STO p_hp = fmax_(remaining - f_hp, (STO)0.0);

            const REAL r_dph = case2 ? r_dph_c2 : (case3 ? r_dph_c3 : r_dph_curr);
            const REAL f_dph = case2 ? f_dph_c2 : (case3 ? f_dph_c3 : fdep_curr);
            const REAL p_dph = case2 ? (REAL)0.0 : (case3 ? p_dph_c3 : (case4 ? p_dph_c4 : (REAL)0.0));
            const REAL f_frc = case2 ? f_frc_c2 : (case3 ? f_frc_c3 : flood_fraction[c]);

            if (log) {
                double* q = &part[(size_t)v * L_COUNT];
                const STO stage_new = r_hp + f_hp + p_hp;
                q[L_STORAGE_STAGE] = (double)stage_new;
                q[L_RIVER_STORAGE] = (double)r_hp;
                q[L_FLOOD_STORAGE] = (double)f_hp;
                q[L_FLOOD_AREA] = (double)(f_frc * ca);
                q[L_STAGE_ERROR] = (double)(stage_new - total_hp);
            }
            river_storage[c] = r_hp;
            flood_storage[c] = f_hp;
            protected_storage[c] = p_hp;
            river_depth[c] = r_dph;
            flood_depth[c] = f_dph;
            protected_depth[c] = p_dph;
            flood_fraction[c] = f_frc;
        }
        if (log) accumulate_log(part, V, current_step);
    }

    // ------------------------------------ compute_adaptive_time_step -----
    // triton/adaptive_time.py:13-50
    int adaptive_time(REAL outer_time_step) const {
        const int64_t n = N();
        const REAL* river_depth = f<REAL>(F_RIVER_DEPTH);
        const REAL* dist = f<REAL>(F_DOWNSTREAM_DISTANCE);
        const uint8_t* dam = f<uint8_t>(F_IS_DAM_RELATED);
        const bool has_res = flags & HAS_RESERVOIR;
        const REAL factor = (REAL)s[S_ADAPTIVE_FACTOR];
        const REAL gravity = (REAL)s[S_GRAVITY];
        int best = 0;
#pragma omp parallel for reduction(max : best) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            int steps = 1;
            if (!(has_res && dam[i])) {
                const REAL depth = fmax_(river_depth[i], (REAL)0.01);
                const REAL dt = factor * dist[i] / (REAL)std::sqrt(gravity * depth);
                const REAL dtc = fmin_(dt, outer_time_step);
                const REAL nf = (REAL)std::floor(outer_time_step / dtc - (REAL)0.01) + (REAL)1.0;
                steps = (int)nf;
            }
            best = std::max(best, steps);
        }
        return best;
    }

    // ------------------------------------------- CaMaFlood.step_advance ---
    // cmfgpu/models/cama_flood_model.py:104-173
    int step_advance(double outer_seconds, int fixed_count) {
        int count = fixed_count;
        if (flags & HAS_ADAPTIVE) count = adaptive_time((REAL)outer_seconds);
        if (count < 1) count = 1;
        const double sub = outer_seconds / (double)count;
        const REAL time_step = (REAL)sub;
        const bool has_bif = flags & HAS_BIFURCATION;
        const bool has_levee = flags & HAS_LEVEE;
        for (int k = 0; k < count; ++k) {
            outflow(time_step);
            if (flags & HAS_RESERVOIR) reservoir(time_step);
            if (has_bif) bifurcation_outflow(time_step, has_levee);
            inflow();
            if (has_bif) bifurcation_inflow();
            flood_stage(time_step, k);
            if (has_levee) levee_stage(k);
        }
        return count;
    }
};

template <class REAL, class STO>
int run_op(int op, void* const* ptrs, const int64_t* dims, const double* scalars,
           int flags, double arg0, int arg1) {
    Model<REAL, STO> m{ptrs, dims, scalars, flags, {}, {}, {}, {}};
    m.build_upstream();
    switch (op) {
        case 0: return m.step_advance(arg0, arg1);
        case 1: return m.adaptive_time((REAL)arg0);
        default: return -1;
    }
}

}  // namespace

extern "C" {

// Threads of the OpenMP loops (the CPU baseline times one thread and all).
void cmf_oracle_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int cmf_oracle_max_threads() { return omp_get_max_threads(); }

// Advance `num_outer` outer steps of `outer_seconds` each.  With HAS_ADAPTIVE
// the sub-step count comes from the CFL rule; otherwise `fixed_count` is used.
// `counts_out[k]` receives the sub-step count of outer step k.  precision 0 is
// (float, double), 1 is (double, double), 2 is (float, float).
int cmf_oracle_advance(int precision, void* const* ptrs, const int64_t* dims,
                       const double* scalars, int flags, double outer_seconds,
                       int fixed_count, int num_outer, int32_t* counts_out) {
    for (int k = 0; k < num_outer; ++k) {
        int c = precision == 0 ? run_op<float, double>(0, ptrs, dims, scalars, flags, outer_seconds, fixed_count)
              : precision == 2 ? run_op<float, float>(0, ptrs, dims, scalars, flags, outer_seconds, fixed_count)
                               : run_op<double, double>(0, ptrs, dims, scalars, flags, outer_seconds, fixed_count);
        if (counts_out) counts_out[k] = c;
    }
    return 0;
}

int cmf_oracle_adaptive_steps(int precision, void* const* ptrs, const int64_t* dims,
                              const double* scalars, int flags, double outer_seconds) {
    return precision == 0 ? run_op<float, double>(1, ptrs, dims, scalars, flags, outer_seconds, 0)
         : precision == 2 ? run_op<float, float>(1, ptrs, dims, scalars, flags, outer_seconds, 0)
                          : run_op<double, double>(1, ptrs, dims, scalars, flags, outer_seconds, 0);
}

// Single-kernel entry points, for pinning each kernel on its own.
// op: 0 outflow, 1 inflow, 2 flood_stage, 3 bifurcation_outflow,
//     4 bifurcation_inflow, 5 levee_stage, 6 reservoir, 7 levee_bifurcation_outflow
int cmf_oracle_kernel(int precision, int op, void* const* ptrs, const int64_t* dims,
                      const double* scalars, int flags, double time_step, int current_step) {
    auto go = [&](auto& m, auto ts) {
        m.build_upstream();
        switch (op) {
            case 0: m.outflow(ts); break;
            case 1: m.inflow(); break;
            case 2: m.flood_stage(ts, current_step); break;
            case 3: m.bifurcation_outflow(ts, false); break;
            case 4: m.bifurcation_inflow(); break;
            case 5: m.levee_stage(current_step); break;
            case 6: {
                // the canonical re-sum needs the catchment terms of the
                // preceding outflow pass: recompute them from the stored flows
                const int32_t* dn = m.template f<int32_t>(F_DOWNSTREAM_IDX);
                const auto* qr = m.template f<std::remove_reference_t<decltype(ts)>>(F_RIVER_OUTFLOW);
                const auto* qf = m.template f<std::remove_reference_t<decltype(ts)>>(F_FLOOD_OUTFLOW);
                using R = std::remove_reference_t<decltype(ts)>;
                for (int64_t i = 0; i < m.N(); ++i) {
                    const R pos = fmax_(qr[i], (R)0) + fmax_(qf[i], (R)0);
                    const R neg = fmin_(qr[i], (R)0) + fmin_(qf[i], (R)0);
                    m.pos_term[i] = pos * ts;
                    m.neg_term[i] = (dn[i] == i) ? (R)0 : -neg * ts;
                }
                m.reservoir(ts);
                break;
            }
            case 7: m.bifurcation_outflow(ts, true); break;
            default: return -1;
        }
        return 0;
    };
    if (precision == 0) {
        Model<float, double> m{ptrs, dims, scalars, flags, {}, {}, {}, {}};
        return go(m, (float)time_step);
    }
    if (precision == 2) {
        Model<float, float> m{ptrs, dims, scalars, flags, {}, {}, {}, {}};
        return go(m, (float)time_step);
    }
    Model<double, double> m{ptrs, dims, scalars, flags, {}, {}, {}, {}};
    return go(m, (double)time_step);
}

float cmf_oracle_cbrtf(float x) { return cbrt_real(x); }
double cmf_oracle_cbrt(double x) { return cbrt_real(x); }
float cmf_oracle_logf(float x) { return log_real(x); }
float cmf_oracle_expf(float x) { return exp_real(x); }
int cmf_oracle_num_fields(void) { return F_COUNT; }

}  // extern "C"
