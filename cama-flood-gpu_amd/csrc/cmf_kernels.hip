// gfx950 kernels of the CaMa-Flood per-sub-step hydrodynamic path.
//
// Each kernel is one thread per element (catchment, path, levee or
// reservoir) of one trial; the arithmetic of every element follows the
// reference Triton kernel it replaces, operation for operation:
//   k_outflow            cmfgpu/phys/triton/outflow.py:15-315 (batched :386-717)
//   k_inflow_limit       cmfgpu/phys/triton/outflow.py:318-371
//   k_reservoir          cmfgpu/phys/triton/reservoir.py:17-184
//   k_bif_outflow        cmfgpu/phys/triton/bifurcation.py:13-104, levee.py:565-684
//   k_bif_inflow         cmfgpu/phys/triton/bifurcation.py:106-151
//   k_stage              cmfgpu/phys/triton/storage.py:16-193 (log :196-425)
//   k_levee              cmfgpu/phys/triton/levee.py:18-304 (log :307-562)
//   k_cfl                cmfgpu/phys/triton/adaptive_time.py:13-103
// The reference's float atomic scatters (outflow.py:309-314,376-383,
// bifurcation.py:103-104,150-151, reservoir.py:68-77,175-184) become
// gathers: every target sums its contributions itself, in a fixed order, from
// contiguous upstream ranges (k_outgoing_sum, k_inflow_sum, k_limiter,
// k_stage) and per-catchment bifurcation event lists (k_bif_outgoing,
// k_gbif_sum).  The CFL maximum is a wave64 shuffle reduction; water-balance
// sums are FP64 per-block partials folded in a fixed order (k_log_final).

#include "cmf_device.hpp"

namespace cmf {

// Default block size, and the largest one the tunable kernels accept (their
// block size is a launch parameter, a multiple of 64).
constexpr int kBlock = 256;
constexpr int kMaxBlock = 512;
// Width of a group block table entry (catchments of one group per block of
// a grouped network's catchment kernels, which run kGroupRecord threads).
#ifndef CMF_GROUP_RECORD
#define CMF_GROUP_RECORD 512
#endif
constexpr int kGroupRecord = CMF_GROUP_RECORD;
static_assert(kGroupRecord % 128 == 0 && kGroupRecord <= kMaxBlock, "group records: 128..512 catchments");

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// With CMF_XCD_REMAP the blocks are renumbered so that every XCD walks one
// contiguous stretch of the breadth-first layout (a bijection on
// [0, gridDim.x)).  Measured slower on the 6-arcmin workload (the plain
// rate drops about 5 %), so it is off by default.
#ifndef CMF_XCD_REMAP
#define CMF_XCD_REMAP 0
#endif
constexpr int64_t kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t nb = gridDim.x, b = blockIdx.x;
    if (!CMF_XCD_REMAP || nb < kNumXcd) return b;
    const int64_t q = nb / kNumXcd, r = nb % kNumXcd, x = b % kNumXcd, i = b / kNumXcd;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Store `val` unless the location already holds it bit for bit (saves the
// write of values that did not change, such as the zero flood state of dry
// catchments).
__device__ __forceinline__ unsigned int bits_of(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned long long bits_of(double x) {
    return (unsigned long long)__double_as_longlong(x);
}
template <class X>
__device__ __forceinline__ void store_if_changed(X* p, X old, X val) {
    if (bits_of(old) != bits_of(val)) *p = val;
}


// =========================================================== outflow ====
// FUSED: the unlimited flows go to the qraw scratch (the limiter and the
// inflow pass read them); inflows are not zeroed because the fused stage
// pass writes them whole.  Otherwise flows are written in place, as in the
// reference, and river/flood inflow and the bifurcation total are zeroed.
// Both write the outgoing-storage terms (pos_term / neg_term) of each
// catchment, for the passes that sum them (k_outgoing_sum, k_limiter).
// FUSED writes the water surfaces and the total storage, which no later pass
// of the fused schedule reads, on the catchment's last sub-step only
// (`store_last`: that flag for one shared count).  `blk`: the block's index
// in the catchment grid.
// Block-local schedule (k_fused_local): per catchment flags of a network,
// for one block size, saying which parts of the sub-step the block holding
// the catchment completes itself (LL: the storage limiter, SL: the stage) and
// which scratch values other blocks or the list passes read (the *_G bits).
enum : uint8_t {
    kLocalLim = 1, kLocalStage = 2, kNegG = 4, kPosG = 8, kQrawG = 16, kLimG = 32
};

// The outflow of one catchment as the block-local schedule keeps it.
template <class REAL, class STO>
struct OutflowLocal {
    REAL q_r, q_f, pos, neg;
    STO rs, fs;
};

// PROV (fused, not block-local): instead of the outgoing terms, the storage
// limiter formed from the catchment's own positive term alone, and a flag
// at the downstream catchment when this one flows back into it (see
// k_limiter_fix).
template <class REAL, class STO, bool FUSED, class VT, bool PROV = false>
__device__ __forceinline__ void outflow_one(const VT& v, int substep, int64_t blk, int store_last,
                                            OutflowLocal<REAL, STO>* local = nullptr, uint8_t lflags = 0) {
    int64_t t, i;
    if (!v.catchment_in(blk, t, i)) return;
    if (!v.active_in_block(t, i, substep)) return;
    const int64_t e = v.el(t, i);
    const int64_t d = v.dn[i];
    const int64_t ed = v.el(t, d);
    const bool mouth = (d == i);
    const REAL time_step = v.dt(t, i);
    const REAL gravity = v.gravity;

    const REAL r_out = v.river_outflow[e];
    const REAL r_man = v.prm(v.river_manning, CMF_BATCHED_RIVER_MANNING, e, i);
    const REAL r_dep = v.river_depth[e];
    const REAL r_wid = v.prm(v.river_width, CMF_BATCHED_RIVER_WIDTH, e, i);
    const REAL r_len = v.prm(v.river_length, CMF_BATCHED_RIVER_LENGTH, e, i);
    const REAL r_hgt = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, i);
    const REAL f_out = v.flood_outflow[e];
    const REAL f_man = v.prm(v.flood_manning, CMF_BATCHED_FLOOD_MANNING, e, i);
    const REAL f_dep = v.flood_depth[e];
    const REAL c_elv = v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, e, i);
    const REAL dn_dist = v.prm(v.dist, CMF_BATCHED_DOWNSTREAM_DISTANCE, e, i);
    const STO rs_s = v.river_storage[e], fs_s = v.flood_storage[e];
    const REAL rs = (REAL)rs_s;
    const REAL fs = (REAL)fs_s;
    // A stage pass folds protected storage into the total and leaves +0.0 in
    // every catchment without a levee, so after the first sub-step of an
    // outer step only levee catchments hold any.
    const bool lev_i = v.has(CMF_HAS_LEVEE) && v.is_levee[i];
    const REAL ps = (FUSED && substep > 0 && !lev_i) ? (REAL)0.0 : (REAL)v.protected_storage[e];
    const REAL r_csd = v.rcsd[e], f_csd = v.fcsd[e], f_csa = v.fcsa[e];

    const REAL total_storage = rs + fs + ps;
    const REAL r_dep_dn = v.river_depth[ed];
    const REAL r_hgt_dn = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, ed, d);
    const REAL c_elv_dn = v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, ed, d);
    // wse: water surface; dwse: its drop to the downstream surface;
    // upd_rcsd / fsd: the higher of the two above the river bed / the bank
    REAL wse, pwse, dwse, upd_rcsd, fsd;
    if (precise_surfaces(v)) {
        // in double from the REAL depth, elevation and height (sums of two
        // or three floats: exact), each difference rounded once
        const double rel = (double)c_elv - (double)r_hgt;
        const double w = (double)r_dep + rel;
        const double w_dn = mouth ? (double)c_elv : (double)r_dep_dn + ((double)c_elv_dn - (double)r_hgt_dn);
        const double mw = vmax(w, w_dn);
        wse = (REAL)w;
        pwse = lev_i ? (REAL)vmin((double)c_elv + (double)v.protected_depth[e], w) : wse;
        dwse = (REAL)(w - w_dn);
        upd_rcsd = mouth ? r_dep : (REAL)(mw - rel);
        fsd = (REAL)((mouth ? w : mw) - (double)c_elv);
    } else {
        const REAL river_elevation = c_elv - r_hgt;
        wse = r_dep + river_elevation;
        pwse = wse;
        if (lev_i) pwse = vmin(c_elv + v.protected_depth[e], wse);
        const REAL river_elevation_dn = c_elv_dn - r_hgt_dn;
        REAL wse_dn = r_dep_dn + river_elevation_dn;
        if (mouth) wse_dn = c_elv;
        const REAL max_wse = vmax(wse, wse_dn);
        dwse = wse - wse_dn;
        upd_rcsd = mouth ? r_dep : (max_wse - river_elevation);
        const REAL flood_surface = mouth ? wse : max_wse;
        fsd = flood_surface - c_elv;
    }

    const REAL river_slope = dwse / dn_dist;
    const REAL flood_slope = vclamp(river_slope, (REAL)-0.005, (REAL)0.005);

    const REAL r_sifd = vmax(dsqrt(upd_rcsd * r_csd), (REAL)1e-6);
    const REAL upd_fcsd = vmax(fsd, (REAL)0.0);
    const REAL f_sifd = vmax(dsqrt(upd_fcsd * f_csd), (REAL)1e-6);
    const REAL upd_fcsa = vmax(fs / r_len - f_dep * r_wid, (REAL)0.0);

    const REAL r_cs_area = upd_rcsd * r_wid;
    const bool river_cond = (r_sifd > (REAL)1e-5) && (r_cs_area > (REAL)1e-5);
    REAL q_r = (REAL)0.0;
    if (river_cond) {
        const REAL unit_r = r_out / r_wid;
        const REAL num_r = r_wid * (unit_r + gravity * time_step * r_sifd * river_slope);
        const REAL den_r = (REAL)1.0 + gravity * time_step * (r_man * r_man) * dabs(unit_r)
                         * ((REAL)1.0 / (r_sifd * r_sifd * dcbrt(r_sifd)));
        q_r = num_r / den_r;
    }
    const bool flood_cond = (f_sifd > (REAL)1e-5) && (upd_fcsa > (REAL)1e-5);
    REAL q_f = (REAL)0.0;
    if (flood_cond) {
        const REAL f_imp_area = vmax(dsqrt(upd_fcsa * vmax(f_csa, (REAL)1e-6)), (REAL)1e-6);
        const REAL num_f = f_out + gravity * time_step * f_imp_area * flood_slope;
        const REAL den_f = (REAL)1.0 + gravity * time_step * (f_man * f_man) * dabs(f_out)
                         * ((REAL)1.0 / (f_sifd * dcbrt(f_sifd))) / f_imp_area;
        q_f = num_f / den_f;
    }
    if ((q_r * q_f) < (REAL)0.0) q_f = (REAL)0.0;
    const bool negative = (q_r < (REAL)0.0) && !mouth;
    if (negative) {
        const REAL total_neg = (-q_r - q_f) * time_step;
        const REAL limit = vmin((REAL)0.05 * total_storage / total_neg, (REAL)1.0);
        q_r = q_r * limit;
        q_f = q_f * limit;
    }
    if (v.has(CMF_HAS_RESERVOIR) && v.is_dam_up[i]) {
        REAL bed_slope = (c_elv - c_elv_dn) / dn_dist;
        bed_slope = vmax(bed_slope, v.min_kin);
        const REAL kin_riv_vel = ((REAL)1.0 / r_man) * dsqrt(bed_slope) * dcbrt(r_dep * r_dep);
        REAL kin_riv = r_wid * r_dep * kin_riv_vel;
        kin_riv = vmin(kin_riv, rs / time_step);
        kin_riv = vmax(kin_riv, (REAL)0.0);
        const REAL bed_slope_f = vmin(bed_slope, (REAL)0.005);
        const REAL kin_fld_vel = ((REAL)1.0 / f_man) * dsqrt(bed_slope_f) * dcbrt(f_dep * f_dep);
        const REAL kin_fld_area = vmax(fs / r_len - f_dep * r_wid, (REAL)0.0);
        REAL kin_fld = kin_fld_area * kin_fld_vel;
        kin_fld = vmin(kin_fld, fs / time_step);
        kin_fld = vmax(kin_fld, (REAL)0.0);
        q_r = kin_riv;
        q_f = kin_fld;
    }

    if (v.tape_q_r) {
        // the input state goes to the tape with non-temporal stores: the
        // reverse sweep reads it long after, so it need not displace the
        // fields the rest of the sub-step reads from the caches
        __builtin_nontemporal_store(r_out, v.tape_q_r + e);
        __builtin_nontemporal_store(f_out, v.tape_q_f + e);
        __builtin_nontemporal_store(r_csd, v.tape_rcsd + e);
        __builtin_nontemporal_store(f_csd, v.tape_fcsd + e);
        __builtin_nontemporal_store(f_csa, v.tape_fcsa + e);
        __builtin_nontemporal_store(r_dep, v.tape_rd + e);
        __builtin_nontemporal_store(f_dep, v.tape_fd + e);
        __builtin_nontemporal_store(rs_s, v.tape_rs + e);
        __builtin_nontemporal_store(fs_s, v.tape_fs + e);
        // levee catchments only: every other catchment holds zero protected
        // storage (the adjoint takes it as zero there)
        if (v.tape_ps && lev_i) {
            v.tape_ps[e] = v.protected_storage[e];
            v.tape_pd[e] = v.protected_depth[e];
        }
        if (v.tape_ri) v.tape_ri[e] = v.res_inflow[e];
    }
    if (FUSED && local) {
        // block-local schedule: the scratch only where another block or a
        // list pass reads it
        if (lflags & kQrawG) {
            v.qraw_r[e] = q_r;
            v.qraw_f[e] = q_f;
        }
    } else if (FUSED) {
        v.qraw_r[e] = q_r;
        v.qraw_f[e] = q_f;
    } else {
        v.river_outflow[e] = q_r;
        v.flood_outflow[e] = q_f;
        v.river_inflow[e] = (STO)0;
        v.flood_inflow[e] = (STO)0;
        if (v.has(CMF_HAS_BIFURCATION)) v.gbif[e] = (STO)0;
    }
    const bool fin = !FUSED || v.last_substep(t, i, substep, store_last);
    if (fin && v.has(CMF_HAS_WATER_SURFACE)) v.wse[e] = wse;
    if (fin && v.has(CMF_HAS_PROTECTED_WATER_SURFACE)) v.pwse[e] = pwse;
    v.rcsd[e] = upd_rcsd;
    store_if_changed(v.fcsd + e, f_csd, upd_fcsd);
    store_if_changed(v.fcsa + e, f_csa, upd_fcsa);
    if (fin && v.has(CMF_HAS_TOTAL_STORAGE)) v.total_storage[e] = (STO)total_storage;
    const REAL pos = vmax(q_r, (REAL)0.0) + vmax(q_f, (REAL)0.0);
    const REAL neg = vmin(q_r, (REAL)0.0) + vmin(q_f, (REAL)0.0);
    const REAL pos_t = pos * time_step;
    const REAL neg_t = mouth ? (REAL)0.0 : -neg * time_step;
    if (FUSED && local) {
        if (lflags & kPosG) v.pos_term[e] = pos_t;
        if (lflags & kNegG) v.neg_term[e] = neg_t;
        local->q_r = q_r;
        local->q_f = q_f;
        local->pos = pos_t;
        local->neg = neg_t;
        local->rs = rs_s;
        local->fs = fs_s;
        return;
    }
    if constexpr (PROV) {
        // The limiter (outflow.py:346-352) sums the negative terms of the
        // upstream catchments, then the own positive term, then the path
        // terms.  Without backflow from upstream those terms are zeros, the
        // sum is the own term, and the limiter is known here.  A catchment
        // with backflow flags its downstream catchment, whose limiter the
        // fix-up pass forms in full, as it does for every path end.
        STO acc = (STO)0;
        acc += (STO)pos_t;
        const REAL og = (REAL)acc;
        const REAL rate = (REAL)(rs_s + fs_s);
        const REAL l = (og > (REAL)1e-8) ? vmin(rate / og, (REAL)1.0) : (REAL)1.0;
        v.lim[e] = l;
        if (v.has(CMF_HAS_BIFURCATION) && fin) v.limit_rate[e] = l;
        if (neg_t != (REAL)0.0)
            __hip_atomic_fetch_or(v.bmask + (ed >> 5), 1u << (ed & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    v.pos_term[e] = pos_t;
    v.neg_term[e] = neg_t;
}

template <class REAL, class STO, bool FUSED, class VT = View<REAL, STO>>
__global__ __launch_bounds__(kMaxBlock) void k_outflow(VT v, int substep, int store_last) {
    outflow_one<REAL, STO, FUSED>(v, substep, xcd_block(), store_last);
}

// fused, provisional limiters instead of outgoing terms (the three-pass
// schedule, before k_limiter_fix)
template <class REAL, class STO, class VT>
__global__ __launch_bounds__(kMaxBlock) void k_outflow_prov(VT v, int substep, int store_last) {
    outflow_one<REAL, STO, true, VT, true>(v, substep, xcd_block(), store_last);
}

// Outgoing total of catchment j from the catchment terms: upstream negative
// terms in ascending network index, then j's own positive term.
template <class REAL, class STO, bool CONTIG, class VT>
__device__ __forceinline__ STO outgoing_terms(const VT& v, int64_t t, int64_t j) {
    STO acc = (STO)0;
    const int2 rg = v.up_range(j);
    const int32_t b = rg.x, en = rg.y;
    for (int32_t k = b; k < en; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == j) continue;  // a mouth's range ends with the mouth
        acc += (STO)v.neg_term[v.el(t, u)];
    }
    acc += (STO)v.pos_term[v.el(t, j)];
    return acc;
}

// per-op: outgoing_storage += catchment terms (compute_outflow's scatter)
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_outgoing_sum(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.N * v.T) return;
    int64_t t, j;
    v.split_c(g, t, j);
    const int64_t e = v.el(t, j);
    v.outgoing[e] = v.outgoing[e] + outgoing_terms<REAL, STO, CONTIG>(v, t, j);
}

// ========================================================= reservoir ====
// Replaces the reservoir catchment's outflow and its outgoing terms.  FUSED
// writes the flow into the qraw scratch.
template <class REAL, class STO, bool FUSED>
__global__ __launch_bounds__(kMaxBlock) void k_reservoir(View<REAL, STO> v, int substep) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.R * v.T) return;
    const int64_t t = g / v.R, r = g - t * v.R;
    const int64_t c = v.res_idx[r];
    if (!v.active(t, c, substep)) return;
    const int64_t d = v.dn[c];
    const bool mouth = (d == c);
    const int64_t e = v.el(t, c);
    const REAL time_step = v.dt(t, c);
    const REAL rs = (REAL)v.river_storage[e];
    const REAL fs = (REAL)v.flood_storage[e];
    const REAL total = rs + fs;
    const REAL total_inflow = (REAL)v.res_inflow[e];
    v.res_inflow[e] = (STO)0;
    const REAL runoff = v.prm(v.runoff, CMF_BATCHED_RUNOFF, e, c);
    const REAL inflow = total_inflow + runoff;
    const REAL cv = v.res_cons[r], ev = v.res_emer[r], av = v.res_adj[r];
    const REAL qn = v.res_qn[r], qa = v.res_qa[r], qf = v.res_qf[r];
    REAL out = (REAL)0;
    if (total <= cv) out = qn * dsqrt(total / cv);
    if ((total > cv) && (total <= av)) {
        const REAL frac2 = (total - cv) / (av - cv);
        out = qn + dexp((REAL)3.0 * dlog(frac2)) * (qa - qn);
    }
    if ((total > av) && (total <= ev)) {
        const bool flood_period = inflow >= qf;
        const REAL frac3 = (total - av) / (ev - av);
        const REAL tmp = qa + dexp((REAL)0.1 * dlog(frac3)) * (qf - qa);
        if (flood_period) {
            const REAL o_flood = qn + ((total - cv) / (ev - cv)) * (inflow - qn);
            out = vmax(o_flood, tmp);
        } else {
            out = tmp;
        }
    }
    if (total > ev) out = (inflow >= qf) ? inflow : qf;
    out = vclamp(out, (REAL)0.0, total / time_step);

    if (FUSED) {
        v.qraw_r[e] = out;
        v.qraw_f[e] = (REAL)0.0;
    } else {
        v.river_outflow[e] = out;
        v.flood_outflow[e] = (REAL)0.0;
    }
    v.total_storage[e] = (STO)total;
    const REAL new_pos = vmax(out, (REAL)0.0);
    const REAL new_neg = vmin(out, (REAL)0.0);
    v.pos_term[e] = new_pos * time_step;
    v.neg_term[e] = mouth ? (REAL)0.0 : -(new_neg * time_step);
}

// per-op: re-sum the outgoing totals touched by a reservoir (its catchment
// and its downstream catchment).  Two reservoirs may share a target; they
// write the same value.
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_res_regather(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.R * v.T) return;
    const int64_t t = g / v.R, r = g - t * v.R;
    const int64_t c = v.res_idx[r];
    const int64_t d = v.dn[c];
    v.outgoing[v.el(t, c)] = outgoing_terms<REAL, STO, CONTIG>(v, t, c);
    if (d != c) v.outgoing[v.el(t, d)] = outgoing_terms<REAL, STO, CONTIG>(v, t, d);
}

// ====================================================== bifurcation ======
// Water surface, protected water surface and total storage of catchment i
// from the state the stage pass left, by the expressions of the outflow pass
// (outflow.py:185-200, levee.py:565-600); bit-identical to the values that
// pass stores in the water_surface_elevation, protected_water_surface_elevation
// and total_storage fields.
template <class REAL, class STO, class VT>
__device__ __forceinline__ double surface_of_d(const VT& v, int64_t e, int64_t i) {
    const REAL c_elv = v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, e, i);
    const REAL r_hgt = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, i);
    return (double)v.river_depth[e] + ((double)c_elv - (double)r_hgt);
}
template <class REAL, class STO, class VT>
__device__ __forceinline__ double protected_surface_of_d(const VT& v, int64_t e, int64_t i, double w) {
    if (!(v.has(CMF_HAS_LEVEE) && v.is_levee[i])) return w;
    return vmin((double)v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, e, i) + (double)v.protected_depth[e], w);
}
template <class REAL, class STO, class VT>
__device__ __forceinline__ REAL surface_of(const VT& v, int64_t e, int64_t i) {
    const REAL c_elv = v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, e, i);
    const REAL r_hgt = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, i);
    return v.river_depth[e] + (c_elv - r_hgt);
}
template <class REAL, class STO, class VT>
__device__ __forceinline__ REAL protected_surface_of(const VT& v, int64_t e, int64_t i, REAL wse) {
    if (!(v.has(CMF_HAS_LEVEE) && v.is_levee[i])) return wse;
    return vmin(v.prm(v.elev, CMF_BATCHED_CATCHMENT_ELEVATION, e, i) + v.protected_depth[e], wse);
}
template <class REAL, class STO>
__device__ __forceinline__ REAL total_of(const View<REAL, STO>& v, int64_t e) {
    return (REAL)v.river_storage[e] + (REAL)v.flood_storage[e] + (REAL)v.protected_storage[e];
}

// One path-trial element `g`.  LBFIX > 0: the level count known at compile
// time (loop unrolled, the per-level flows kept in registers).  FROM_STATE:
// the end catchments' surfaces and total storages are formed from the state
// (surface_of, total_of), so the path pass can share the launch of the
// catchment outflow pass; otherwise they are read from the fields that pass
// wrote.  Besides the per-level state and the outgoing terms, writes path_raw:
// the sum of the limited per-level flows, in level order, as
// compute_bifurcation_inflow forms it (bifurcation.py:132-137).
template <class REAL, class STO, bool LEVEE, int LBFIX, bool FROM_STATE, class VT>
__device__ __forceinline__ void bif_path(const VT& v, int substep, int64_t g) {
    if (g >= v.P * v.T) return;
    int64_t t, p;
    v.split_p(g, t, p);
    const int64_t ci = v.bif_ci[p], di = v.bif_di[p];
    if (!v.active(t, ci, substep)) return;
    const int64_t pe = v.pel(t, p);
    const int64_t eci = v.el(t, ci), edi = v.el(t, di);
    const REAL time_step = v.dt(t, ci);
    const REAL gravity = v.gravity;
    const REAL length = v.prm(v.bif_length, CMF_BATCHED_BIFURCATION_LENGTH, pe, p);
    REAL w0, w1, ts0, ts1, max_pw = (REAL)0;
    // precise surfaces: formed in double from the state as the outflow pass
    // forms them (the river depth is unchanged since that pass)
    const bool precise = precise_surfaces(v);
    double w0d = 0.0, w1d = 0.0, max_wd = 0.0, max_pwd = 0.0;
    if (precise) {
        w0d = surface_of_d<REAL, STO>(v, eci, ci);
        w1d = surface_of_d<REAL, STO>(v, edi, di);
        max_wd = vmax(w0d, w1d);
        if (LEVEE) max_pwd = vmax(protected_surface_of_d<REAL, STO>(v, eci, ci, w0d), protected_surface_of_d<REAL, STO>(v, edi, di, w1d));
    }
    if (FROM_STATE) {
        w0 = surface_of<REAL, STO>(v, eci, ci);
        w1 = surface_of<REAL, STO>(v, edi, di);
        if (LEVEE) max_pw = vmax(protected_surface_of<REAL, STO>(v, eci, ci, w0), protected_surface_of<REAL, STO>(v, edi, di, w1));
        ts0 = total_of(v, eci);
        ts1 = total_of(v, edi);
    } else {
        w0 = v.wse[eci];
        w1 = v.wse[edi];
        if (LEVEE) max_pw = vmax(v.pwse[eci], v.pwse[edi]);
        ts0 = (REAL)v.total_storage[eci];
        ts1 = (REAL)v.total_storage[edi];
    }
    const REAL max_w = vmax(w0, w1);
    REAL slope = (precise ? (REAL)(w0d - w1d) : (w0 - w1)) / length;
    slope = vclamp(slope, (REAL)-0.005, (REAL)0.005);
    const int LB = LBFIX > 0 ? LBFIX : (int)v.LB;
    REAL qs[LBFIX > 0 ? LBFIX : 1];
    REAL sum = (REAL)0;
#pragma unroll
    for (int l = 0; l < (LBFIX > 0 ? LBFIX : 64); ++l) {
        if (LBFIX == 0 && l >= LB) break;
        const int64_t k = pe * LB + l;
        const int64_t ks = p * LB + l;
        const REAL csd_old = v.bif_csd[k];
        const REAL bel = v.prm(v.bif_elev, CMF_BATCHED_BIFURCATION_ELEVATION, k, ks);
        REAL upd_csd, sifd;
        const REAL above = precise ? (REAL)(max_wd - (double)bel) : max_w - bel;
        if (!LEVEE) {
            upd_csd = vmax(above, (REAL)0.0);
            sifd = vmax(dsqrt(upd_csd * csd_old), dsqrt(upd_csd * (REAL)0.01));
        } else if (l == 0) {
            upd_csd = vmax(above, (REAL)0.0);
            const REAL semi = dsqrt(upd_csd * csd_old);
            sifd = (semi <= (REAL)0.0) ? upd_csd : semi;
        } else {
            upd_csd = vmax(precise ? (REAL)(max_pwd - (double)bel) : max_pw - bel, (REAL)0.0);
            sifd = upd_csd;
        }
        const REAL w = v.prm(v.bif_width, CMF_BATCHED_BIFURCATION_WIDTH, k, ks);
        const REAL man = v.prm(v.bif_manning, CMF_BATCHED_BIFURCATION_MANNING, k, ks);
        const REAL q_old = v.bif_outflow[k];
        if (v.tape_bq) {
            v.tape_bq[k] = q_old;
            v.tape_bcsd[k] = csd_old;
        }
        REAL q = (REAL)0.0;
        if (sifd > (REAL)1e-5) {
            const REAL unit = q_old / w;
            const REAL num = w * (unit + gravity * time_step * sifd * slope);
            const REAL den = (REAL)1.0 + gravity * time_step * (man * man) * dabs(unit)
                           * ((REAL)1.0 / (sifd * sifd * dcbrt(sifd)));
            q = num / den;
        }
        sum += q;
        store_if_changed(v.bif_csd + k, csd_old, upd_csd);
        if (LBFIX > 0) qs[LBFIX > 0 ? l : 0] = q;
        else v.bif_outflow[k] = q;
    }
    const REAL limit = vmin((REAL)0.05 * vmin(ts0, ts1) / (dabs(sum) * time_step), (REAL)1.0);
    sum *= limit;
    REAL raw = (REAL)0;
#pragma unroll
    for (int l = 0; l < (LBFIX > 0 ? LBFIX : 64); ++l) {
        if (LBFIX == 0 && l >= LB) break;
        const int64_t k = pe * LB + l;
        const REAL q = (LBFIX > 0 ? qs[LBFIX > 0 ? l : 0] : v.bif_outflow[k]) * limit;
        v.bif_outflow[k] = q;
        if (v.tape_bqlim) v.tape_bqlim[k] = q;
        raw += q;
    }
    const REAL pos = vmax(sum, (REAL)0.0);
    const REAL neg = vmin(sum, (REAL)0.0);
    const REAL pos_t = pos * time_step, neg_t = -neg * time_step;
    v.path_pos[pe] = pos_t;
    v.path_neg[pe] = neg_t;
    v.path_raw[pe] = raw;
    if (v.eterm) {
        // the same terms at the two ends' event slots, where the limiter
        // fix-up reads each path end's terms in event order
        const int2 sl = v.pev[p];
        v.eterm[v.evel(t, sl.x)] = pos_t;
        v.eterm[v.evel(t, sl.y)] = neg_t;
    }
}

// per-op: compute_bifurcation_outflow
template <class REAL, class STO, bool LEVEE, int LBFIX>
__global__ __launch_bounds__(kMaxBlock) void k_bif_outflow(View<REAL, STO> v, int substep) {
    bif_path<REAL, STO, LEVEE, LBFIX, false>(v, substep, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// fused, with bifurcation: the catchment outflow pass and the path pass in
// one launch.  Blocks [0, path_blocks) run the paths, first, so that their
// longer threads overlap the catchment blocks; the rest run the catchments.
template <class REAL, class STO, bool LEVEE, int LBFIX, class VT = View<REAL, STO>>
__global__ __launch_bounds__(kMaxBlock) void k_outflow_paths(VT v, int substep, int64_t path_blocks,
                                                         int store_last) {
    const int64_t b = blockIdx.x;
    if (b < path_blocks) bif_path<REAL, STO, LEVEE, LBFIX, true>(v, substep, b * blockDim.x + threadIdx.x);
    else outflow_one<REAL, STO, true>(v, substep, b - path_blocks, store_last);
}

// the same with provisional limiters (the three-pass schedule)
template <class REAL, class STO, bool LEVEE, int LBFIX, class VT>
__global__ __launch_bounds__(kMaxBlock) void k_outflow_paths_prov(VT v, int substep, int64_t path_blocks,
                                                              int store_last) {
    const int64_t b = blockIdx.x;
    if (b < path_blocks) bif_path<REAL, STO, LEVEE, LBFIX, true>(v, substep, b * blockDim.x + threadIdx.x);
    else outflow_one<REAL, STO, true, VT, true>(v, substep, b - path_blocks, store_last);
}

// fused: bifurcation total of catchment j (compute_bifurcation_inflow,
// bifurcation.py:128-151) in ascending path order, each path net formed from
// the path's level sum and the limiters of its two ends (l_own: j's); the
// upstream end also writes the path's limited per-level flows.
template <class REAL, class STO, class Src, class VT>
__device__ __forceinline__ STO bif_nets(const VT& v, int64_t t, int64_t j, REAL l_own, const Src& src) {
    STO acc = (STO)0;
    const int32_t b = v.ev_ptr[j], en = v.ev_ptr[j + 1];
    for (int32_t k = b; k < en; ++k) {
        const int32_t code = v.ev_code[k];
        const int64_t pe = v.pel(t, code >> 1);
        const int64_t other = v.ev_other[k];
        const REAL l_other = src.lim(v, v.el(t, other), other);
        const bool down_end = (code & 1) != 0;
        const REAL l0 = down_end ? l_other : l_own;   // limiter of the upstream end
        const REAL l1 = down_end ? l_own : l_other;   // of the downstream end
        const REAL raw = v.path_raw[pe];
        const REAL net = (raw >= (REAL)0.0) ? raw * l0 : raw * l1;
        acc += down_end ? (STO)(-net) : (STO)net;
        if (!down_end) {
            REAL* q = v.bif_outflow + pe * v.LB;
            if (v.LB == 5) {
                // all five loads in flight together, then the stores
                const REAL q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
                q[0] = (q0 >= (REAL)0.0) ? q0 * l0 : q0 * l1;
                q[1] = (q1 >= (REAL)0.0) ? q1 * l0 : q1 * l1;
                q[2] = (q2 >= (REAL)0.0) ? q2 * l0 : q2 * l1;
                q[3] = (q3 >= (REAL)0.0) ? q3 * l0 : q3 * l1;
                q[4] = (q4 >= (REAL)0.0) ? q4 * l0 : q4 * l1;
            } else {
                for (int64_t l = 0; l < v.LB; ++l) q[l] = (q[l] >= (REAL)0.0) ? q[l] * l0 : q[l] * l1;
            }
        }
    }
    return acc;
}

// Sum of the bifurcation events of catchment j in ascending path order.
// code = 2 p + role; role 0: j is the path's upstream end, 1: downstream end.
template <class REAL, class STO>
__device__ __forceinline__ STO bif_event_sum(const View<REAL, STO>& v, int64_t t, int64_t j,
                                             STO acc, bool outgoing_terms_mode) {
    const int32_t b = v.ev_ptr[j], en = v.ev_ptr[j + 1];
    for (int32_t k = b; k < en; ++k) {
        const int32_t code = v.ev_code[k];
        const int64_t pe = v.pel(t, code >> 1);
        if (outgoing_terms_mode) {
            acc += (STO)((code & 1) ? v.path_neg[pe] : v.path_pos[pe]);
        } else {
            const REAL net = v.path_net[pe];
            acc += (code & 1) ? (STO)(-net) : (STO)net;
        }
    }
    return acc;
}

// The same from the first-event record `ev` of j (ev.x >= 0), the further
// events (if any) from the list.
template <class REAL, class STO, class VT>
__device__ __forceinline__ STO bif_terms_from(const VT& v, int64_t t, int64_t j, int2 ev, STO acc) {
    const int64_t pe = v.pel(t, ev.x >> 1);
    acc += (STO)((ev.x & 1) ? v.path_neg[pe] : v.path_pos[pe]);
    if (ev.y < 0) {
        const int32_t en = v.ev_ptr[j + 1];
        for (int32_t k = v.ev_ptr[j] + 1; k < en; ++k) {
            const int32_t code = v.ev_code[k];
            const int64_t pk = v.pel(t, code >> 1);
            acc += (STO)((code & 1) ? v.path_neg[pk] : v.path_pos[pk]);
        }
    }
    return acc;
}

// per-op: outgoing_storage += bifurcation terms
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_bif_outgoing(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.N * v.T) return;
    int64_t t, j;
    v.split_c(g, t, j);
    if (v.ev_ptr[j] == v.ev_ptr[j + 1]) return;
    const int64_t e = v.el(t, j);
    v.outgoing[e] = bif_event_sum<REAL, STO>(v, t, j, v.outgoing[e], true);
}

// per-op: path nets from the limit_rate field compute_inflow wrote
template <class REAL, class STO>
__global__ __launch_bounds__(kMaxBlock) void k_bif_inflow(View<REAL, STO> v, int substep) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.P * v.T) return;
    int64_t t, p;
    v.split_p(g, t, p);
    const int64_t ci = v.bif_ci[p], di = v.bif_di[p];
    if (!v.active(t, ci, substep)) return;
    const int64_t pe = v.pel(t, p);
    const REAL l0 = v.limit_rate[v.el(t, ci)], l1 = v.limit_rate[v.el(t, di)];
    REAL raw_sum = (REAL)0;
    for (int64_t l = 0; l < v.LB; ++l) {
        const int64_t k = pe * v.LB + l;
        const REAL raw = v.bif_outflow[k];
        raw_sum += raw;
        v.bif_outflow[k] = (raw >= (REAL)0.0) ? raw * l0 : raw * l1;
    }
    v.path_net[pe] = (raw_sum >= (REAL)0.0) ? raw_sum * l0 : raw_sum * l1;
}

// per-op: global_bifurcation_outflow += path nets
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_gbif_sum(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.N * v.T) return;
    int64_t t, j;
    v.split_c(g, t, j);
    if (v.ev_ptr[j] == v.ev_ptr[j + 1]) return;
    const int64_t e = v.el(t, j);
    v.gbif[e] = bif_event_sum<REAL, STO>(v, t, j, v.gbif[e], false);
}

// ============================================================ inflow ====
template <class REAL, class STO>
__device__ __forceinline__ REAL limit_of(const View<REAL, STO>& v, STO outgoing, int64_t e) {
    const REAL og = (REAL)outgoing;
    const REAL rate = (REAL)(v.river_storage[e] + v.flood_storage[e]);
    return (og > (REAL)1e-8) ? vmin(rate / og, (REAL)1.0) : (REAL)1.0;
}

// per-op: apply the storage limiter in place
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_inflow_limit(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.N * v.T) return;
    int64_t t, i;
    v.split_c(g, t, i);
    const int64_t e = v.el(t, i);
    const int64_t ed = v.el(t, v.dn[i]);
    const REAL l_own = limit_of(v, v.outgoing[e], e);
    const REAL l_dn = limit_of(v, v.outgoing[ed], ed);
    const REAL qr = v.river_outflow[e], qf = v.flood_outflow[e];
    v.river_outflow[e] = (qr >= (REAL)0.0) ? qr * l_own : qr * l_dn;
    v.flood_outflow[e] = (qf >= (REAL)0.0) ? qf * l_own : qf * l_dn;
    if (v.has(CMF_HAS_BIFURCATION)) v.limit_rate[e] = l_own;
}

// per-op: river/flood (and reservoir) inflow += limited upstream flows
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_inflow_sum(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= v.N * v.T) return;
    int64_t t, j;
    v.split_c(g, t, j);
    const int2 rg = v.up_range(j);
    const int32_t b = rg.x, en = rg.y;
    if (b == en) return;
    const int64_t e = v.el(t, j);
    STO ri = v.river_inflow[e], fi = v.flood_inflow[e];
    const bool res = v.has(CMF_HAS_RESERVOIR) && v.is_reservoir[j];
    STO rt = res ? v.res_inflow[e] : (STO)0;
    for (int32_t k = b; k < en; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == j) continue;
        const int64_t eu = v.el(t, u);
        const REAL ur = v.river_outflow[eu], uf = v.flood_outflow[eu];
        ri += (STO)ur;
        fi += (STO)uf;
        if (res) rt += (STO)(ur + uf);
    }
    v.river_inflow[e] = ri;
    v.flood_inflow[e] = fi;
    if (res) v.res_inflow[e] = rt;
}

// ======================================================= flood stage ====
struct LogPart {
    double q[CMF_LOG_COUNT];
};

// Floodplain depth table row of one catchment.  With LFIX == 10 (the level
// count of every configuration, known at compile time) the whole row is
// loaded into ten named registers up front, one round trip, and read through
// a switch that folds to a register once the level loops are unrolled (an
// array here would be placed in LDS or scratch, since the array promotion
// runs before unrolling); other counts read the row level by level.
template <class REAL, int LFIX>
struct TableRow {
    const REAL* row;
    REAL h0, h1, h2, h3, h4, h5, h6, h7, h8, h9;
    __device__ __forceinline__ void load(const REAL* r) {
        row = r;
        if (LFIX == 10) {
            h0 = r[0]; h1 = r[1]; h2 = r[2]; h3 = r[3]; h4 = r[4];
            h5 = r[5]; h6 = r[6]; h7 = r[7]; h8 = r[8]; h9 = r[9];
        }
    }
    __device__ __forceinline__ REAL operator[](int k) const {
        if (LFIX != 10) return row[k];
        switch (k) {
            case 0: return h0;
            case 1: return h1;
            case 2: return h2;
            case 3: return h3;
            case 4: return h4;
            case 5: return h5;
            case 6: return h6;
            case 7: return h7;
            case 8: return h8;
            default: return h9;
        }
    }
};

template <class REAL, class STO>
struct StageOut {
    STO total_hp, rsf_hp, fsf_hp;
    REAL rdep, fdep, ffr, ca;
};

// Storage update of one catchment (storage.py:81-94): the new total storage
// and, with LOG, the sums that depend only on it.
template <class REAL, class STO>
struct StagePre {
    STO total_hp;
    STO fs_old, ps_old;
};

template <class REAL, class STO, bool LOG>
__device__ __forceinline__ StagePre<REAL, STO> stage_update_p(
    STO rs, STO fs, STO ps, REAL ro, REAL rin, REAL fin, REAL rout, REAL fout, REAL gbif, REAL time_step,
    LogPart& lp) {
    StagePre<REAL, STO> o;
    o.fs_old = fs;
    o.ps_old = ps;
    const REAL inflow_extra = (REAL)0.0;

    STO rsu = rs + (STO)((rin - rout) * time_step);
    STO fsu = fs + ((rsu < (STO)0.0) ? rsu : (STO)0.0) + (STO)((fin - fout - gbif) * time_step);
    rsu = vmax(rsu, (STO)0.0);
    rsu = (fsu < (STO)0.0) ? vmax(rsu + fsu, (STO)0.0) : rsu;
    fsu = vmax(fsu, (STO)0.0);
    const STO total_next = rsu + fsu + ps + (STO)(ro * time_step) + (STO)(inflow_extra * time_step);
    o.total_hp = vmax(total_next, (STO)0.0);
    if (LOG) {
        const STO total_pre = rs + fs + ps;
        lp.q[CMF_LOG_STORAGE_PRE] += (double)total_pre;
        lp.q[CMF_LOG_STORAGE_NEXT] += (double)total_next;
        lp.q[CMF_LOG_STORAGE_NEW] += (double)o.total_hp;
        lp.q[CMF_LOG_INFLOW] += (double)((rin + fin + inflow_extra) * time_step);
        lp.q[CMF_LOG_OUTFLOW] += (double)((rout + fout) * time_step);
        lp.q[CMF_LOG_INFLOW_ERROR] += (double)(total_pre - total_next
            + (STO)((rin + fin + ro + inflow_extra - rout - fout - gbif) * time_step));
    }
    return o;
}

template <class REAL, class STO, bool LOG>
__device__ __forceinline__ StagePre<REAL, STO> stage_update(
    const View<REAL, STO>& v, int64_t e, int64_t i, REAL rin, REAL fin, REAL rout, REAL fout, REAL gbif,
    REAL time_step, LogPart& lp) {
    return stage_update_p<REAL, STO, LOG>(v.river_storage[e], v.flood_storage[e], v.protected_storage[e],
                                          v.prm(v.runoff, CMF_BATCHED_RUNOFF, e, i), rin, fin, rout, fout, gbif,
                                          time_step, lp);
}

// Whether the new total exceeds the channel (the table scan is needed).
template <class REAL, class STO>
__device__ __forceinline__ bool stage_is_wet(const View<REAL, STO>& v, int64_t e, int64_t i, STO total_hp) {
    const REAL rh = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, i);
    const REAL rw = v.prm(v.river_width, CMF_BATCHED_RIVER_WIDTH, e, i);
    const REAL rl = v.prm(v.river_length, CMF_BATCHED_RIVER_LENGTH, e, i);
    return (REAL)total_hp > rl * rw * rh;
}

// Flood-stage diagnosis of one catchment from its new total storage
// (storage.py:105-181), with LOG the sums of the stage (non-levee catchments).
// `ca_pre` >= 0 is the catchment area already read (else read when needed).
template <class REAL, class STO, int LFIX, bool LOG, class VT>
__device__ __forceinline__ StageOut<REAL, STO> stage_diag_p(
    const VT& v, int64_t e, int64_t i, STO total_hp, REAL rl, REAL rw, REAL rh, REAL ca_pre,
    bool non_levee, TableRow<REAL, LFIX>& row, bool& row_loaded, LogPart& lp) {
    StageOut<REAL, STO> o;
    const REAL total = (REAL)total_hp;
    o.total_hp = total_hp;
    const REAL river_max_storage = rl * rw * rh;
    const int L = LFIX > 0 ? LFIX : (int)v.L;
    o.ca = (REAL)0;
    if (!(total > river_max_storage)) {
        // dry: the table scan leaves level = -1 (storage.py:110,137)
        o.rsf_hp = total_hp;
        o.fdep = (REAL)0.0;
        o.ffr = (REAL)0.0;
        o.rdep = (REAL)o.rsf_hp / (rl * rw);
        o.fsf_hp = vmax(total_hp - o.rsf_hp, (STO)0.0);
        if (LOG) o.ca = ca_pre >= (REAL)0 ? ca_pre : v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, i);
    } else {
        o.ca = ca_pre >= (REAL)0 ? ca_pre : v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, i);
        const REAL ca = o.ca;
        if (!row_loaded) {
            row.load(v.bat(CMF_BATCHED_FLOOD_DEPTH_TABLE) ? v.table + e * v.L : v.table + i * v.L);
            row_loaded = true;
        }
        const REAL catchment_width = ca / rl;
        const REAL width_increment = catchment_width / (REAL)L;
        int level = 0;
        REAL S_accum = river_max_storage;
        REAL prev_H = (REAL)0.0;
        REAL prev_W = rw;
        REAL prev_total_storage = river_max_storage;
        REAL prev_flood_depth = (REAL)0.0;
        REAL next_flood_depth = (REAL)0.0;
        // S is non-decreasing, so once total <= S no later level is above:
        // stopping there leaves the values the full scan would.
#pragma unroll
        for (int k = 0; k < (LFIX > 0 ? LFIX : 64); ++k) {
            if (LFIX == 0 && k >= L) break;
            const REAL H = row[k];
            const REAL W = rw + (REAL)(k + 1) * width_increment;
            const REAL dS = rl * (REAL)0.5 * (prev_W + W) * (H - prev_H);
            const REAL S = S_accum + dS;
            next_flood_depth = H;
            if (!(total > S)) break;
            level += 1;
            prev_total_storage = S;
            prev_flood_depth = H;
            S_accum = S;
            prev_H = H;
            prev_W = W;
        }
        const REAL prev_total_width = rw + (REAL)level * width_increment;
        if (level == L) {
            o.fdep = prev_flood_depth + (total - prev_total_storage) / (prev_total_width * rl);
            o.ffr = (REAL)1.0;
        } else {
            const REAL flood_grad = (next_flood_depth - prev_flood_depth) / width_increment;
            const REAL diff_width = dsqrt(prev_total_width * prev_total_width
                + (REAL)2.0 * (total - prev_total_storage) / (flood_grad * rl)) - prev_total_width;
            o.fdep = prev_flood_depth + diff_width * flood_grad;
            o.ffr = vclamp((prev_total_width + diff_width - rw) * rl / ca, (REAL)0.0, (REAL)1.0);
        }
        const REAL cap = river_max_storage + rl * rw * o.fdep;
        o.rsf_hp = vmin((STO)cap, total_hp);
        o.rdep = (REAL)o.rsf_hp / (rl * rw);
        o.fsf_hp = vmax(total_hp - o.rsf_hp, (STO)0.0);
    }

    if (LOG && non_levee) {
        const STO stage_new = o.rsf_hp + o.fsf_hp;
        lp.q[CMF_LOG_STORAGE_STAGE] += (double)stage_new;
        lp.q[CMF_LOG_RIVER_STORAGE] += (double)(REAL)o.rsf_hp;
        lp.q[CMF_LOG_FLOOD_STORAGE] += (double)(REAL)o.fsf_hp;
        lp.q[CMF_LOG_FLOOD_AREA] += (double)(o.ffr * o.ca);
        lp.q[CMF_LOG_STAGE_ERROR] += (double)(stage_new - total_hp);
    }
    return o;
}

template <class REAL, class STO, int LFIX, bool LOG>
__device__ __forceinline__ StageOut<REAL, STO> stage_diag(
    const View<REAL, STO>& v, int64_t e, int64_t i, STO total_hp, bool non_levee, TableRow<REAL, LFIX>& row,
    bool& row_loaded, LogPart& lp) {
    return stage_diag_p<REAL, STO, LFIX, LOG>(
        v, e, i, total_hp, v.prm(v.river_length, CMF_BATCHED_RIVER_LENGTH, e, i),
        v.prm(v.river_width, CMF_BATCHED_RIVER_WIDTH, e, i), v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, i),
        (REAL)-1, non_levee, row, row_loaded, lp);
}

// Storage update and flood-stage diagnosis of one catchment (values only).
template <class REAL, class STO, int LFIX, bool LOG>
__device__ __forceinline__ StageOut<REAL, STO> stage_one(
    const View<REAL, STO>& v, int64_t e, int64_t i, REAL rin, REAL fin, REAL rout, REAL fout, REAL gbif,
    REAL time_step, bool non_levee, TableRow<REAL, LFIX>& row, bool& row_loaded, LogPart& lp) {
    const StagePre<REAL, STO> pre = stage_update<REAL, STO, LOG>(v, e, i, rin, fin, rout, fout, gbif, time_step, lp);
    return stage_diag<REAL, STO, LFIX, LOG>(v, e, i, pre.total_hp, non_levee, row, row_loaded, lp);
}

// ------------------------------------------------------ levee partition --
template <class REAL, class STO>
struct LeveeOut {
    STO r_hp, f_hp, p_hp;
    REAL r_dph, f_dph, p_dph, f_frc;
};

// What the reverse of the levee stage needs of one evaluation: the case, the
// storages the cases compare against, the case-3 operands, and whether the
// case-3 fill storage moves with the bankfull storage.
template <class REAL>
struct LeveeTrace {
    bool case2, case3, case4, none, found_B, fil_moves;
    REAL total, rs_c, base_sto, s_top, levee_distance, dsto_fil_B, dwth_fil_B, gradient_B, term_B;
    REAL r_cand, f_cand, f_cand_raw;
};

// Levee stage of catchment c (levee row lv) from its post-flood-stage state
// (triton/levee.py:62-304 with the partition of :18-59).  TR: also record
// the evaluation in *tr (the reverse's recompute; the arithmetic is the same).
template <class REAL, class STO, int LFIX, bool TR = false>
__device__ __forceinline__ LeveeOut<REAL, STO> levee_one_p(
    const View<REAL, STO>& v, REAL rl, REAL rw, REAL rh, REAL ca, REAL crown, REAL lfrac, REAL base,
    STO rs_hp, STO fs_hp, REAL fdep_curr, REAL rdep_curr, REAL ffr_curr, const TableRow<REAL, LFIX>& row,
    LeveeTrace<REAL>* tr = nullptr) {
    const int L = LFIX > 0 ? LFIX : (int)v.L;
    const STO total_hp = rs_hp + fs_hp;
    const REAL rs_c = (REAL)rs_hp, fs_c = (REAL)fs_hp, total = (REAL)total_hp;

    const REAL river_max_storage = rl * rw * rh;
    // Below bankfull none of the cases applies: the base and crown storages
    // formed below are river_max_storage plus non-negative terms, and with
    // crown >= base the protected-side threshold s_top is not below the base
    // storage, so all three tests are false and the partition keeps the state
    // (the "none" branch at the end).  With negative geometry the full scan
    // decides.
    if (total < river_max_storage && crown >= base && rl > (REAL)0 && rw >= (REAL)0 && ca >= (REAL)0 &&
        lfrac >= (REAL)0) {
        LeveeOut<REAL, STO> o;
        o.r_dph = rdep_curr;
        o.f_dph = fdep_curr;
        o.p_dph = (REAL)0.0;
        o.f_frc = ffr_curr;
        const STO remaining = vmax(total_hp - rs_hp, (STO)0.0);
        o.r_hp = rs_hp;
        o.f_hp = fs_hp;
        o.p_hp = vmax(remaining - fs_hp, (STO)0.0);
        if (TR) {
            tr->case2 = tr->case3 = tr->case4 = false;
            tr->none = true;
            tr->total = total;
            tr->rs_c = rs_c;
        }
        return o;
    }
    const REAL dwth_inc = (ca / rl) / (REAL)L;
    const REAL levee_distance = lfrac * (ca / rl);
    const int ilev = (int)(lfrac * (REAL)L);
    // The scan of levee.py:149-211.  The quotients it takes at one level only
    // (the base and crown interpolation ratios, the gradient of the level where
    // the protected side stops filling) are formed after the loop from the
    // operands recorded at that level: the same operations on the same values,
    // one division each instead of one per level.
    REAL s_curr = river_max_storage;
    REAL dhgt_pre = (REAL)0.0;
    REAL dwth_pre = rw;
    bool found_base = false, found_fill = false;
    REAL b_s = (REAL)0.0, b_pre = (REAL)0.0, b_seg = (REAL)1.0, b_w = (REAL)0.0;   // base level operands
    REAL f_s = (REAL)0.0, f_pre = (REAL)0.0, f_seg = (REAL)1.0, f_w = (REAL)0.0;   // crown level operands
    bool ib_found = false;                                                         // base found by level ilev
    REAL ib_s = (REAL)0.0, ib_pre = (REAL)0.0, ib_seg = (REAL)1.0, ib_w = (REAL)0.0;
    REAL dsto_fil_B = (REAL)0.0, dwth_fil_B = (REAL)0.0, ddph_fil_B = (REAL)0.0, seg_B = (REAL)0.0;
    bool found_B = false, B_at_ilev = false, fil_moves = false;
#pragma unroll
    for (int k = 0; k < (LFIX > 0 ? LFIX : 64); ++k) {
        if (LFIX == 0 && k >= L) break;
        const REAL depth_val = row[k];
        REAL dhgt_seg = depth_val - dhgt_pre;
        dhgt_seg = vmax(dhgt_seg, (REAL)1e-6);
        const REAL dwth_mid = dwth_pre + (REAL)0.5 * dwth_inc;
        const REAL dsto_seg = rl * dwth_mid * dhgt_seg;
        const REAL s_next = s_curr + dsto_seg;
        if ((base > dhgt_pre) && (base <= depth_val)) {
            b_s = s_curr; b_pre = dhgt_pre; b_seg = dhgt_seg; b_w = dwth_pre;
            found_base = true;
        }
        if ((crown > dhgt_pre) && (crown <= depth_val)) {
            f_s = s_curr; f_pre = dhgt_pre; f_seg = dhgt_seg; f_w = dwth_pre;
            found_fill = true;
        }
        if (k == ilev) {
            ib_found = found_base; ib_s = b_s; ib_pre = b_pre; ib_seg = b_seg; ib_w = b_w;
        }
        if ((k >= ilev) && !found_B) {
            const REAL wedge = (levee_distance + rw) * (crown - depth_val) * rl;
            const REAL threshold = s_next + wedge;
            if (total < threshold) {
                B_at_ilev = (k == ilev);
                seg_B = dhgt_seg;
                found_B = true;
            } else {
                dsto_fil_B = threshold;
                dwth_fil_B = dwth_inc * (REAL)(k + 1) - levee_distance;
                ddph_fil_B = depth_val - base;
                fil_moves = true;
            }
        }
        s_curr = s_next;
        dhgt_pre = depth_val;
        dwth_pre += dwth_inc;
    }
    // storage at a height inside level (s, pre, seg, w): levee.py:163-170,178-185
    auto interp_sto = [&](REAL h, REAL s, REAL pre, REAL seg, REAL w) -> REAL {
        const REAL ratio = (h - pre) / seg;
        return s + rl * (w + (REAL)0.5 * ratio * dwth_inc) * (ratio * seg);
    };
    const REAL s_base_extra = s_curr + rl * dwth_pre * (base - dhgt_pre);
    const REAL base_sto = found_base ? interp_sto(base, b_s, b_pre, b_seg, b_w)
                                     : ((base > dhgt_pre) ? s_base_extra : river_max_storage);
    const REAL s_fill_extra = s_curr + rl * dwth_pre * (crown - dhgt_pre);
    const REAL fill_sto = found_fill ? interp_sto(crown, f_s, f_pre, f_seg, f_w)
                                     : ((crown > dhgt_pre) ? s_fill_extra : river_max_storage);
    const REAL gradient_B = found_B ? seg_B / dwth_inc : (REAL)0.0;
    if (B_at_ilev) {
        const REAL base_sto_ilev = ib_found ? interp_sto(base, ib_s, ib_pre, ib_seg, ib_w) : river_max_storage;
        dsto_fil_B = base_sto_ilev + (levee_distance + rw) * (crown - base) * rl;
        fil_moves = true;
    }
    const REAL dhgt_dif = crown - base;
    const REAL s_top = base_sto + (levee_distance + rw) * dhgt_dif * rl;
    const bool case4 = total >= fill_sto;
    const bool case3 = !case4 && (total >= s_top);
    const bool case2 = !case4 && !case3 && (total >= base_sto);

    LeveeOut<REAL, STO> o;
    REAL r_cand = rs_c, f_cand = fs_c, f_cand_raw = fs_c;
    REAL term_B_rec = (REAL)0;
    o.r_dph = rdep_curr;
    o.f_dph = fdep_curr;
    o.p_dph = (REAL)0.0;
    o.f_frc = ffr_curr;
    if (case2) {
        const REAL dsto_add_c2 = total - base_sto;
        const REAL dwth_add_c2 = levee_distance + rw;
        const REAL f_dph_c2 = base + dsto_add_c2 / dwth_add_c2 / rl;
        const REAL r_sto_c2 = river_max_storage + rl * rw * f_dph_c2;
        r_cand = r_sto_c2;
        f_cand_raw = total - r_sto_c2;
        f_cand = vmax(f_cand_raw, (REAL)0.0);
        o.r_dph = r_sto_c2 / rl / rw;
        o.f_dph = f_dph_c2;
        o.f_frc = lfrac;
    } else if (case3) {
        const REAL dsto_add_B = total - dsto_fil_B;
        REAL p_dph_B, f_frc_B;
        if (found_B) {
            const REAL term_B = dwth_fil_B * dwth_fil_B + (REAL)2.0 * dsto_add_B / rl / (gradient_B + (REAL)1e-9);
            term_B_rec = term_B;
            const REAL dwth_add_B = -dwth_fil_B + dsqrt(vmax(term_B, (REAL)0.0));
            const REAL ddph_add_B = dwth_add_B * gradient_B;
            p_dph_B = base + ddph_fil_B + ddph_add_B;
            f_frc_B = (dwth_fil_B + levee_distance) / (dwth_inc * (REAL)L);
        } else {
            const REAL ddph_add_B_extra = dsto_add_B / (dwth_fil_B * rl + (REAL)1e-9);
            p_dph_B = base + ddph_fil_B + ddph_add_B_extra;
            f_frc_B = (REAL)1.0;
        }
        const REAL r_sto_c3 = river_max_storage + rl * rw * crown;
        r_cand = r_sto_c3;
        f_cand_raw = s_top - r_sto_c3;
        f_cand = vmax(f_cand_raw, (REAL)0.0);
        o.r_dph = r_sto_c3 / rl / rw;
        o.f_dph = crown;
        o.p_dph = p_dph_B;
        o.f_frc = vclamp(f_frc_B, (REAL)0.0, (REAL)1.0);
    } else if (case4) {
        const REAL dsto_add_c4 = (fdep_curr - crown) * (levee_distance + rw) * rl;
        r_cand = rs_c;
        f_cand_raw = s_top + dsto_add_c4 - rs_c;
        f_cand = vmax(f_cand_raw, (REAL)0.0);
        o.p_dph = fdep_curr;
    }
    const bool none = !(case2 || case3 || case4);
    if (TR) {
        tr->case2 = case2;
        tr->case3 = case3;
        tr->case4 = case4;
        tr->none = none;
        tr->found_B = found_B;
        tr->fil_moves = fil_moves;
        tr->total = total;
        tr->rs_c = rs_c;
        tr->base_sto = base_sto;
        tr->s_top = s_top;
        tr->levee_distance = levee_distance;
        tr->dsto_fil_B = dsto_fil_B;
        tr->dwth_fil_B = dwth_fil_B;
        tr->gradient_B = gradient_B;
        tr->term_B = term_B_rec;
        tr->r_cand = r_cand;
        tr->f_cand = f_cand;
        tr->f_cand_raw = f_cand_raw;
    }
    STO r_hp = vmin((STO)r_cand, total_hp);
    r_hp = (case4 || none) ? rs_hp : r_hp;
    const STO remaining = vmax(total_hp - r_hp, (STO)0.0);
    STO f_hp = vmin(vmax((STO)f_cand, (STO)0.0), remaining);
    f_hp = case2 ? remaining : (none ? fs_hp : f_hp);
    o.r_hp = r_hp;
    o.f_hp = f_hp;
    o.p_hp = vmax(remaining - f_hp, (STO)0.0);
    return o;
}

// levee_one_p with the catchment and levee parameters read from the fields
template <class REAL, class STO, int LFIX>
__device__ __forceinline__ LeveeOut<REAL, STO> levee_one(
    const View<REAL, STO>& v, int64_t e, int64_t c, int64_t lv, int64_t le, STO rs_hp, STO fs_hp,
    REAL fdep_curr, REAL rdep_curr, REAL ffr_curr, const TableRow<REAL, LFIX>& row) {
    return levee_one_p<REAL, STO, LFIX>(
        v, v.prm(v.river_length, CMF_BATCHED_RIVER_LENGTH, e, c), v.prm(v.river_width, CMF_BATCHED_RIVER_WIDTH, e, c),
        v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, c), v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, c),
        v.prm(v.levee_crown, CMF_BATCHED_LEVEE_CROWN_HEIGHT, le, lv), v.prm(v.levee_frac, CMF_BATCHED_LEVEE_FRACTION, le, lv),
        v.prm(v.levee_base, CMF_BATCHED_LEVEE_BASE_HEIGHT, le, lv), rs_hp, fs_hp, fdep_curr, rdep_curr, ffr_curr, row);
}

template <class REAL, class STO>
__device__ __forceinline__ void levee_log(const LeveeOut<REAL, STO>& o, STO total_hp, REAL ca, LogPart& lp) {
    const STO stage_new = o.r_hp + o.f_hp + o.p_hp;
    lp.q[CMF_LOG_STORAGE_STAGE] += (double)stage_new;
    lp.q[CMF_LOG_RIVER_STORAGE] += (double)o.r_hp;
    lp.q[CMF_LOG_FLOOD_STORAGE] += (double)o.f_hp;
    lp.q[CMF_LOG_FLOOD_AREA] += (double)(o.f_frc * ca);
    lp.q[CMF_LOG_STAGE_ERROR] += (double)(stage_new - total_hp);
}

// Block-level FP64 sums of the log partials in a fixed order.
__device__ __forceinline__ void block_log_partials(LogPart& lp, double* out_block) {
    __shared__ double sh[kMaxBlock / 64][CMF_LOG_COUNT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < CMF_LOG_COUNT; ++k) {
        const double s = wave_sum_double(lp.q[k]);
        if (lane == 0) sh[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < CMF_LOG_COUNT) {
        double acc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) acc += sh[w][threadIdx.x];
        out_block[threadIdx.x] = acc;
    }
}

// Store the post-stage state of one catchment.  `fin`: also the fields no
// later pass of the sub-step loop reads for a catchment without levee (flood
// fraction, protected depth), which the fused schedule writes on the
// catchment's last sub-step only.
template <class REAL, class STO>
__device__ __forceinline__ void store_state(const View<REAL, STO>& v, int64_t e, const StageOut<REAL, STO>& so,
                                            STO fs_old, STO ps_old, bool fin = true) {
    v.river_storage[e] = so.rsf_hp;
    store_if_changed(v.flood_storage + e, fs_old, so.fsf_hp);
    store_if_changed(v.protected_storage + e, ps_old, (STO)0.0);
    v.river_depth[e] = so.rdep;
    v.flood_depth[e] = so.fdep;
    if (fin) {
        v.protected_depth[e] = so.fdep;
        v.flood_fraction[e] = so.ffr;
    }
}

// per-op compute_flood_stage(_log): storage update and flood stage from
// the inflow fields (storage.py:16-193, :196-425).
template <class REAL, class STO, bool LOG, int LFIX>
__global__ __launch_bounds__(kBlock) void k_stage(View<REAL, STO> v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LogPart lp;
    if (LOG) {
#pragma unroll
        for (int k = 0; k < CMF_LOG_COUNT; ++k) lp.q[k] = 0.0;
    }
    if (g < v.N * v.T) {
        int64_t t, j;
        v.split_c(g, t, j);
        const int64_t e = v.el(t, j);
        const REAL rin = (REAL)v.river_inflow[e];
        const REAL fin = (REAL)v.flood_inflow[e];
        const REAL gb = v.has(CMF_HAS_BIFURCATION) ? (REAL)v.gbif[e] : (REAL)0.0;
        const bool is_lev = v.has(CMF_HAS_LEVEE) && v.is_levee[j];
        TableRow<REAL, LFIX> row;
        bool row_loaded = false;
        const STO fs_old = v.flood_storage[e], ps_old = v.protected_storage[e];
        const StageOut<REAL, STO> so = stage_one<REAL, STO, LFIX, LOG>(
            v, e, j, rin, fin, v.river_outflow[e], v.flood_outflow[e], gb, v.dt(t, j), !is_lev, row, row_loaded, lp);
        v.outgoing[e] = (STO)0.0;
        store_state(v, e, so, fs_old, ps_old);
    }
    if (LOG) block_log_partials(lp, v.partials + (int64_t)blockIdx.x * CMF_LOG_COUNT);
}

// fused: storage limiter of every catchment from the terms of the outflow
// pass (outgoing_storage of compute_outflow, then compute_inflow's limiter,
// outflow.py:346-352) into the limiter scratch; the limit_rate field, which no
// later pass reads, on the catchment's last sub-step only (`store_last`).
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kMaxBlock) void k_limiter(View<REAL, STO> v, int substep, int store_last) {
    int64_t t, j;
    if (!v.catchment_of(xcd_block() * blockDim.x + threadIdx.x, t, j)) return;
    if (!v.active_in_block(t, j, substep)) return;
    const int64_t e = v.el(t, j);
    const int2 ev = v.has(CMF_HAS_BIFURCATION) ? v.ev1[j] : make_int2(-1, 0);
    STO og = outgoing_terms<REAL, STO, CONTIG>(v, t, j);
    if (ev.x >= 0) og = bif_terms_from<REAL, STO>(v, t, j, ev, og);
    const REAL l = limit_of(v, og, e);
    v.lim[e] = l;
    if (v.has(CMF_HAS_BIFURCATION) && v.last_substep(t, j, substep, store_last)) v.limit_rate[e] = l;
}

// The same, two catchments per thread (g and g + blockDim.x of a block's
// 2 * blockDim.x) with their loads interleaved: the pass is bound by
// dependent round trips (upstream range, then its terms), so two independent
// chains per thread keep twice the loads in flight.  With GroupView the
// block is one record of the group block table, blockDim.x half its width.
template <class REAL, class STO, bool CONTIG, class VT = View<REAL, STO>>
__global__ __launch_bounds__(kMaxBlock) void k_limiter_pair(VT v, int substep, int store_last) {
    int64_t t0 = 0, j0 = 0, t1 = 0, j1 = 0;
    bool a0, a1;
    if constexpr (VT::kGrouped) {
        const int4 r = v.rec();
        j0 = (int64_t)r.x + threadIdx.x;
        j1 = j0 + blockDim.x;
        const bool on = v.active_in_block(0, j0, substep);
        a0 = on && j0 < r.y;
        a1 = on && j1 < r.y;
    } else {
        const int64_t NT = v.N * v.T;
        const int64_t g0 = (int64_t)blockIdx.x * (2 * (int64_t)blockDim.x) + threadIdx.x;
        const int64_t g1 = g0 + blockDim.x;
        a0 = g0 < NT;
        a1 = g1 < NT;
        if (a0) {
            v.split_c(g0, t0, j0);
            a0 = v.active(t0, j0, substep);
        }
        if (a1) {
            v.split_c(g1, t1, j1);
            a1 = v.active(t1, j1, substep);
        }
    }
    if (!a0 && !a1) return;
    const int64_t e0 = v.el(t0, j0), e1 = v.el(t1, j1);
    int32_t b0 = 0, n0 = 0, b1 = 0, n1 = 0;
    REAL p0 = (REAL)0, p1 = (REAL)0;
    STO rs0 = (STO)0, fs0 = (STO)0, rs1 = (STO)0, fs1 = (STO)0;
    int2 ev0 = make_int2(-1, 0), ev1 = make_int2(-1, 0);
    if (v.has(CMF_HAS_BIFURCATION)) {
        if (a0) ev0 = v.ev1[j0];
        if (a1) ev1 = v.ev1[j1];
    }
    if (a0) {
        const int2 r0 = v.up_range(j0);
        b0 = r0.x;
        n0 = r0.y - r0.x;
        p0 = v.pos_term[e0];
        rs0 = v.river_storage[e0];
        fs0 = v.flood_storage[e0];
    }
    if (a1) {
        const int2 r1 = v.up_range(j1);
        b1 = r1.x;
        n1 = r1.y - r1.x;
        p1 = v.pos_term[e1];
        rs1 = v.river_storage[e1];
        fs1 = v.flood_storage[e1];
    }
    // outgoing_terms of both, one loop so that both ranges' loads issue together
    STO acc0 = (STO)0, acc1 = (STO)0;
    const int32_t nmax = n0 > n1 ? n0 : n1;
    for (int32_t k = 0; k < nmax; ++k) {
        if (k < n0) {
            const int64_t u = CONTIG ? (int64_t)(b0 + k) : (int64_t)v.up_idx[b0 + k];
            if (!(CONTIG && u == j0)) acc0 += (STO)v.neg_term[v.el(t0, u)];
        }
        if (k < n1) {
            const int64_t u = CONTIG ? (int64_t)(b1 + k) : (int64_t)v.up_idx[b1 + k];
            if (!(CONTIG && u == j1)) acc1 += (STO)v.neg_term[v.el(t1, u)];
        }
    }
    acc0 += (STO)p0;
    acc1 += (STO)p1;
    if (ev0.x >= 0) acc0 = bif_terms_from<REAL, STO>(v, t0, j0, ev0, acc0);
    if (ev1.x >= 0) acc1 = bif_terms_from<REAL, STO>(v, t1, j1, ev1, acc1);
    // limit_of (outflow.py:346-352) of both
    const REAL og0 = (REAL)acc0, og1 = (REAL)acc1;
    const REAL r0 = (REAL)(rs0 + fs0), r1 = (REAL)(rs1 + fs1);
    const REAL l0 = (og0 > (REAL)1e-8) ? vmin(r0 / og0, (REAL)1.0) : (REAL)1.0;
    const REAL l1 = (og1 > (REAL)1e-8) ? vmin(r1 / og1, (REAL)1.0) : (REAL)1.0;
    if (a0) {
        v.lim[e0] = l0;
        if (v.has(CMF_HAS_BIFURCATION) && v.last_substep(t0, j0, substep, store_last)) v.limit_rate[e0] = l0;
    }
    if (a1) {
        v.lim[e1] = l1;
        if (v.has(CMF_HAS_BIFURCATION) && v.last_substep(t1, j1, substep, store_last)) v.limit_rate[e1] = l1;
    }
}

// The storage limiter of element (t, x) in full (outflow.py:306-352): the
// negative terms of its upstream catchments (range rg) in ascending index,
// its own positive term, then its path terms (events [eb, ee) of the event
// list, in list order as bif_terms_from, read from the event slots the path
// pass filled: eterm), each catchment term formed from the unlimited flows by
// the outflow pass's own expressions, so that every sum is the limiter
// pass's.  An upstream catchment u of x drains into x != u, so it is no
// mouth.  Bound by dependent round trips: with contiguous ranges the flows
// of the first kHoist upstream catchments and the first two path terms are
// loaded with the catchment's own fields, before any sum.
template <class REAL, class STO, bool CONTIG, class VT>
__device__ __forceinline__ REAL limiter_exact(const VT& v, int64_t t, int64_t x, int2 rg, int32_t eb, int32_t ee) {
    constexpr int kHoist = 4;
    const int64_t e = v.el(t, x);
    const REAL dt = v.dt(t, x);
    const STO rs = v.river_storage[e], fs = v.flood_storage[e];
    const REAL qr = v.qraw_r[e], qf = v.qraw_f[e];
    REAL hr[kHoist], hf[kHoist];
    if (CONTIG) {
#pragma unroll
        for (int k = 0; k < kHoist; ++k) {
            const int64_t u = (int64_t)rg.x + k;
            const bool in = u < (int64_t)rg.y && u != x;
            const int64_t eu = v.el(t, in ? u : x);
            hr[k] = in ? v.qraw_r[eu] : (REAL)0.0;
            hf[k] = in ? v.qraw_f[eu] : (REAL)0.0;
        }
    }
    const REAL p0 = (eb < ee) ? v.eterm[v.evel(t, eb)] : (REAL)0.0;
    const REAL p1 = (eb + 1 < ee) ? v.eterm[v.evel(t, eb + 1)] : (REAL)0.0;
    STO acc = (STO)0;
    for (int32_t k = rg.x; k < rg.y; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == x) continue;
        REAL ur, uf;
        if (CONTIG && k - rg.x < kHoist) {
            // registers of the hoisted loads (k - rg.x is the slot)
            ur = (REAL)0.0;
            uf = (REAL)0.0;
#pragma unroll
            for (int h = 0; h < kHoist; ++h)
                if (h == k - rg.x) { ur = hr[h]; uf = hf[h]; }
        } else {
            const int64_t eu = v.el(t, u);
            ur = v.qraw_r[eu];
            uf = v.qraw_f[eu];
        }
        const REAL neg = vmin(ur, (REAL)0.0) + vmin(uf, (REAL)0.0);
        acc += (STO)(-neg * dt);
    }
    const REAL pos = vmax(qr, (REAL)0.0) + vmax(qf, (REAL)0.0);
    acc += (STO)(pos * dt);
    for (int32_t k = eb; k < ee; ++k) acc += (STO)((k == eb) ? p0 : (k == eb + 1) ? p1 : v.eterm[v.evel(t, k)]);
    const REAL og = (REAL)acc;
    const REAL rate = (REAL)(rs + fs);
    return (og > (REAL)1e-8) ? vmin(rate / og, (REAL)1.0) : (REAL)1.0;
}

template <class REAL, class STO, class VT>
__device__ __forceinline__ void store_limiter(const VT& v, int64_t t, int64_t x, REAL l, int substep,
                                              int store_last) {
    const int64_t e = v.el(t, x);
    v.lim[e] = l;
    if (v.has(CMF_HAS_BIFURCATION) && v.last_substep(t, x, substep, store_last)) v.limit_rate[e] = l;
}

// Fix-up pass of the provisional-limiter schedule, after the outflow pass
// (PROV), before the stage pass: the full limiter of every path end and of
// every element an upstream backflow flagged; every other limiter is exact
// as the outflow pass left it.
//   blocks [0, list_blocks): the path ends, once per trial, from the static
//     records `list` ({catchment, upstream range start, end, first event};
//     entry list_n holds the end of the last one's events in .w), so that
//     the catchment's fields, its upstream flows and its path terms (eterm)
//     are one round trip away;
//   the next res_blocks: the reservoir catchments, once per trial, whose
//     release k_reservoir wrote over the flows the outflow pass formed the
//     provisional limiter from;
//   the other blocks: 32-element words of the flag bits, one per lane (the
//     whole network, or with `rec` the words of one catchment record per
//     block of one wave: the group records of the sub-step), cleared after
//     use, the flagged elements dealt out one per lane of the wave; path
//     ends among them are left to the list blocks (the limiter is formed
//     before that test so that its loads issue with the test's).  A word two
//     records share may be read by both: each of its flags is then formed
//     twice, to the same value.  VT must give any element its own time-step
//     slot (not GroupView, whose slot is the block's record's).
template <class REAL, class STO, bool CONTIG, class VT>
__global__ __launch_bounds__(kMaxBlock) void k_limiter_fix(VT v, int substep, int store_last,
                                                        const int4* __restrict__ list, int64_t list_n,
                                                        int64_t list_blocks, int64_t res_blocks,
                                                        const int4* __restrict__ rec) {
    int64_t b = blockIdx.x;
    if (b >= list_blocks && b < list_blocks + res_blocks) {
        const int64_t g = (b - list_blocks) * blockDim.x + threadIdx.x;
        if (g >= v.R * v.T) return;
        const int64_t t = g / v.R, r = g - t * v.R;
        const int64_t x = v.res_idx[r];
        if (!v.active(t, x, substep)) return;
        const bool ev = v.has(CMF_HAS_BIFURCATION) && v.eterm;
        const int32_t eb = ev ? v.ev_ptr[x] : 0, ee = ev ? v.ev_ptr[x + 1] : 0;
        const REAL l = limiter_exact<REAL, STO, CONTIG>(v, t, x, v.up_range(x), eb, ee);
        store_limiter<REAL, STO>(v, t, x, l, substep, store_last);
        return;
    }
    if (b >= list_blocks) b -= res_blocks;
    if (b < list_blocks) {
        const int64_t g = b * blockDim.x + threadIdx.x;
        if (g >= list_n * v.T) return;
        int64_t t, k;
        v.list_split(g, t, k);
        const int4 r = list[k];
        const int32_t ee = list[k + 1].w;
        if (!v.active(t, r.x, substep)) return;
        const REAL l = limiter_exact<REAL, STO, CONTIG>(v, t, r.x, make_int2(r.y, r.z), r.w, ee);
        store_limiter<REAL, STO>(v, t, r.x, l, substep, store_last);
        return;
    }
    // the flagged elements of the wave's words are dealt out one per lane
    // (wave-wide prefix of the bit counts, then each lane finds the word and
    // bit of its flag), so that a word with several flags costs no longer
    // chain than one with a single flag.  Every lane takes part in the
    // shuffles: no lane leaves before them.
    const int lane = (int)(threadIdx.x & 63);
    int64_t w;
    if (rec) {
        const int4 r = rec[b - list_blocks];
        const int64_t wa = (int64_t)r.x >> 5, wb = ((int64_t)r.y - 1) >> 5;
        w = (lane <= wb - wa) ? wa + lane : -1;
    } else {
        w = (b - list_blocks) * blockDim.x + threadIdx.x;
        if (32 * w >= v.N * v.T) w = -1;
    }
    const uint32_t f = (w >= 0) ? v.bmask[w] : 0u;
    if (f != 0u) v.bmask[w] = 0u;
    const int cnt = __popc(f);
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    const int wlo = (int)(w & 0xffffffff), whi = (int)(w >> 32);
    for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        // the last lane s whose flags start at or before flag k
        int s = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
            if (__shfl(excl, s + step, 64) <= k) s += step;
        uint32_t fw = __shfl(f, s, 64);
        const int r = k - __shfl(excl, s, 64);
        const int64_t ws = (int64_t)(uint32_t)__shfl(wlo, s, 64) | ((int64_t)__shfl(whi, s, 64) << 32);
        if (k >= total) continue;
        for (int i = 0; i < r; ++i) fw &= fw - 1u;
        int64_t t, x;
        v.elem_split(32 * ws + __builtin_ctz(fw), t, x);
        const bool pend = list_n > 0 && v.ev1[x].x >= 0;
        const REAL l = limiter_exact<REAL, STO, CONTIG>(v, t, x, v.up_range(x), 0, 0);
        if (!pend) store_limiter<REAL, STO>(v, t, x, l, substep, store_last);
    }
}

// Where the stage pass reads the limiters and the unlimited flows: the
// network scratch (StageGlobal), or the LDS rows of the block-local schedule
// (StageLds: entries of the block's catchments [base, base + blockDim.x)).
struct StageGlobal {
    template <class REAL, class STO>
    __device__ __forceinline__ REAL lim(const View<REAL, STO>& v, int64_t e, int64_t) const { return v.lim[e]; }
    template <class REAL, class STO>
    __device__ __forceinline__ REAL qr(const View<REAL, STO>& v, int64_t e, int64_t) const { return v.qraw_r[e]; }
    template <class REAL, class STO>
    __device__ __forceinline__ REAL qf(const View<REAL, STO>& v, int64_t e, int64_t) const { return v.qraw_f[e]; }
};
template <class REAL>
struct StageLds {
    const REAL *l, *r, *f;
    int64_t base;
    template <class STO>
    __device__ __forceinline__ REAL lim(const View<REAL, STO>&, int64_t, int64_t i) const { return l[i - base]; }
    template <class STO>
    __device__ __forceinline__ REAL qr(const View<REAL, STO>&, int64_t, int64_t i) const { return r[i - base]; }
    template <class STO>
    __device__ __forceinline__ REAL qf(const View<REAL, STO>&, int64_t, int64_t i) const { return f[i - base]; }
};

// Stage of one catchment in the fused schedule: limited outflow, inflow sums,
// bifurcation totals, storage update and flood stage, and with LEV the levee
// stage of levee catchments.
//
// Bound by memory latency: every field of the catchment that does not depend
// on another load is read up front (one round trip), then the downstream
// limiter, the upstream ranges, the bifurcation events and, for catchments
// near bankfull or with a levee, the floodplain row and levee parameters (a
// second round trip); the arithmetic follows.
template <class REAL, class STO, bool CONTIG, bool LOG, int LFIX, bool LEV, class Src, class VT>
__device__ __forceinline__ void stage_fused_one(const VT& v, int64_t t, int64_t j, int substep,
                                                int store_inflow, LogPart& lp, const Src& src) {
    const int64_t e = v.el(t, j);
    const REAL dt = v.dt(t, j);
    // round trip 1
    const int64_t d = v.dn[j];
    const REAL l_own = src.lim(v, e, j);
    const REAL qr = src.qr(v, e, j), qf = src.qf(v, e, j);
    const int2 rg = v.up_range(j);
    const int32_t b = rg.x, en = rg.y;
    const STO rs = v.river_storage[e], fs = v.flood_storage[e];
    const REAL ro = v.prm(v.runoff, CMF_BATCHED_RUNOFF, e, j);
    const REAL rh = v.prm(v.river_height, CMF_BATCHED_RIVER_HEIGHT, e, j);
    const REAL rw = v.prm(v.river_width, CMF_BATCHED_RIVER_WIDTH, e, j);
    const REAL rl = v.prm(v.river_length, CMF_BATCHED_RIVER_LENGTH, e, j);
    // with byte flags (VT::kByteFlags) a byte each says whether the catchment
    // has a levee and bifurcation events, and the levee row index is read
    // only where there is a levee; otherwise the index and the event range
    // are read up front
    constexpr bool BYTES = VT::kByteFlags;
    const int32_t lv0 = (LEV && !BYTES) ? v.levee_index[j] : -1;
    const bool lev_flag = v.has(CMF_HAS_LEVEE) &&
                          ((LEV && !BYTES) ? lv0 >= 0
                                           : (v.is_levee ? v.is_levee[j] != 0 : (LEV && v.levee_index[j] >= 0)));
    const bool res = v.has(CMF_HAS_RESERVOIR) && v.is_reservoir[j];
    const bool events = v.has(CMF_HAS_BIFURCATION) &&
                        ((BYTES && v.ev_flag) ? v.ev_flag[j] != 0 : v.ev_ptr[j] != v.ev_ptr[j + 1]);
    // round trip 2
    const int32_t lv = !LEV ? -1 : (BYTES ? (lev_flag ? v.levee_index[j] : -1) : lv0);
    const bool lev_cell = LEV ? (lv >= 0) : lev_flag;
    // after the first sub-step's stage pass every catchment without a levee
    // holds +0.0 protected storage
    const STO ps = (substep > 0 && !lev_flag) ? (STO)0.0 : v.protected_storage[e];
    // the downstream limiter scales only the catchment's own negative flows
    const int64_t ed = v.el(t, d);
    const bool own_back = !(qr >= (REAL)0.0) || !(qf >= (REAL)0.0);
    const REAL l_dn = own_back ? src.lim(v, ed, d) : (REAL)1.0;
    TableRow<REAL, LFIX> row;
    bool row_loaded = false;
    REAL ca = (REAL)-1;
    const bool near = lev_flag || (REAL)(rs + fs + ps) > (REAL)0.9 * (rl * rw * rh);
    if (LFIX > 0 && near) {
        row.load(v.bat(CMF_BATCHED_FLOOD_DEPTH_TABLE) ? v.table + e * v.L : v.table + j * v.L);
        row_loaded = true;
        ca = v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, j);
    }
    REAL crown = (REAL)0, lfrac = (REAL)0, lbase = (REAL)0;
    if (LEV && lev_cell) {
        const int64_t le = t * v.V + lv;
        crown = v.prm(v.levee_crown, CMF_BATCHED_LEVEE_CROWN_HEIGHT, le, lv);
        lfrac = v.prm(v.levee_frac, CMF_BATCHED_LEVEE_FRACTION, le, lv);
        lbase = v.prm(v.levee_base, CMF_BATCHED_LEVEE_BASE_HEIGHT, le, lv);
    }
    const REAL rout = (qr >= (REAL)0.0) ? qr * l_own : qr * l_dn;
    const REAL fout = (qf >= (REAL)0.0) ? qf * l_own : qf * l_dn;
    v.river_outflow[e] = rout;
    v.flood_outflow[e] = fout;
    if (v.acc_outflow) v.acc_outflow[e] += (double)(rout + fout) * (double)dt;
    // upstream u has downstream j: its negative flow is limited by j
    STO ri = (STO)0, fi = (STO)0;
    STO rt = res ? v.res_inflow[e] : (STO)0;
    for (int32_t k = b; k < en; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == j) continue;
        const int64_t eu = v.el(t, u);
        const REAL lu = src.lim(v, eu, u);
        const REAL ur0 = src.qr(v, eu, u), uf0 = src.qf(v, eu, u);
        const REAL ur1 = (ur0 >= (REAL)0.0) ? ur0 * lu : ur0 * l_own;
        const REAL uf1 = (uf0 >= (REAL)0.0) ? uf0 * lu : uf0 * l_own;
        ri += (STO)ur1;
        fi += (STO)uf1;
        if (res) rt += (STO)(ur1 + uf1);
    }
    if (res) v.res_inflow[e] = rt;
    STO gsum = (STO)0;
    REAL gb = (REAL)0.0;
    if (v.has(CMF_HAS_BIFURCATION)) {
        if (events) gsum = bif_nets<REAL, STO>(v, t, j, l_own, src);
        gb = (REAL)gsum;
    }
    const bool last = v.last_substep(t, j, substep, store_inflow);
    if (last) {
        v.river_inflow[e] = ri;
        v.flood_inflow[e] = fi;
        if (v.has(CMF_HAS_BIFURCATION)) v.gbif[e] = gsum;
    }
    const StagePre<REAL, STO> pre = stage_update_p<REAL, STO, LOG>(rs, fs, ps, ro, (REAL)ri, (REAL)fi, rout, fout,
                                                                   gb, dt, lp);
    const StageOut<REAL, STO> so = stage_diag_p<REAL, STO, LFIX, LOG>(v, e, j, pre.total_hp, rl, rw, rh, ca,
                                                                     !lev_cell, row, row_loaded, lp);
    if (LEV && lev_cell) {
        if (!row_loaded) row.load(v.bat(CMF_BATCHED_FLOOD_DEPTH_TABLE) ? v.table + e * v.L : v.table + j * v.L);
        const REAL cav = ca >= (REAL)0 ? ca : v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, j);
        const LeveeOut<REAL, STO> lo = levee_one_p<REAL, STO, LFIX>(
            v, rl, rw, rh, cav, crown, lfrac, lbase, so.rsf_hp, so.fsf_hp, so.fdep, so.rdep, so.ffr, row);
        if (LOG) levee_log(lo, so.rsf_hp + so.fsf_hp, cav, lp);
        v.river_storage[e] = lo.r_hp;
        v.flood_storage[e] = lo.f_hp;
        v.protected_storage[e] = lo.p_hp;
        v.river_depth[e] = lo.r_dph;
        v.flood_depth[e] = lo.f_dph;
        v.protected_depth[e] = lo.p_dph;
        v.flood_fraction[e] = lo.f_frc;
    } else {
        // a separate levee pass (k_levee) reads the flood fraction
        store_state(v, e, so, pre.fs_old, pre.ps_old, last || (v.has(CMF_HAS_LEVEE) && !LEV));
    }
}

// fused, after k_limiter: the stage of every catchment from the scratch.
template <class REAL, class STO, bool CONTIG, bool LOG, int LFIX, bool LEV, class VT = View<REAL, STO>>
__global__ __launch_bounds__(kMaxBlock) void k_stage_fused(VT v, int substep, int store_inflow) {
    const int64_t g = xcd_block() * blockDim.x + threadIdx.x;
    LogPart lp;
    if (LOG) {
#pragma unroll
        for (int k = 0; k < CMF_LOG_COUNT; ++k) lp.q[k] = 0.0;
    }
    int64_t t, j;
    if (v.catchment_of(g, t, j) && v.active_in_block(t, j, substep))
        stage_fused_one<REAL, STO, CONTIG, LOG, LFIX, LEV>(v, t, j, substep, store_inflow, lp, StageGlobal{});
    if (LOG) block_log_partials(lp, v.partials + v.partial_slot() * CMF_LOG_COUNT);
}

// ------------------------------------------------ block-local schedule ---
// One trial, ungrouped, contiguous upstream ranges, no reservoirs.  A sub-step
// is k_fused_local over the whole network, then k_limiter_list over the
// catchments whose limiter it left (list `lim_list`), then k_stage_list over
// those whose stage it left (`stage_list`).  Block k of the catchment grid
// holds catchments [k B, (k + 1) B) (B = blockDim.x):
//   1. the outflow of every catchment (and, in leading blocks of their own,
//      the bifurcation paths, as k_outflow_paths); flows and negative
//      outgoing terms into LDS;
//   2. after a barrier, the storage limiter of each catchment whose upstream
//      range lies in the block and that is no path end (kLocalLim): the same
//      ordered sum as k_limiter, from LDS;
//   3. after a barrier, the stage of each catchment whose limiter, upstream
//      catchments' limiters and downstream limiter all came out of step 2
//      (kLocalStage).  No other block reads such a catchment's state in this
//      launch (its upstream catchments are in the block, it is no path end),
//      so its new state can be written here.
// Every value is the one the three-pass schedule computes, by the same
// operations in the same order.
template <class REAL, class STO, bool CONTIG, bool LOG, int LFIX, bool LEV, bool PATHS, bool PLEVEE, int LBFIX>
__global__ __launch_bounds__(kMaxBlock) void k_fused_local(View<REAL, STO> v, int substep, int64_t path_blocks,
                                                         int store_last, const uint8_t* __restrict__ lflags) {
    const int64_t b = blockIdx.x;
    if (PATHS && b < path_blocks) {
        bif_path<REAL, STO, PLEVEE, LBFIX, true>(v, substep, b * blockDim.x + threadIdx.x);
        return;
    }
    __shared__ REAL sh_qr[kMaxBlock], sh_qf[kMaxBlock], sh_neg[kMaxBlock], sh_lim[kMaxBlock];
    const int64_t blk = PATHS ? b - path_blocks : b;
    const int64_t base = blk * (int64_t)blockDim.x;
    const int64_t j = base + threadIdx.x;
    const bool valid = j < v.N;
    const uint8_t fl = valid ? lflags[j] : (uint8_t)0;
    OutflowLocal<REAL, STO> o;
    o.q_r = o.q_f = o.pos = o.neg = (REAL)0;
    o.rs = o.fs = (STO)0;
    if (valid) outflow_one<REAL, STO, true>(v, substep, blk, store_last, &o, fl);
    sh_qr[threadIdx.x] = o.q_r;
    sh_qf[threadIdx.x] = o.q_f;
    sh_neg[threadIdx.x] = o.neg;
    __syncthreads();
    if (fl & kLocalLim) {
        // outgoing_terms (upstream negative terms in ascending index, then
        // the own positive term), then limit_of
        STO acc = (STO)0;
        const int2 rg = v.up_range(j);
        const int32_t b0 = rg.x, b1 = rg.y;
        for (int32_t k = b0; k < b1; ++k) {
            if (k == j) continue;   // a mouth's range ends with the mouth
            acc += (STO)sh_neg[k - base];
        }
        acc += (STO)o.pos;
        const REAL og = (REAL)acc;
        const REAL rate = (REAL)(o.rs + o.fs);
        const REAL l = (og > (REAL)1e-8) ? vmin(rate / og, (REAL)1.0) : (REAL)1.0;
        sh_lim[threadIdx.x] = l;
        if (fl & kLimG) v.lim[j] = l;
        if (v.has(CMF_HAS_BIFURCATION) && store_last) v.limit_rate[j] = l;
    }
    __syncthreads();
    LogPart lp;
    if (LOG) {
#pragma unroll
        for (int k = 0; k < CMF_LOG_COUNT; ++k) lp.q[k] = 0.0;
    }
    if (fl & kLocalStage)
        stage_fused_one<REAL, STO, CONTIG, LOG, LFIX, LEV>(v, 0, j, substep, store_last, lp,
                                                           StageLds<REAL>{sh_lim, sh_qr, sh_qf, base});
    if (LOG) block_log_partials(lp, v.partials + blk * CMF_LOG_COUNT);
}

// The storage limiter of the listed catchments (one trial, ungrouped), from
// the terms in the network scratch: k_limiter's arithmetic.
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kMaxBlock) void k_limiter_list(View<REAL, STO> v, int store_last,
                                                          const int32_t* __restrict__ list, int64_t count) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const int64_t j = list[g];
    const int2 ev = v.has(CMF_HAS_BIFURCATION) ? v.ev1[j] : make_int2(-1, 0);
    STO og = outgoing_terms<REAL, STO, CONTIG>(v, 0, j);
    if (ev.x >= 0) og = bif_terms_from<REAL, STO>(v, 0, j, ev, og);
    const REAL l = limit_of(v, og, j);
    v.lim[j] = l;
    if (v.has(CMF_HAS_BIFURCATION) && store_last) v.limit_rate[j] = l;
}

// The stage of the listed catchments from the scratch; log partials in slots
// [slot0, slot0 + gridDim.x).
template <class REAL, class STO, bool CONTIG, bool LOG, int LFIX, bool LEV>
__global__ __launch_bounds__(kMaxBlock) void k_stage_list(View<REAL, STO> v, int substep, int store_inflow,
                                                        const int32_t* __restrict__ list, int64_t count,
                                                        int64_t slot0) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LogPart lp;
    if (LOG) {
#pragma unroll
        for (int k = 0; k < CMF_LOG_COUNT; ++k) lp.q[k] = 0.0;
    }
    if (g < count)
        stage_fused_one<REAL, STO, CONTIG, LOG, LFIX, LEV>(v, 0, (int64_t)list[g], substep, store_inflow, lp,
                                                           StageGlobal{});
    if (LOG) block_log_partials(lp, v.partials + (slot0 + blockIdx.x) * CMF_LOG_COUNT);
}

// Fold the per-block partials into log slot current_step: one block, each
// sum reduced over the blocks in a fixed order.  `mask` selects the sums the
// launching kernel produced.
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_log_final(View<REAL, STO> v, int64_t nblocks, uint32_t mask) {
    __shared__ double sh[kBlock];
    const int step = v.current_step[0];
    for (int k = 0; k < CMF_LOG_COUNT; ++k) {
        if (!(mask & (1u << k))) continue;
        double acc = 0.0;
        for (int64_t b = threadIdx.x; b < nblocks; b += kBlock) acc += v.partials[b * CMF_LOG_COUNT + k];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kBlock / 2; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) v.log_sums[k][step] += sh[0] * 1e-9;
        __syncthreads();
    }
}

// grouped networks: one block per group g sums the partials of the group's
// blocks [gblk[g], gblk[g+1]) in a fixed order into log_sums[k][g *
// log_buffer_size + step]; groups past their sub-step count are skipped.
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_log_final_grouped(View<REAL, STO> v, const int32_t* gblk,
                                                              uint32_t mask, int substep) {
    __shared__ double sh[kBlock];
    const int64_t g = blockIdx.x;
    if (substep >= 0 && substep >= v.max_sub_steps[g]) return;
    const int step = v.current_step[0];
    const int64_t b0 = gblk[g], b1 = gblk[g + 1];
    for (int k = 0; k < CMF_LOG_COUNT; ++k) {
        if (!(mask & (1u << k))) continue;
        double acc = 0.0;
        for (int64_t b = b0 + threadIdx.x; b < b1; b += kBlock) acc += v.partials[b * CMF_LOG_COUNT + k];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kBlock / 2; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) v.log_sums[k][g * v.logsize + step] += sh[0] * 1e-9;
        __syncthreads();
    }
}

// ============================================================= levee ====
// per-op compute_levee_stage(_log): one thread per levee, on the state the
// flood stage left.
template <class REAL, class STO, bool LOG, int LFIX>
__global__ __launch_bounds__(kBlock) void k_levee(View<REAL, STO> v, int substep) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LogPart lp;
    if (LOG) {
#pragma unroll
        for (int k = 0; k < CMF_LOG_COUNT; ++k) lp.q[k] = 0.0;
    }
    if (g < v.V * v.T) {
        const int64_t t = g / v.V, lv = g - t * v.V;
        const int64_t c = v.levee_idx[lv];
        if (v.active(t, c, substep)) {
            const int64_t e = v.el(t, c);
            TableRow<REAL, LFIX> row;
            row.load(v.bat(CMF_BATCHED_FLOOD_DEPTH_TABLE) ? v.table + e * v.L : v.table + c * v.L);
            const STO rs_hp = v.river_storage[e], fs_hp = v.flood_storage[e];
            const LeveeOut<REAL, STO> lo = levee_one<REAL, STO, LFIX>(
                v, e, c, lv, t * v.V + lv, rs_hp, fs_hp, v.flood_depth[e], v.river_depth[e],
                v.flood_fraction[e], row);
            if (LOG) levee_log(lo, rs_hp + fs_hp, v.prm(v.catchment_area, CMF_BATCHED_CATCHMENT_AREA, e, c), lp);
            v.river_storage[e] = lo.r_hp;
            v.flood_storage[e] = lo.f_hp;
            v.protected_storage[e] = lo.p_hp;
            v.river_depth[e] = lo.r_dph;
            v.flood_depth[e] = lo.f_dph;
            v.protected_depth[e] = lo.p_dph;
            v.flood_fraction[e] = lo.f_frc;
        }
    }
    if (LOG) block_log_partials(lp, v.partials + (int64_t)blockIdx.x * CMF_LOG_COUNT);
}

// =============================================================== CFL ====
// Sub-step count of one catchment (adaptive_time.py:30-48): dam-related
// catchments count 1.
// CFL sub-step count of a catchment from its river depth and downstream distance
template <class REAL, class STO>
__device__ __forceinline__ int cfl_count(const View<REAL, STO>& v, REAL river_depth, REAL dist, REAL outer) {
    const REAL depth = vmax(river_depth, (REAL)0.01);
    const REAL dt = v.adaptive_factor * dist / dsqrt(v.gravity * depth);
    const REAL dtc = vmin(dt, outer);
    const REAL nf = floor(outer / dtc - (REAL)0.01) + (REAL)1.0;
    return (int)nf;
}
template <class REAL, class STO>
__device__ __forceinline__ int cfl_steps(const View<REAL, STO>& v, int64_t t, int64_t i, REAL outer) {
    if (v.has(CMF_HAS_RESERVOIR) && v.is_dam_related[i]) return 1;
    const int64_t e = v.el(t, i);
    return cfl_count(v, v.river_depth[e], v.prm(v.dist, CMF_BATCHED_DOWNSTREAM_DISTANCE, e, i), outer);
}

__device__ __forceinline__ int block_max_int(int best) {
    __shared__ int sh[kMaxBlock / 64];
    best = wave_max_int(best);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
    __syncthreads();
    int m = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = max(m, sh[w]);
    return m;
}

// grid (chunks, trials-or-1): block (bx, t) reduces catchments of trial t,
// four independent grid strides per iteration (their loads in flight
// together); one atomic per block, so few blocks.  Without per-trial dt every
// trial folds into slot 0.
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_cfl(View<REAL, STO> v) {
    const int64_t t = blockIdx.y;
    const REAL outer = v.outer_time_step[0];
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.N; i += 4 * stride) {
        const int c0 = cfl_steps(v, t, i, outer);
        const int c1 = (i + stride < v.N) ? cfl_steps(v, t, i + stride, outer) : 0;
        const int c2 = (i + 2 * stride < v.N) ? cfl_steps(v, t, i + 2 * stride, outer) : 0;
        const int c3 = (i + 3 * stride < v.N) ? cfl_steps(v, t, i + 3 * stride, outer) : 0;
        best = max(best, max(max(c0, c1), max(c2, c3)));
    }
    const int m = block_max_int(best);
    if (threadIdx.x == 0) atomicMax(&v.max_sub_steps[v.has(CMF_PER_TRIAL_DT) ? t : 0], m);
}

// One trial, or the trial-inner layout with T dividing 64 (element e =
// catchment * T + trial):
// threads walk the elements contiguously with a stride that is a multiple of
// 64, so every element a thread meets belongs to the trial lane % T; eight
// elements' loads are issued before any is used.  Lanes of one trial are
// folded with shuffles over lane distances that are multiples of T, each
// block leaves one maximum per trial in `part` [gridDim.x, T], and
// k_cfl_fold takes the maximum over the blocks (no atomics on a few hot
// addresses from every block).
//
// OUTER_ARG (the adaptive advance): the outer time step comes as an argument
// instead of from outer_time_step, which k_cfl_fold_tail sets.
constexpr int kCflInnerBlocks = 1024;
template <class REAL, class STO, bool OUTER_ARG = false>
__global__ __launch_bounds__(kBlock) void k_cfl_inner(View<REAL, STO> v, int log2T, int32_t* part,
                                                      double outer_arg) {
    __shared__ int sh[kBlock / 64][64];
    const int64_t NT = v.N * v.T;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const REAL outer = OUTER_ARG ? (REAL)outer_arg : v.outer_time_step[0];
    int best = 0;
    for (int64_t e0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; e0 < NT; e0 += 8 * stride) {
        REAL depth[8], dist[8];
        bool dam[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t e = e0 + k * stride < NT ? e0 + k * stride : e0;
            const int64_t i = e >> log2T;
            depth[k] = v.river_depth[e];
            dist[k] = v.prm(v.dist, CMF_BATCHED_DOWNSTREAM_DISTANCE, e, i);
            dam[k] = v.has(CMF_HAS_RESERVOIR) && v.is_dam_related[i];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (e0 + k * stride < NT) best = max(best, dam[k] ? 1 : cfl_count(v, depth[k], dist[k], outer));
    }
    const int T = 1 << log2T;
    for (int off = 32; off >= T; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < T) sh[wave][lane] = best;
    __syncthreads();
    if ((int)threadIdx.x < T) {
        int m = sh[0][threadIdx.x];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = max(m, sh[w][threadIdx.x]);
        part[(int64_t)blockIdx.x * T + threadIdx.x] = m;
    }
}
// one block per trial: the maximum of its blocks' entries into its count
// (slot 0 without per-trial time steps)
__global__ __launch_bounds__(kBlock) void k_cfl_fold(const int32_t* part, int64_t nblocks, int64_t T,
                                                     int32_t* counts, int per_trial) {
    const int64_t t = blockIdx.x;
    int m = 0;
    for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) m = max(m, part[b * T + t]);
    m = block_max_int(m);
    if (threadIdx.x == 0) atomicMax(&counts[per_trial ? t : 0], m);
}

// What the adaptive advance needs after k_cfl_inner<..., OUTER_ARG = true>,
// in one block: the counts folded from the blocks' maxima (written whole: no
// clearing before), the time steps outer / max(count, 1) in double, rounded
// once (cama_flood_model.py:127,132), the outer time step, current_step =
// count - 1 (one slot, when given) and the counts again at `host` (the device
// address of pinned host memory, which the host reads without a copy).
template <class REAL>
struct CflTail {
    int32_t* counts;        // max_sub_steps: nslots entries
    REAL* time_step;        // nslots entries
    REAL* outer_time_step;  // one entry
    int32_t* current_step;  // or null
    int32_t* host;          // nslots entries
    double outer;
    int per_trial;          // nslots = T, else 1 (every trial folds into slot 0)
};
template <class REAL>
__global__ __launch_bounds__(kBlock) void k_cfl_fold_tail(const int32_t* part, int64_t nblocks, int log2T,
                                                          CflTail<REAL> tail) {
    __shared__ int sh[kBlock / 64][64];
    const int T = 1 << log2T;
    // thread k folds trial k % T over the blocks k / T, k / T + kBlock / T, ...
    const int tt = (int)threadIdx.x & (T - 1);
    int m = 0;
    for (int64_t b = threadIdx.x >> log2T; b < nblocks; b += kBlock >> log2T) m = max(m, part[b * T + tt]);
    // then over the lanes of one trial (distances that are multiples of T)
    // and over the waves
    for (int off = 32; off >= T; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) < (unsigned)T) sh[threadIdx.x >> 6][threadIdx.x & 63] = m;
    __syncthreads();
    if ((int)threadIdx.x < T) {
        int c = sh[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w) c = max(c, sh[w][threadIdx.x]);
        sh[0][threadIdx.x] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) tail.outer_time_step[0] = (REAL)tail.outer;
    const int nslots = tail.per_trial ? T : 1;
    if ((int)threadIdx.x < nslots) {
        int c = sh[0][threadIdx.x];
        if (!tail.per_trial)
            for (int k = 1; k < T; ++k) c = max(c, sh[0][k]);
        const int cc = c < 1 ? 1 : c;
        tail.counts[threadIdx.x] = c;
        tail.time_step[threadIdx.x] = (REAL)(tail.outer / (double)cc);
        if (tail.current_step && nslots == 1) tail.current_step[0] = cc - 1;
        __hip_atomic_store(tail.host + threadIdx.x, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __threadfence_system();
}

// grouped networks: kCflRecords consecutive table entries (each a single
// group's catchments) per block of kGroupRecord threads, one catchment of
// each entry per thread; the records are read first and every catchment load
// is issued before any is used.  The entries' maxima are merged over runs of
// one group, then one integer atomic max per run into the group's count (a
// large group's entries otherwise queue on one address).
constexpr int kCflRecords = 16;
template <class REAL, class STO>
__global__ __launch_bounds__(kGroupRecord) void k_cfl_grouped(View<REAL, STO> v, int64_t nrec) {
    __shared__ int sh[kCflRecords][kGroupRecord / 64];
    __shared__ int rec_max[kCflRecords], rec_group[kCflRecords];
    const REAL outer = v.outer_time_step[0];
    int4 r[kCflRecords];
#pragma unroll
    for (int k = 0; k < kCflRecords; ++k) {
        const int64_t q = (int64_t)blockIdx.x * kCflRecords + k;
        r[k] = q < nrec ? v.blk_rec[q] : make_int4(0, 0, 0, -1);
    }
    REAL depth[kCflRecords], dist[kCflRecords];
    bool dam[kCflRecords];
#pragma unroll
    for (int k = 0; k < kCflRecords; ++k) {
        // an index inside the record, or catchment 0 for lanes past its end
        const int64_t i = (int64_t)r[k].x + threadIdx.x < r[k].y ? (int64_t)r[k].x + threadIdx.x : 0;
        depth[k] = v.river_depth[v.el(0, i)];
        dist[k] = v.prm(v.dist, CMF_BATCHED_DOWNSTREAM_DISTANCE, v.el(0, i), i);
        dam[k] = v.has(CMF_HAS_RESERVOIR) && v.is_dam_related[i];
    }
#pragma unroll
    for (int k = 0; k < kCflRecords; ++k) {
        const bool in = (int64_t)r[k].x + threadIdx.x < r[k].y;
        const int c = in ? (dam[k] ? 1 : cfl_count(v, depth[k], dist[k], outer)) : 0;
        const int m = wave_max_int(c);
        if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x < kCflRecords) {
        int m = sh[threadIdx.x][0];
        for (int w = 1; w < kGroupRecord / 64; ++w) m = max(m, sh[threadIdx.x][w]);
        int g = -1;
#pragma unroll
        for (int k = 0; k < kCflRecords; ++k)
            if ((int)threadIdx.x == k) g = r[k].w;   // static indices: r stays in registers
        rec_max[threadIdx.x] = m;
        rec_group[threadIdx.x] = g;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int g = rec_group[0], m = rec_max[0];
        for (int k = 1; k <= kCflRecords; ++k) {
            if (k == kCflRecords || rec_group[k] != g) {
                if (g >= 0) atomicMax(&v.max_sub_steps[g], m);
                if (k == kCflRecords) break;
                g = rec_group[k];
                m = rec_max[k];
            } else {
                m = max(m, rec_max[k]);
            }
        }
    }
}

// The per-sub-step group block tables (level k: the records of the groups
// with count > k, group g's from level_pos[k * G + g] on): record r of group
// g is copied into the first min(count, counts[g]) levels.
__global__ void k_level_tables(const int4* rec, int64_t nrec, const int32_t* gblk, const int32_t* counts,
                               const int32_t* level_pos, int64_t G, int count, int4* out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int4 x = rec[r];
    const int64_t g = x.w;
    const int64_t within = r - gblk[g];
    const int c = min(count, counts[g]);
    for (int k = 0; k < c; ++k) out[(int64_t)level_pos[(int64_t)k * G + g] + within] = x;
}

// time_step[t] = outer / count[t], in double then rounded (cama_flood_model.py:127,132);
// with one slot and `current_step` given, also current_step = count - 1
template <class REAL>
__global__ void k_set_dt(REAL* time_step, const int32_t* counts, int64_t n, double outer, int32_t* current_step) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = counts[t] < 1 ? 1 : counts[t];
    time_step[t] = (REAL)(outer / (double)c);
    if (current_step && n == 1) current_step[0] = c - 1;
}

__global__ void k_zero_int(int32_t* p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// ==================================================== runoff mapping ===
// Catchment runoff from a gridded field through the catchment x grid-cell
// area-weight matrix (CSR, rows in the caller's catchment order):
// out[r] = scale * sum_k w[k] * grid[idx[k]], summed in double in index order
// (out[out_index[r]] when out_index is given: the rows in an order whose grid
// gathers stay close, the results scattered to the catchments' elements).
template <class REAL>
__global__ __launch_bounds__(kBlock) void k_map_runoff(int64_t rows, const int64_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ idx,
                                                       const float* __restrict__ w,
                                                       const float* __restrict__ grid, double scale,
                                                       const int32_t* __restrict__ out_index,
                                                       REAL* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double acc = 0.0;
    for (int64_t k = indptr[r], e = indptr[r + 1]; k < e; ++k) acc += (double)w[k] * (double)grid[idx[k]];
    out[out_index ? (int64_t)out_index[r] : r] = (REAL)(scale * acc);
}

// ============================================= explicit instantiation ===
// (the C ABI in cmf_capi.hip includes this file and instantiates what it launches)

}  // namespace cmf
