// Hand-written adjoint (reverse mode) of one fused sub-step of the base
// physics: local-inertial river and floodplain outflow
// (cmfgpu/phys/triton/outflow.py:15-315), the storage limiter and inflow sums
// (:318-383), the storage update and flood-stage diagnosis
// (cmfgpu/phys/triton/storage.py:16-193), of the bifurcation paths
// (cmfgpu/phys/triton/bifurcation.py:13-151: per-level local-inertial flow,
// the 5 % storage limiter of a path, the path terms of the outgoing totals and
// the path nets limited by the end catchments; with levees the variant of
// cmfgpu/phys/triton/levee.py:565-684) and of the levee stage
// (cmfgpu/phys/triton/levee.py:62-304 with the partition :18-59) and of the
// reservoirs (the kinematic flow upstream of a dam, outflow.py:251-272, and
// the release by storage regime, reservoir.py:17-184, from the inflow the
// previous sub-step gathered).
//
// The sub-step maps the state x = (river/flood outflow, river/flood/flood-area
// cross-section depth, river/flood storage, river/flood depth) and the
// parameters (river Manning n, river height H) to the next state.  Given the
// adjoint of the next state, the passes below give the adjoint of x and add
// the parameter gradients, with dt held at its forward value.  Branch decisions
// (max/min/clamp picks, wet or dry, floodplain level, limiter regimes) are
// those of the forward evaluation, recomputed bit for bit from the saved input
// state; derivatives are formed in double precision.
//
// Passes (each over catchments or paths, one grid-wide boundary between them):
//   B0  (paths) recompute the limited level flows, level sums and outgoing terms
//   A0  recompute the unlimited outflows and outgoing terms  } skipped when the
//   A1  recompute the outgoing totals and storage limiters   } forward taped them
//   A2  reverse of the storage update and stage diagnosis
//   A3  reverse of the limited flows, inflow sums and path nets, then of the
//       limiter
//   B2  (paths) reverse of the path nets, the path terms, the path limiter
//       and the per-level momentum update
//   A4  reverse of the outflow momentum update
//   A5  gather of the downstream water-surface adjoints (or deferred into the
//       A2 pass of the next, earlier sub-step, which loops over the same
//       upstream range anyway)
// A path pass writes what belongs to its end catchments into per-path slots;
// the next catchment pass gathers them over the catchment's events in path
// order, so no pass adds into another element's memory.

namespace cmf {

enum {
    ADJ_QR = 0, ADJ_QF, ADJ_RCSD, ADJ_FCSD, ADJ_FCSA, ADJ_RS, ADJ_FS, ADJ_RD, ADJ_FD, ADJ_COUNT
};

// Adjoint arrays are held in REAL (the profile's flux precision); the
// arithmetic of every pass is double; the gradients accumulate in double.
template <class REAL>
struct AdjView {
    // saved input state of the sub-step
    const REAL *q_r, *q_f, *rcsd, *fcsd, *fcsa, *rd, *fd;
    const void *rs_raw, *fs_raw;
    const REAL* lout[ADJ_COUNT];     // adjoint of the sub-step outputs
    REAL* lin[ADJ_COUNT];            // adjoint of the sub-step inputs (written)
    double *g_n, *g_h;               // gradient accumulators
    // forward intermediates: unlimited flows and limiters (taped, or
    // recomputed into scratch by A0/A1), outgoing terms (A0 scratch)
    const REAL *qr, *qf, *lim;
    REAL *w_qr, *w_qf, *w_lim, *pos, *neg;
    REAL *a_rout, *a_fout, *a_ri, *a_fi, *a_qr, *a_qf, *a_og, *d_wd;
    REAL dt;
    // bifurcation paths (bif: the module is open): saved input path state
    // [P, LB] and its adjoints after (lout_b*) and before (lin_b*) the sub-step
    bool bif;
    const REAL *bq, *bcsd, *lout_bq, *lout_bcsd;
    REAL *lin_bq, *lin_bcsd;
    // B0: limited level flows [P, LB], their sum, the outgoing terms at the
    // upstream (pos) and downstream (neg) end
    REAL *p_qlim, *p_raw, *p_pos, *p_neg;
    // A2: adjoint of each catchment's bifurcation total; B2: adjoint of the
    // end water surfaces, total storages and (levee ends) protected depths
    REAL *a_gb, *p_w0, *p_w1, *p_t0, *p_t1, *p_d0, *p_d1;
    // levees (lev: the module is open): protected storage (STO) and
    // protected depth join the state
    bool lev;
    const void* ps_raw;
    const REAL *pd, *lout_ps, *lout_pd;
    REAL *lin_ps, *lin_pd;
    // reservoirs (res: the module is open): the inflow each reservoir
    // gathered in the previous sub-step (STO) joins the state
    bool res;
    const void* ri_raw;
    const REAL* lout_ri;
    REAL* lin_ri;
    int32_t* res_row;     // catchment -> reservoir row (reservoir catchments only)
};

constexpr int kAdjMaxLevels = 8;

template <class STO>
__device__ __forceinline__ STO adj_sto(const void* p, int64_t i) { return static_cast<const STO*>(p)[i]; }

// Protected storage of catchment i in the saved state: the taped value at a
// levee catchment, zero elsewhere (a stage pass leaves zero outside levee
// catchments, and the adjoint's driver refuses an initial state that holds
// any there).
template <class REAL, class STO>
__device__ __forceinline__ REAL adj_ps(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t i) {
    return (a.lev && v.is_levee[i]) ? (REAL)adj_sto<STO>(a.ps_raw, i) : (REAL)0.0;
}

// ---------------------------------------------------- outflow trace ----
// The arithmetic of k_outflow (FUSED, no reservoir override), same
// operations in the same order, keeping what the reverse needs.
template <class REAL>
struct OutTrace {
    bool mouth, river_cond, flood_cond, killed, negative, limited, mx_own, slope_in;
    REAL wse, wse_dn, slope, fslope, upd_rcsd, r_sifd_raw, r_sifd, fsd, upd_fcsd, f_sifd_raw, f_sifd;
    REAL upd_fcsa_raw, upd_fcsa, f_csa_m, f_imp_raw, f_imp, unit_r, den_r, qr0, den_f, qf0, qf1, tot, total_neg;
    REAL ratio, limit, qr, qf;
    // upstream of a dam: the kinematic flows (raw, and which clamps held)
    bool dam, kr_cap, kr_pos, kf_cap, kf_pos;
    REAL kr_raw, kf_raw, kf_area_raw, kf_coef;
    // a reservoir catchment: the release (regime 1-4, picks and clamps)
    bool reservoir, rel_flood, rel_oflood, rel_spill, rel_pos, rel_cap;
    int regime;
    REAL rel_total, rel_inflow, rel_frac, rel_raw;
};

// RES: the reservoir module is open (a compile-time switch, so that the
// kernels of networks without reservoirs carry none of its registers)
template <class REAL, class STO, bool RES>
__device__ __forceinline__ OutTrace<REAL> outflow_trace(const View<REAL, STO>& v, const AdjView<REAL>& a,
                                                        int64_t i, REAL time_step) {
    OutTrace<REAL> o;
    const int64_t d = v.dn[i];
    o.mouth = (d == i);
    const REAL gravity = v.gravity;
    const REAL r_out = a.q_r[i];
    const REAL r_man = v.river_manning[i];
    const REAL r_dep = a.rd[i];
    const REAL r_wid = v.river_width[i];
    const REAL r_len = v.river_length[i];
    const REAL r_hgt = v.river_height[i];
    // the floodplain flow, Manning's n and cross-section area are read where
    // the floodplain flows (flood_cond) or below a dam
    const REAL f_dep = a.fd[i];
    const REAL c_elv = v.elev[i];
    const REAL dn_dist = v.dist[i];
    const REAL rs = (REAL)adj_sto<STO>(a.rs_raw, i);
    const REAL fs = (REAL)adj_sto<STO>(a.fs_raw, i);
    const REAL ps = adj_ps(v, a, i);
    const REAL r_csd = a.rcsd[i], f_csd = a.fcsd[i];

    o.tot = rs + fs + ps;
    const REAL r_dep_dn = a.rd[d];
    const REAL r_hgt_dn = v.river_height[d];
    const REAL c_elv_dn = v.elev[d];
    // the surfaces, their drop and the depths above bed and bank as the
    // forward forms them: in FP32 as the reference does, or with precise
    // surfaces in double, each difference rounded once (the derivatives are
    // the same; only the values that pick the branches differ)
    REAL dwse;
    if (precise_surfaces(v)) {
        const double rel = (double)c_elv - (double)r_hgt;
        const double w = (double)r_dep + rel;
        const double w_dn = o.mouth ? (double)c_elv : (double)r_dep_dn + ((double)c_elv_dn - (double)r_hgt_dn);
        const double mw = vmax(w, w_dn);
        o.wse = (REAL)w;
        o.wse_dn = (REAL)w_dn;
        o.mx_own = (w > w_dn || w_dn != w_dn);
        dwse = (REAL)(w - w_dn);
        o.upd_rcsd = o.mouth ? r_dep : (REAL)(mw - rel);
        o.fsd = (REAL)((o.mouth ? w : mw) - (double)c_elv);
    } else {
        const REAL river_elevation = c_elv - r_hgt;
        o.wse = r_dep + river_elevation;
        const REAL river_elevation_dn = c_elv_dn - r_hgt_dn;
        o.wse_dn = r_dep_dn + river_elevation_dn;
        if (o.mouth) o.wse_dn = c_elv;
        o.mx_own = (o.wse > o.wse_dn || o.wse_dn != o.wse_dn);
        const REAL max_wse = vmax(o.wse, o.wse_dn);
        dwse = o.wse - o.wse_dn;
        o.upd_rcsd = o.mouth ? r_dep : (max_wse - river_elevation);
        o.fsd = (o.mouth ? o.wse : max_wse) - c_elv;
    }
    o.slope = dwse / dn_dist;
    o.fslope = vclamp(o.slope, (REAL)-0.005, (REAL)0.005);
    o.slope_in = (o.slope > (REAL)-0.005) && (o.slope < (REAL)0.005);
    o.r_sifd_raw = dsqrt(o.upd_rcsd * r_csd);
    o.r_sifd = vmax(o.r_sifd_raw, (REAL)1e-6);
    o.upd_fcsd = vmax(o.fsd, (REAL)0.0);
    o.f_sifd_raw = dsqrt(o.upd_fcsd * f_csd);
    o.f_sifd = vmax(o.f_sifd_raw, (REAL)1e-6);
    o.upd_fcsa_raw = fs / r_len - f_dep * r_wid;
    o.upd_fcsa = vmax(o.upd_fcsa_raw, (REAL)0.0);

    const REAL r_cs_area = o.upd_rcsd * r_wid;
    o.river_cond = (o.r_sifd > (REAL)1e-5) && (r_cs_area > (REAL)1e-5);
    REAL q_r = (REAL)0.0;
    o.unit_r = r_out / r_wid;
    o.den_r = (REAL)1;
    if (o.river_cond) {
        const REAL num_r = r_wid * (o.unit_r + gravity * time_step * o.r_sifd * o.slope);
        o.den_r = (REAL)1.0 + gravity * time_step * (r_man * r_man) * dabs(o.unit_r)
                * ((REAL)1.0 / (o.r_sifd * o.r_sifd * dcbrt(o.r_sifd)));
        q_r = num_r / o.den_r;
    }
    o.qr0 = q_r;
    o.flood_cond = (o.f_sifd > (REAL)1e-5) && (o.upd_fcsa > (REAL)1e-5);
    REAL q_f = (REAL)0.0;
    o.f_csa_m = (REAL)1e-6;
    o.f_imp_raw = (REAL)0;
    o.f_imp = (REAL)1;
    o.den_f = (REAL)1;
    if (o.flood_cond) {
        const REAL f_out = a.q_f[i];
        const REAL f_man = v.flood_manning[i];
        o.f_csa_m = vmax(a.fcsa[i], (REAL)1e-6);
        o.f_imp_raw = dsqrt(o.upd_fcsa * o.f_csa_m);
        o.f_imp = vmax(o.f_imp_raw, (REAL)1e-6);
        const REAL num_f = f_out + gravity * time_step * o.f_imp * o.fslope;
        o.den_f = (REAL)1.0 + gravity * time_step * (f_man * f_man) * dabs(f_out)
                * ((REAL)1.0 / (o.f_sifd * dcbrt(o.f_sifd))) / o.f_imp;
        q_f = num_f / o.den_f;
    }
    o.qf0 = q_f;
    o.killed = (q_r * q_f) < (REAL)0.0;
    if (o.killed) q_f = (REAL)0.0;
    o.qf1 = q_f;
    o.negative = (q_r < (REAL)0.0) && !o.mouth;
    o.limited = false;
    o.limit = (REAL)1;
    o.total_neg = (REAL)1;
    o.ratio = (REAL)1;
    if (o.negative) {
        o.total_neg = (-q_r - q_f) * time_step;
        o.ratio = (REAL)0.05 * o.tot / o.total_neg;
        o.limited = (o.ratio < (REAL)1.0);
        o.limit = vmin(o.ratio, (REAL)1.0);
        q_r = q_r * o.limit;
        q_f = q_f * o.limit;
    }
    o.dam = false;
    if (RES && v.is_dam_up[i]) {
        // kinematic flow upstream of a dam (outflow.py:251-272)
        o.dam = true;
        REAL bed_slope = (c_elv - c_elv_dn) / dn_dist;
        bed_slope = vmax(bed_slope, v.min_kin);
        const REAL kin_riv_vel = ((REAL)1.0 / r_man) * dsqrt(bed_slope) * dcbrt(r_dep * r_dep);
        o.kr_raw = r_wid * r_dep * kin_riv_vel;
        o.kr_cap = !(o.kr_raw < rs / time_step);
        REAL kin_riv = vmin(o.kr_raw, rs / time_step);
        o.kr_pos = kin_riv > (REAL)0.0;
        kin_riv = vmax(kin_riv, (REAL)0.0);
        const REAL bed_slope_f = vmin(bed_slope, (REAL)0.005);
        o.kf_coef = ((REAL)1.0 / v.flood_manning[i]) * dsqrt(bed_slope_f);
        const REAL kin_fld_vel = o.kf_coef * dcbrt(f_dep * f_dep);
        o.kf_area_raw = fs / r_len - f_dep * r_wid;
        const REAL kin_fld_area = vmax(o.kf_area_raw, (REAL)0.0);
        o.kf_raw = kin_fld_area * kin_fld_vel;
        o.kf_cap = !(o.kf_raw < fs / time_step);
        REAL kin_fld = vmin(o.kf_raw, fs / time_step);
        o.kf_pos = kin_fld > (REAL)0.0;
        kin_fld = vmax(kin_fld, (REAL)0.0);
        q_r = kin_riv;
        q_f = kin_fld;
    }
    o.reservoir = RES && v.is_reservoir[i];
    if (o.reservoir) {
        // release by storage regime (k_reservoir, reservoir.py:114-163)
        const int64_t r = a.res_row[i];
        const REAL total = rs + fs;
        const REAL inflow = (REAL)adj_sto<STO>(a.ri_raw, i) + v.runoff[i];
        const REAL cv = v.res_cons[r], ev = v.res_emer[r], av = v.res_adj[r];
        const REAL qn = v.res_qn[r], qa = v.res_qa[r], qf = v.res_qf[r];
        o.rel_total = total;
        o.rel_inflow = inflow;
        REAL out = (REAL)0;
        o.regime = 0;
        o.rel_flood = o.rel_oflood = o.rel_spill = false;
        o.rel_frac = (REAL)0;
        if (total <= cv) {
            out = qn * dsqrt(total / cv);
            o.regime = 1;
            o.rel_frac = total / cv;
        }
        if ((total > cv) && (total <= av)) {
            const REAL frac2 = (total - cv) / (av - cv);
            out = qn + dexp((REAL)3.0 * dlog(frac2)) * (qa - qn);
            o.regime = 2;
            o.rel_frac = frac2;
        }
        if ((total > av) && (total <= ev)) {
            o.rel_flood = inflow >= qf;
            const REAL frac3 = (total - av) / (ev - av);
            const REAL tmp = qa + dexp((REAL)0.1 * dlog(frac3)) * (qf - qa);
            o.rel_frac = frac3;
            if (o.rel_flood) {
                const REAL o_flood = qn + ((total - cv) / (ev - cv)) * (inflow - qn);
                o.rel_oflood = (o_flood > tmp || tmp != tmp);
                out = vmax(o_flood, tmp);
            } else {
                out = tmp;
            }
            o.regime = 3;
        }
        if (total > ev) {
            o.rel_spill = inflow >= qf;
            out = o.rel_spill ? inflow : qf;
            o.regime = 4;
        }
        o.rel_raw = out;
        o.rel_pos = out > (REAL)0.0;
        const REAL lo = vmax(out, (REAL)0.0);
        o.rel_cap = !(lo < total / time_step);
        q_r = vclamp(out, (REAL)0.0, total / time_step);
        q_f = (REAL)0.0;
    }
    o.qr = q_r;
    o.qf = q_f;
    return o;
}

// protected water surface of catchment i from its river surface wse, as
// protected_surface_of forms it (levee.py:565-600)
template <class REAL, class STO>
__device__ __forceinline__ REAL adj_pwse(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t i, REAL wse) {
    if (!v.is_levee[i]) return wse;
    return vmin(v.elev[i] + a.pd[i], wse);
}
// the same two in double, as surface_of_d / protected_surface_of_d form them
// with precise surfaces
template <class REAL, class STO>
__device__ __forceinline__ double adj_surface_d(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t i) {
    return (double)a.rd[i] + ((double)v.elev[i] - (double)v.river_height[i]);
}
template <class REAL, class STO>
__device__ __forceinline__ double adj_pwse_d(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t i, double w) {
    if (!v.is_levee[i]) return w;
    return vmin((double)v.elev[i] + (double)a.pd[i], w);
}

// ------------------------------------------------------- path trace ----
// The arithmetic of bif_path (non-levee variant, end surfaces and totals
// formed from the state) for path p on the saved input state, same
// operations in the same order, keeping what the reverse needs.
template <class REAL>
struct PathTrace {
    int64_t ci, di;
    bool max_own, slope_in, min_own, limited;
    REAL w0, w1, max_w, slope, ts0, ts1, sum, ratio, limit, sum_l;
    REAL upd[kAdjMaxLevels], sq_a[kAdjMaxLevels], sq_b[kAdjMaxLevels], sifd[kAdjMaxLevels];
    REAL unit[kAdjMaxLevels], den[kAdjMaxLevels], q[kAdjMaxLevels];
};

template <class REAL, class STO>
__device__ __forceinline__ void path_trace(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t p,
                                           REAL time_step, PathTrace<REAL>& o) {
    const int LB = (int)v.LB;
    o.ci = v.bif_ci[p];
    o.di = v.bif_di[p];
    const int64_t ci = o.ci, di = o.di;
    const REAL gravity = v.gravity;
    const REAL length = v.bif_length[p];
    o.w0 = a.rd[ci] + (v.elev[ci] - v.river_height[ci]);
    o.w1 = a.rd[di] + (v.elev[di] - v.river_height[di]);
    o.ts0 = (REAL)adj_sto<STO>(a.rs_raw, ci) + (REAL)adj_sto<STO>(a.fs_raw, ci) + adj_ps(v, a, ci);
    o.ts1 = (REAL)adj_sto<STO>(a.rs_raw, di) + (REAL)adj_sto<STO>(a.fs_raw, di) + adj_ps(v, a, di);
    o.max_w = vmax(o.w0, o.w1);
    REAL max_pw = (REAL)0;
    if (a.lev) max_pw = vmax(adj_pwse(v, a, ci, o.w0), adj_pwse(v, a, di, o.w1));
    // with precise surfaces the forward forms the surfaces, their drop and
    // the depths above the sills in double (bif_path)
    const bool prec = precise_surfaces(v);
    double w0d = 0.0, w1d = 0.0, mwd = 0.0, mpwd = 0.0;
    if (prec) {
        w0d = adj_surface_d(v, a, ci);
        w1d = adj_surface_d(v, a, di);
        mwd = vmax(w0d, w1d);
        if (a.lev) mpwd = vmax(adj_pwse_d(v, a, ci, w0d), adj_pwse_d(v, a, di, w1d));
        o.max_own = (w0d > w1d || w1d != w1d);
    } else {
        o.max_own = (o.w0 > o.w1 || o.w1 != o.w1);
    }
    const REAL sx = (prec ? (REAL)(w0d - w1d) : (o.w0 - o.w1)) / length;
    o.slope_in = (sx > (REAL)-0.005) && (sx < (REAL)0.005);
    o.slope = vclamp(sx, (REAL)-0.005, (REAL)0.005);
    REAL sum = (REAL)0;
#pragma unroll
    for (int l = 0; l < kAdjMaxLevels; ++l) {
        if (l >= LB) break;
        const int64_t k = p * LB + l;
        const REAL csd_old = a.bcsd[k];
        const REAL bel = v.bif_elev[k];
        REAL upd, sa, sb = (REAL)0, sifd;
        if (!a.lev) {
            upd = vmax(prec ? (REAL)(mwd - (double)bel) : o.max_w - bel, (REAL)0.0);
            sa = dsqrt(upd * csd_old);
            sb = dsqrt(upd * (REAL)0.01);
            sifd = vmax(sa, sb);
        } else if (l == 0) {
            upd = vmax(prec ? (REAL)(mwd - (double)bel) : o.max_w - bel, (REAL)0.0);
            sa = dsqrt(upd * csd_old);
            sifd = (sa <= (REAL)0.0) ? upd : sa;
        } else {
            upd = vmax(prec ? (REAL)(mpwd - (double)bel) : max_pw - bel, (REAL)0.0);
            sa = (REAL)0;
            sifd = upd;
        }
        const REAL w = v.bif_width[k], man = v.bif_manning[k];
        const REAL unit = a.bq[k] / w;
        REAL q = (REAL)0.0, den = (REAL)1;
        if (sifd > (REAL)1e-5) {
            const REAL num = w * (unit + gravity * time_step * sifd * o.slope);
            den = (REAL)1.0 + gravity * time_step * (man * man) * dabs(unit)
                * ((REAL)1.0 / (sifd * sifd * dcbrt(sifd)));
            q = num / den;
        }
        sum += q;
        o.upd[l] = upd;
        o.sq_a[l] = sa;
        o.sq_b[l] = sb;
        o.sifd[l] = sifd;
        o.unit[l] = unit;
        o.den[l] = den;
        o.q[l] = q;
    }
    o.sum = sum;
    o.min_own = (o.ts0 < o.ts1 || o.ts1 != o.ts1);
    o.ratio = (REAL)0.05 * vmin(o.ts0, o.ts1) / (dabs(sum) * time_step);
    o.limited = o.ratio < (REAL)1.0;
    o.limit = vmin(o.ratio, (REAL)1.0);
    o.sum_l = sum * o.limit;
}

// B0: limited level flows, their sum and the outgoing terms of every path
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_adj_path_fwd(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= v.P) return;
    PathTrace<REAL> o;
    path_trace(v, a, p, a.dt, o);
    const int LB = (int)v.LB;
    REAL raw = (REAL)0;
#pragma unroll
    for (int l = 0; l < kAdjMaxLevels; ++l) {
        if (l >= LB) break;
        const REAL q = o.q[l] * o.limit;
        a.p_qlim[p * LB + l] = q;
        raw += q;
    }
    a.p_raw[p] = raw;
    a.p_pos[p] = vmax(o.sum_l, (REAL)0.0) * a.dt;
    a.p_neg[p] = -vmin(o.sum_l, (REAL)0.0) * a.dt;
}

// Path terms of catchment j's outgoing total, in the order of bif_terms_from.
template <class REAL, class STO>
__device__ __forceinline__ STO adj_path_terms(const View<REAL, STO>& v, const AdjView<REAL>& a, int64_t j, STO acc) {
    for (int32_t k = v.ev_ptr[j]; k < v.ev_ptr[j + 1]; ++k) {
        const int32_t code = v.ev_code[k];
        acc += (STO)((code & 1) ? a.p_neg[code >> 1] : a.p_pos[code >> 1]);
    }
    return acc;
}

// ---------------------------------------------------- levee reverse ----
// Reverse of levee_one_p (levee.py:62-304, partition :18-59) from its trace.
// On entry Lr, Lf, Lrd, Lfd hold the adjoints of the levee stage's river and
// flood storage and depth (Lp, Lpd: protected storage and depth); on return
// Lr, Lf, Lrd, Lfd hold the adjoints of its inputs (the flood stage's river
// and flood storage, river and flood depth) and Lrms that of the bankfull
// storage.  Every storage the cases compare against (base, crown, protected
// threshold, case-3 fill) is the bankfull storage plus terms of the fixed
// profile, so each moves one for one with it.
template <class REAL, class STO>
__device__ __forceinline__ void levee_reverse(const LeveeTrace<REAL>& t, STO rs_hp, STO fs_hp, REAL rl, REAL rw,
                                              REAL crown, double& Lr, double& Lf, double Lp, double& Lrd,
                                              double& Lfd, double Lpd, double& Lrms) {
    (void)crown;
    const STO total_hp = rs_hp + fs_hp;
    // the partition's picks
    const bool keep_r = t.case4 || t.none;
    const bool r_cand_pick = !keep_r && ((STO)t.r_cand < total_hp || total_hp != total_hp);
    const STO r_hp = keep_r ? rs_hp : vmin((STO)t.r_cand, total_hp);
    const bool rem_pos = total_hp - r_hp > (STO)0.0;
    const STO remaining = vmax(total_hp - r_hp, (STO)0.0);
    const STO f_cl = vmax((STO)t.f_cand, (STO)0.0);
    const bool f_cl_pick = (f_cl < remaining || remaining != remaining);
    const STO f_hp = t.case2 ? remaining : (t.none ? fs_hp : vmin(f_cl, remaining));
    const bool p_pos = remaining - f_hp > (STO)0.0;
    // p = max(remaining - f, 0)
    double Lrem = 0.0, Lfh = Lf;
    if (p_pos) {
        Lrem += Lp;
        Lfh -= Lp;
    }
    // f = case2 ? remaining : none ? fs : min(max(f_cand, 0), remaining)
    double Lfc = 0.0, Lrs_in = 0.0, Lfs_in = 0.0;
    if (t.case2) Lrem += Lfh;
    else if (t.none) Lfs_in += Lfh;
    else if (f_cl_pick) {
        if ((STO)t.f_cand > (STO)0.0) Lfc += Lfh;
    } else {
        Lrem += Lfh;
    }
    // remaining = max(total - r, 0)
    double Ltot_hp = 0.0, Lrh = Lr;
    if (rem_pos) {
        Ltot_hp += Lrem;
        Lrh -= Lrem;
    }
    // r = case4 or none ? rs : min(r_cand, total)
    double Lrc = 0.0;
    if (keep_r) Lrs_in += Lrh;
    else if (r_cand_pick) Lrc += Lrh;
    else Ltot_hp += Lrh;
    // f_cand = max(f_cand_raw, 0)
    const double Lfr = (t.f_cand_raw > (REAL)0.0) ? Lfc : 0.0;
    const double drl = (double)rl, drw = (double)rw;
    double Ltotal = 0.0, Lrdep = 0.0, Lfdep = 0.0;
    Lrms = 0.0;
    if (t.case2) {
        // f_dph = base + (total - base_sto) / (ld + rw) / rl;
        // r_sto = rms + rl rw f_dph; f_cand_raw = total - r_sto; r_dph = r_sto / rl / rw
        const double Lrsto = Lrc + Lrd / (drl * drw) - Lfr;
        Ltotal += Lfr;
        Lrms += Lrsto;
        const double Lfd2 = Lfd + Lrsto * drl * drw;
        const double c = 1.0 / (((double)t.levee_distance + drw) * drl);
        Ltotal += Lfd2 * c;
        Lrms -= Lfd2 * c;
    } else if (t.case3) {
        // r_sto = rms + rl rw crown; f_cand_raw = s_top - r_sto; r_dph = r_sto / rl / rw;
        // f_dph = crown; p_dph from total - fill storage
        const double Lrsto = Lrc + Lrd / (drl * drw) - Lfr;
        Lrms += Lrsto + Lfr;
        double Ladd = 0.0;
        if (t.found_B) {
            if (t.term_B > (REAL)0.0) {
                const double g = (double)t.gradient_B;
                Ladd = Lpd * g * 0.5 / sqrt((double)t.term_B) * 2.0 / drl / (g + 1e-9);
            }
        } else {
            Ladd = Lpd / ((double)t.dwth_fil_B * drl + 1e-9);
        }
        Ltotal += Ladd;
        if (t.fil_moves) Lrms -= Ladd;
    } else if (t.case4) {
        // f_cand_raw = s_top + (f_dph - crown) (ld + rw) rl - rs; depths: the flood
        // stage's river and flood depth, protected depth = flood depth
        Lrms += Lfr;
        Lfdep += Lfr * ((double)t.levee_distance + drw) * drl;
        Lrs_in -= Lfr;
        Lrdep += Lrd;
        Lfdep += Lfd + Lpd;
    } else {
        Lrdep += Lrd;
        Lfdep += Lfd;
    }
    Ltot_hp += Ltotal;
    Lr = Lrs_in + Ltot_hp;
    Lf = Lfs_in + Ltot_hp;
    Lrd = Lrdep;
    Lfd = Lfdep;
}

// reservoir row of each reservoir catchment
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_adj_res_rows(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= v.R) return;
    a.res_row[v.res_idx[r]] = (int32_t)r;
}

// A0: unlimited outflows and outgoing terms of the saved state
template <class REAL, class STO, bool RES>
__global__ __launch_bounds__(kBlock) void k_adj_outflow_fwd(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.N) return;
    const OutTrace<REAL> o = outflow_trace<REAL, STO, RES>(v, a, i, a.dt);
    a.w_qr[i] = o.qr;
    a.w_qf[i] = o.qf;
    const REAL pos = vmax(o.qr, (REAL)0.0) + vmax(o.qf, (REAL)0.0);
    const REAL neg = vmin(o.qr, (REAL)0.0) + vmin(o.qf, (REAL)0.0);
    a.pos[i] = pos * a.dt;
    a.neg[i] = o.mouth ? (REAL)0.0 : -neg * a.dt;
}

// A1: outgoing totals (upstream negative terms in order, then the own
// positive term) and storage limiters, as k_limiter
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_adj_limiter_fwd(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= v.N) return;
    STO acc = (STO)0;
    const int2 rg = v.up_range(j);
    for (int32_t k = rg.x; k < rg.y; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == j) continue;
        acc += (STO)a.neg[u];
    }
    acc += (STO)a.pos[j];
    if (a.bif) acc = adj_path_terms(v, a, j, acc);
    const REAL og = (REAL)acc;
    const REAL rate = (REAL)(adj_sto<STO>(a.rs_raw, j) + adj_sto<STO>(a.fs_raw, j));
    a.w_lim[j] = (og > (REAL)1e-8) ? vmin(rate / og, (REAL)1.0) : (REAL)1.0;
}

// ------------------------------------------------------ stage reverse ----
// A2: recompute the limited flows, inflow sums, storage update and stage of
// catchment j (k_stage_fused), then reverse them.  Writes the adjoints of the
// limited own flows (a_rout/a_fout, including the next state's direct use),
// of the inflow sums (a_ri/a_fi), of the old storages, and dJ/dH (stage part).
template <class REAL, class STO, bool CONTIG, int LFIX, bool PENDING>
__global__ __launch_bounds__(kBlock) void k_adj_stage(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= v.N) return;
    const REAL dt = a.dt;
    const int64_t d = v.dn[j];
    const REAL l_own = a.lim[j];
    const REAL qr = a.qr[j], qf = a.qf[j];
    // the downstream limiter scales only the catchment's own negative flows
    const REAL l_dn = (!(qr >= (REAL)0.0) || !(qf >= (REAL)0.0)) ? a.lim[d] : (REAL)1.0;
    const REAL rout = (qr >= (REAL)0.0) ? qr * l_own : qr * l_dn;
    const REAL fout = (qf >= (REAL)0.0) ? qf * l_own : qf * l_dn;
    STO ri = (STO)0, fi = (STO)0;
    double pend = 0.0;   // PENDING: downstream water-surface adjoints of the later sub-step
    const int2 rg = v.up_range(j);
    for (int32_t k = rg.x; k < rg.y; ++k) {
        const int64_t u = CONTIG ? (int64_t)k : (int64_t)v.up_idx[k];
        if (CONTIG && u == j) continue;
        const REAL lu = a.lim[u];
        const REAL ur0 = a.qr[u], uf0 = a.qf[u];
        ri += (STO)((ur0 >= (REAL)0.0) ? ur0 * lu : ur0 * l_own);
        fi += (STO)((uf0 >= (REAL)0.0) ? uf0 * lu : uf0 * l_own);
        if (PENDING) pend += (double)a.d_wd[u];
    }
    const REAL rin = (REAL)ri, fin = (REAL)fi;
    // bifurcation total: the path nets in path order, as bif_nets forms them
    STO gsum = (STO)0;
    // a byte says whether j has paths (and a levee) before any index is read
    const bool events = a.bif && (v.ev_flag ? v.ev_flag[j] != 0 : v.ev_ptr[j] != v.ev_ptr[j + 1]);
    if (events) {
        for (int32_t k = v.ev_ptr[j]; k < v.ev_ptr[j + 1]; ++k) {
            const int32_t code = v.ev_code[k];
            const REAL l_other = a.lim[v.ev_other[k]];
            const bool down_end = (code & 1) != 0;
            const REAL l0 = down_end ? l_other : l_own, l1 = down_end ? l_own : l_other;
            const REAL raw = a.p_raw[code >> 1];
            const REAL net = (raw >= (REAL)0.0) ? raw * l0 : raw * l1;
            gsum += down_end ? (STO)(-net) : (STO)net;
        }
    }
    const REAL gb = (REAL)gsum;
    const STO rs = adj_sto<STO>(a.rs_raw, j), fs = adj_sto<STO>(a.fs_raw, j);
    const int32_t lv = (a.lev && v.is_levee[j]) ? v.levee_index[j] : -1;
    const STO ps = (lv >= 0) ? adj_sto<STO>(a.ps_raw, j) : (STO)0.0;
    const REAL ro = v.runoff[j];
    // storage update (storage.py:81-94)
    const STO sa = rs + (STO)((rin - rout) * dt);
    const bool carry = sa < (STO)0.0;
    const STO sb = fs + (carry ? sa : (STO)0.0) + (STO)((fin - fout - gb) * dt);
    const STO rsu1 = vmax(sa, (STO)0.0);
    const bool b_neg = sb < (STO)0.0;
    const STO rsu2 = b_neg ? vmax(rsu1 + sb, (STO)0.0) : rsu1;
    const STO fsu1 = vmax(sb, (STO)0.0);
    const STO total_next = rsu2 + fsu1 + ps + (STO)(ro * dt) + (STO)((REAL)0.0 * dt);
    const STO total_hp = vmax(total_next, (STO)0.0);
    const REAL total = (REAL)total_hp;
    // stage diagnosis (storage.py:105-181)
    const REAL rh = v.river_height[j], rw = v.river_width[j], rl = v.river_length[j];
    const REAL rms = rl * rw * rh;
    const int L = LFIX > 0 ? LFIX : (int)v.L;
    double Ltot = 0.0, LH = 0.0;
    double Lrs_out = (double)a.lout[ADJ_RS][j], Lfs_out = (double)a.lout[ADJ_FS][j];
    double Lrd_out = (double)a.lout[ADJ_RD][j] + pend, Lfd_out = (double)a.lout[ADJ_FD][j];
    // forward of the stage diagnosis: flood depth and river storage
    const bool wet = total > rms;
    TableRow<REAL, LFIX> lrow;
    REAL fdep = (REAL)0.0;
    double dfd_dtot = 0.0, dfd_dpts = 0.0;
    bool cap_pick = false;
    STO rsf = total_hp;
    if (wet) {
        const REAL ca = v.catchment_area[j];
        const REAL* row = v.table + j * v.L;
        const REAL catchment_width = ca / rl;
        const REAL winc = catchment_width / (REAL)L;
        int level = 0;
        REAL S_accum = rms, prev_H = (REAL)0.0, prev_W = rw, pts = rms, pfd = (REAL)0.0, nfd = (REAL)0.0;
        for (int k = 0; k < L; ++k) {
            const REAL H = row[k];
            const REAL W = rw + (REAL)(k + 1) * winc;
            const REAL dS = rl * (REAL)0.5 * (prev_W + W) * (H - prev_H);
            const REAL S = S_accum + dS;
            nfd = H;
            if (!(total > S)) break;
            level += 1;
            pts = S;
            pfd = H;
            S_accum = S;
            prev_H = H;
            prev_W = W;
        }
        const REAL ptw = rw + (REAL)level * winc;
        if (level == L) {
            fdep = pfd + (total - pts) / (ptw * rl);
            dfd_dtot = 1.0 / ((double)ptw * (double)rl);
            dfd_dpts = -dfd_dtot;
        } else {
            const REAL fg = (nfd - pfd) / winc;
            const REAL A = ptw * ptw + (REAL)2.0 * (total - pts) / (fg * rl);
            const REAL dw = dsqrt(A) - ptw;
            fdep = pfd + dw * fg;
            const double sq = sqrt((double)A);
            dfd_dtot = 1.0 / (sq * (double)rl);
            dfd_dpts = -dfd_dtot;
        }
        const REAL cap = rms + rl * rw * fdep;
        cap_pick = ((STO)cap < total_hp);
        rsf = cap_pick ? (STO)cap : total_hp;
    }
    if (lv >= 0) {
        // levee stage on the stage's output, then its reverse: the adjoints of
        // the stage's river / flood storage and depths, and of the bankfull
        // storage (so of the river height)
        const STO fsf = vmax(total_hp - rsf, (STO)0.0);
        const REAL rdep = (REAL)rsf / (rl * rw);
        lrow.load(v.table + j * v.L);
        LeveeTrace<REAL> tr;
        (void)levee_one_p<REAL, STO, LFIX, true>(v, rl, rw, rh, v.catchment_area[j], v.levee_crown[lv],
                                                 v.levee_frac[lv], v.levee_base[lv], rsf, fsf, fdep, rdep,
                                                 (REAL)0.0, lrow, &tr);
        double Lrms = 0.0;
        levee_reverse<REAL, STO>(tr, rsf, fsf, rl, rw, v.levee_crown[lv], Lrs_out, Lfs_out,
                                 (double)a.lout_ps[j], Lrd_out, Lfd_out, (double)a.lout_pd[j], Lrms);
        LH += Lrms * (double)rl * (double)rw;
    }
    double Lrsf = Lrs_out + Lrd_out / ((double)rl * (double)rw);
    if (!wet) {
        // dry: rsf = total, flood depth 0, flood storage max(0, 0)
        Ltot += Lrsf;
    } else {
        if (total_hp - rsf > (STO)0.0) {
            Ltot += Lfs_out;
            Lrsf -= Lfs_out;
        }
        double Lcap = 0.0;
        if (cap_pick) Lcap = Lrsf;
        else Ltot += Lrsf;
        const double rlrw = (double)rl * (double)rw;
        const double Lfdep = Lfd_out + Lcap * rlrw;
        LH += Lcap * rlrw;
        Ltot += Lfdep * dfd_dtot;
        LH += Lfdep * dfd_dpts * rlrw;
    }
    // total_hp = max(total_next, 0); total_next = rsu2 + fsu1 + ...
    const double Ltn = (total_next > (STO)0.0) ? Ltot : 0.0;
    // the protected storage of a catchment without a levee stays zero: its
    // adjoint is never read
    if (lv >= 0) a.lin_ps[j] = (REAL)Ltn;
    double Lrsu1 = 0.0, Lb = 0.0, La = 0.0;
    if (sb > (STO)0.0) Lb += Ltn;
    if (b_neg) {
        if (rsu1 + sb > (STO)0.0) { Lrsu1 += Ltn; Lb += Ltn; }
    } else {
        Lrsu1 += Ltn;
    }
    if (sa > (STO)0.0) La += Lrsu1;
    double Lfs = Lb, Lfin = Lb * (double)dt, Lfout = -Lb * (double)dt;
    if (carry) La += Lb;
    double Lrs = La, Lrin = La * (double)dt, Lrout = -La * (double)dt;
    a.a_rout[j] = (REAL)(Lrout + (double)a.lout[ADJ_QR][j]);
    a.a_fout[j] = (REAL)(Lfout + (double)a.lout[ADJ_QF][j]);
    a.a_ri[j] = (REAL)Lrin;
    a.a_fi[j] = (REAL)Lfin;
    // read by the path reverse and the limiter reverse at path ends only
    if (events) a.a_gb[j] = (REAL)(-Lb * (double)dt);
    a.lin[ADJ_RS][j] = (REAL)Lrs;
    a.lin[ADJ_FS][j] = (REAL)Lfs;
    a.g_h[j] += LH - pend;
}

// A3: adjoints of the unlimited flows and of the limiter of catchment k.
// The limited outflow of k enters its own storage update and the inflow sum
// of its downstream catchment; positive flow is limited by k, negative flow
// by the downstream catchment (outflow.py:364-365).  Then the limiter
// reverse (outflow.py:346-352) gives the adjoint of the outgoing total.
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_adj_route(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= v.N) return;
    const int64_t d = v.dn[k];
    const bool mouth = (d == k);
    const REAL qr = a.qr[k], qf = a.qf[k];
    const REAL l_own = a.lim[k];
    const REAL l_dn = (!(qr >= (REAL)0.0) || !(qf >= (REAL)0.0)) ? a.lim[d] : (REAL)1.0;
    // a reservoir downstream also gathers the limited flows into its inflow
    const double Lres = (a.res && !mouth && v.is_reservoir[d]) ? (double)a.lout_ri[d] : 0.0;
    const double LR = (double)a.a_rout[k] + (mouth ? 0.0 : (double)a.a_ri[d]) + Lres;
    const double LF = (double)a.a_fout[k] + (mouth ? 0.0 : (double)a.a_fi[d]) + Lres;
    a.a_qr[k] = (REAL)(LR * (double)((qr >= (REAL)0.0) ? l_own : l_dn));
    a.a_qf[k] = (REAL)(LF * (double)((qf >= (REAL)0.0) ? l_own : l_dn));
    double Llim = 0.0;
    if (qr >= (REAL)0.0 || mouth) Llim += LR * (double)qr;
    if (qf >= (REAL)0.0 || mouth) Llim += LF * (double)qf;
    // The limiter binds (og > 1e-8 and rate / og < 1) exactly where the
    // forward's limiter min(rate / og, 1) is below 1; elsewhere its adjoint
    // is zero and nothing below is needed.
    if (!(l_own < (REAL)1.0)) {
        a.a_og[k] = (REAL)0.0;
        return;
    }
    const double Lri_k = (double)a.a_ri[k], Lfi_k = (double)a.a_fi[k];
    // the outgoing total of k, as the forward formed it: upstream negative
    // terms in order, then k's own positive term
    const REAL dt = a.dt;
    STO og_acc = (STO)0;
    const int2 rg = v.up_range(k);
    for (int32_t p = rg.x; p < rg.y; ++p) {
        const int64_t u = CONTIG ? (int64_t)p : (int64_t)v.up_idx[p];
        if (CONTIG && u == k) continue;
        const REAL uq = a.qr[u], uf = a.qf[u];
        og_acc += (STO)(-(vmin(uq, (REAL)0.0) + vmin(uf, (REAL)0.0)) * dt);
        if (uq < (REAL)0.0) Llim += ((double)a.a_rout[u] + Lri_k) * (double)uq;
        if (uf < (REAL)0.0) Llim += ((double)a.a_fout[u] + Lfi_k) * (double)uf;
    }
    og_acc += (STO)((vmax(qr, (REAL)0.0) + vmax(qf, (REAL)0.0)) * dt);
    if (a.bif) {
        og_acc = adj_path_terms(v, a, k, og_acc);
    }
    // limiter: lim = og > 1e-8 ? min(rate / og, 1) : 1, rate = rs + fs
    const REAL og = (REAL)og_acc;
    const REAL rate = (REAL)(adj_sto<STO>(a.rs_raw, k) + adj_sto<STO>(a.fs_raw, k));
    double Log = 0.0;
    if (og > (REAL)1e-8 && (rate / og) < (REAL)1.0) {
        if (a.bif) {
            // the limiter of k also scales the nets of its paths
            // (bifurcation.py:137-151: the level sum limited by the upstream
            // end when non-negative, else by the downstream end, added to the
            // upstream end's total and subtracted from the downstream end's)
            // and their limited next level flows (:132-143) on the same rule
            const int LB = (int)v.LB;
            for (int32_t e = v.ev_ptr[k]; e < v.ev_ptr[k + 1]; ++e) {
                const int32_t code = v.ev_code[e];
                const int64_t p = code >> 1;
                const bool down_end = (code & 1) != 0;
                const int64_t other = v.ev_other[e];
                const double Lnet = down_end ? (double)a.a_gb[other] - (double)a.a_gb[k]
                                             : (double)a.a_gb[k] - (double)a.a_gb[other];
                const REAL raw = a.p_raw[p];
                if ((raw >= (REAL)0.0) != down_end) Llim += Lnet * (double)raw;
                for (int l = 0; l < LB; ++l) {
                    const REAL q = a.p_qlim[p * LB + l];
                    if ((q >= (REAL)0.0) != down_end) Llim += (double)a.lout_bq[p * LB + l] * (double)q;
                }
            }
        }
        const double Lrate = Llim / (double)og;
        Log = -Llim * (double)rate / ((double)og * (double)og);
        a.lin[ADJ_RS][k] = (REAL)((double)a.lin[ADJ_RS][k] + Lrate);
        a.lin[ADJ_FS][k] = (REAL)((double)a.lin[ADJ_FS][k] + Lrate);
    }
    a.a_og[k] = (REAL)Log;
}

// B2: reverse of a path's net and limited next level flows (as in A3, their
// share of the adjoint of the limited level flows), its outgoing terms, its
// 5 % storage limiter (bifurcation.py:93-104) and the per-level momentum
// update (:49-89).  Gives
// the adjoint of the input path state and, per end, of the end's water
// surface and total storage.  Eight lanes per path, lane l on level l: the
// level sums pass between the lanes of the group by shuffles, in level order
// (the forward's order for its sum), so the recomputed values are the
// forward's and every result is independent of the launch.
template <class REAL, class STO>
__global__ __launch_bounds__(kBlock) void k_adj_path(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p = gid >> 3;
    const int l = (int)(gid & 7);
    if (p >= v.P) return;   // whole groups of eight leave together
    const int LB = (int)v.LB;
    const bool lv = l < LB;
    const int64_t ci = v.bif_ci[p], di = v.bif_di[p];
    const REAL time_step = a.dt;
    const REAL gravity = v.gravity;
    const REAL length = v.bif_length[p];
    // path values (every lane of the group forms the same ones)
    const REAL w0 = a.rd[ci] + (v.elev[ci] - v.river_height[ci]);
    const REAL w1 = a.rd[di] + (v.elev[di] - v.river_height[di]);
    const REAL ts0 = (REAL)adj_sto<STO>(a.rs_raw, ci) + (REAL)adj_sto<STO>(a.fs_raw, ci) + adj_ps(v, a, ci);
    const REAL ts1 = (REAL)adj_sto<STO>(a.rs_raw, di) + (REAL)adj_sto<STO>(a.fs_raw, di) + adj_ps(v, a, di);
    const REAL max_w = vmax(w0, w1);
    // with levees the upper levels run on the protected surfaces
    REAL pw0 = w0, pw1 = w1, max_pw = (REAL)0;
    if (a.lev) {
        pw0 = adj_pwse(v, a, ci, w0);
        pw1 = adj_pwse(v, a, di, w1);
        max_pw = vmax(pw0, pw1);
    }
    // with precise surfaces the forward forms the surfaces, their drop and
    // the depths above the sills in double (bif_path): every pick below
    // follows those values
    const bool prec = precise_surfaces(v);
    double w0d = 0.0, w1d = 0.0, pw0d = 0.0, pw1d = 0.0, mwd = 0.0, mpwd = 0.0;
    if (prec) {
        w0d = adj_surface_d(v, a, ci);
        w1d = adj_surface_d(v, a, di);
        mwd = vmax(w0d, w1d);
        if (a.lev) {
            pw0d = adj_pwse_d(v, a, ci, w0d);
            pw1d = adj_pwse_d(v, a, di, w1d);
            mpwd = vmax(pw0d, pw1d);
        }
    }
    const bool max_own = prec ? (w0d > w1d || w1d != w1d) : (w0 > w1 || w1 != w1);
    const bool upper = a.lev && l > 0;   // this lane's level on the protected surface
    const REAL sx = (prec ? (REAL)(w0d - w1d) : (w0 - w1)) / length;
    const bool slope_in = (sx > (REAL)-0.005) && (sx < (REAL)0.005);
    const REAL slope = vclamp(sx, (REAL)-0.005, (REAL)0.005);
    // level l of the forward
    const int64_t k = p * LB + (lv ? l : 0);
    REAL upd = (REAL)0, sa = (REAL)0, sb = (REAL)0, sifd = (REAL)0, unit = (REAL)0, den = (REAL)1, q = (REAL)0;
    REAL csd_old = (REAL)0, above = (REAL)0;
    if (lv) {
        csd_old = a.bcsd[k];
        const REAL bel = v.bif_elev[k];
        // the depth above the sill: of the higher river surface, or of the
        // higher protected surface on the upper levels with levees
        if (!upper) above = prec ? (REAL)(mwd - (double)bel) : max_w - bel;
        else above = prec ? (REAL)(mpwd - (double)bel) : max_pw - bel;
        upd = vmax(above, (REAL)0.0);
        if (!a.lev) {
            sa = dsqrt(upd * csd_old);
            sb = dsqrt(upd * (REAL)0.01);
            sifd = vmax(sa, sb);
        } else if (!upper) {
            sa = dsqrt(upd * csd_old);
            sifd = (sa <= (REAL)0.0) ? upd : sa;
        } else {
            sifd = upd;
        }
        const REAL w = v.bif_width[k], man = v.bif_manning[k];
        unit = a.bq[k] / w;
        if (sifd > (REAL)1e-5) {
            const REAL num = w * (unit + gravity * time_step * sifd * slope);
            den = (REAL)1.0 + gravity * time_step * (man * man) * dabs(unit)
                * ((REAL)1.0 / (sifd * sifd * dcbrt(sifd)));
            q = num / den;
        }
    }
    REAL sum = (REAL)0;
    for (int m = 0; m < LB; ++m) sum += __shfl(q, m, 8);
    const REAL ratio = (REAL)0.05 * vmin(ts0, ts1) / (dabs(sum) * time_step);
    const bool limited = ratio < (REAL)1.0;
    const REAL limit = vmin(ratio, (REAL)1.0);
    const REAL sum_l = sum * limit;
    const double dt = (double)time_step, g = (double)gravity;
    // outgoing terms: max(sum_l, 0) dt at the upstream end, -min(sum_l, 0) dt
    // at the downstream end
    double Lsum_l = 0.0;
    if (sum_l > (REAL)0.0) Lsum_l += (double)a.a_og[ci] * dt;
    if (sum_l < (REAL)0.0) Lsum_l -= (double)a.a_og[di] * dt;
    // the limited level flow q_l * limit enters the path net (level sum raw)
    // and the next level flow, each limited by one end
    double Llq = 0.0;
    if (lv) {
        const REAL l0 = a.lim[ci], l1 = a.lim[di];
        const double Lnet = (double)a.a_gb[ci] - (double)a.a_gb[di];
        const REAL raw = a.p_raw[p], ql = a.p_qlim[k];
        Llq = Lnet * (double)(raw >= (REAL)0.0 ? l0 : l1)
            + (double)a.lout_bq[k] * (double)(ql >= (REAL)0.0 ? l0 : l1);
    }
    // sum_l = sum * limit, limited level flow q_l * limit
    const double mine = Llq * (double)q;
    double Llimit = 0.0;
    for (int m = 0; m < LB; ++m) Llimit += __shfl(mine, m, 8);
    Llimit += Lsum_l * (double)sum;
    // limit = min(0.05 min(ts0, ts1) / (|sum| dt), 1)
    double Lts0 = 0.0, Lts1 = 0.0, Lsum = 0.0;
    if (limited) {
        const double as = fabs((double)sum);
        const double Lm = Llimit * 0.05 / (as * dt);
        if (ts0 < ts1 || ts1 != ts1) Lts0 += Lm;
        else Lts1 += Lm;
        const double sgn = sum > (REAL)0.0 ? 1.0 : (sum < (REAL)0.0 ? -1.0 : 0.0);
        Lsum = -Llimit * (double)ratio / as * sgn;
    }
    const double Lql = (Llq + Lsum_l) * (double)limit + Lsum;
    double Lqold = 0.0, Lsifd = 0.0, Lcsd = 0.0, Lslope = 0.0, Lmax_w = 0.0, Lmax_pw = 0.0;
    if (lv) {
        if (sifd > (REAL)1e-5) {
            const double w = (double)v.bif_width[k], man = (double)v.bif_manning[k];
            const double s = (double)sifd, dd = (double)den;
            const double rdd = 1.0 / dd;
            const double Lnum = Lql * rdd, Lden = -Lql * (double)q * rdd;
            const double r3 = rcbrt_pos_d(s);
            const double rs3 = r3 * r3 * r3;     // 1 / s
            const double Q = rs3 * rs3 * r3;     // s^(-7/3)
            const double u = (double)unit, au = fabs(u);
            const double sgn = u > 0.0 ? 1.0 : (u < 0.0 ? -1.0 : 0.0);
            // num = w (q_old / w + g dt sifd slope)
            Lqold += Lnum;
            Lsifd += Lnum * w * g * dt * (double)slope;
            Lslope += Lnum * w * g * dt * s;
            // den = 1 + g dt n^2 |q_old / w| sifd^(-7/3)
            Lqold += Lden * g * dt * man * man * sgn * Q / w;
            Lsifd += Lden * g * dt * man * man * au * (-7.0 / 3.0) * (Q * rs3);
        }
        // the next depth is upd; sifd = max(sqrt(upd csd_old), sqrt(upd 0.01)),
        // with levees sqrt(upd csd_old) (upd where that is not positive) at
        // level 0 and upd above
        double Lupd = (double)a.lout_bcsd[k];
        if (upper) {
            Lupd += Lsifd;
        } else if (a.lev) {
            if (sa > (REAL)0.0) {
                Lupd += Lsifd * 0.5 * (double)csd_old / (double)sa;
                Lcsd += Lsifd * 0.5 * (double)upd / (double)sa;
            } else {
                Lupd += Lsifd;
            }
        } else if (sa > sb || sb != sb) {
            if (sa > (REAL)0.0) {
                Lupd += Lsifd * 0.5 * (double)csd_old / (double)sa;
                Lcsd += Lsifd * 0.5 * (double)upd / (double)sa;
            }
        } else if (sb > (REAL)0.0) {
            Lupd += Lsifd * 0.5 * 0.01 / (double)sb;
        }
        // upd = max(surface - elevation, 0)
        if (above > (REAL)0.0) {
            if (upper) Lmax_pw += Lupd;
            else Lmax_w += Lupd;
        }
        a.lin_bq[k] = (REAL)Lqold;
        a.lin_bcsd[k] = (REAL)Lcsd;
    }
    double Ls = 0.0, Lx = 0.0, Lxp = 0.0;
    for (int m = 0; m < LB; ++m) {
        Ls += __shfl(Lslope, m, 8);
        Lx += __shfl(Lmax_w, m, 8);
        Lxp += __shfl(Lmax_pw, m, 8);
    }
    if (l != 0) return;
    // slope = clamp((w0 - w1) / length, -0.005, 0.005); max_w = max(w0, w1)
    double Lw0 = 0.0, Lw1 = 0.0;
    if (slope_in) {
        const double len = (double)length;
        Lw0 += Ls / len;
        Lw1 -= Ls / len;
    }
    if (max_own) Lw0 += Lx;
    else Lw1 += Lx;
    if (a.lev) {
        // max_pw = max(pw0, pw1); pw = min(elevation + protected depth, wse)
        // at a levee end, wse elsewhere
        double Lp0 = 0.0, Lp1 = 0.0, Ld0 = 0.0, Ld1 = 0.0;
        bool p0_max, lev0, lev1;
        if (prec) {
            const double e0 = (double)v.elev[ci] + (double)a.pd[ci], e1 = (double)v.elev[di] + (double)a.pd[di];
            p0_max = (pw0d > pw1d || pw1d != pw1d);
            lev0 = v.is_levee[ci] && (e0 < w0d || w0d != w0d);
            lev1 = v.is_levee[di] && (e1 < w1d || w1d != w1d);
        } else {
            const REAL e0 = v.elev[ci] + a.pd[ci], e1 = v.elev[di] + a.pd[di];
            p0_max = (pw0 > pw1 || pw1 != pw1);
            lev0 = v.is_levee[ci] && (e0 < w0 || w0 != w0);
            lev1 = v.is_levee[di] && (e1 < w1 || w1 != w1);
        }
        if (p0_max) Lp0 = Lxp;
        else Lp1 = Lxp;
        if (lev0) Ld0 = Lp0;
        else Lw0 += Lp0;
        if (lev1) Ld1 = Lp1;
        else Lw1 += Lp1;
        a.p_d0[p] = (REAL)Ld0;
        a.p_d1[p] = (REAL)Ld1;
    }
    a.p_w0[p] = (REAL)Lw0;
    a.p_w1[p] = (REAL)Lw1;
    a.p_t0[p] = (REAL)Lts0;
    a.p_t1[p] = (REAL)Lts1;
}

// A4: reverse of the outflow momentum update of catchment i.  The final
// flows enter the limited flows (a_qr/a_qf) and the outgoing terms (own total
// for the positive part, the downstream total for the negative part); the
// cross-section depths enter the next sub-step.
template <class REAL, class STO, bool RES>
__global__ __launch_bounds__(kBlock) void k_adj_outflow(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.N) return;
    const REAL dtr = a.dt;
    const OutTrace<REAL> o = outflow_trace<REAL, STO, RES>(v, a, i, dtr);
    const double dt = (double)dtr, g = (double)v.gravity;
    const int64_t d = v.dn[i];
    const double og_own = (double)a.a_og[i], og_dn = o.mouth ? 0.0 : (double)a.a_og[d];
    double Lqr = (double)a.a_qr[i], Lqf = (double)a.a_qf[i];
    if (o.qr > (REAL)0.0) Lqr += og_own * dt;
    if (o.qf > (REAL)0.0) Lqf += og_own * dt;
    if (o.qr < (REAL)0.0) Lqr -= og_dn * dt;
    if (o.qf < (REAL)0.0) Lqf -= og_dn * dt;
    // a reservoir's release or the kinematic flow upstream of a dam replaces
    // the momentum update's flows, whose reverse then gets no adjoint
    double Lrs_x = 0.0, Lfs_x = 0.0, Lrd_x = 0.0, Lfd_x = 0.0, gn_x = 0.0, Lri = 0.0;
    if (o.reservoir) {
        const int64_t r = a.res_row[i];
        const double cv = (double)v.res_cons[r], ev = (double)v.res_emer[r], av = (double)v.res_adj[r];
        const double qn = (double)v.res_qn[r], qa = (double)v.res_qa[r], qf = (double)v.res_qf[r];
        const double total = (double)o.rel_total, fr = (double)o.rel_frac;
        double Ltotal = 0.0, Lraw = 0.0;
        // q = clamp(raw, 0, total / dt)
        if (o.rel_cap) Ltotal += Lqr / dt;
        else if (o.rel_pos) Lraw = Lqr;
        if (o.regime == 1) {
            if (fr > 0.0) Ltotal += Lraw * qn * 0.5 / sqrt(fr) / cv;
        } else if (o.regime == 2) {
            Ltotal += Lraw * 3.0 * fr * fr * (qa - qn) / (av - cv);
        } else if (o.regime == 3) {
            if (o.rel_flood && o.rel_oflood) {
                Ltotal += Lraw * ((double)o.rel_inflow - qn) / (ev - cv);
                Lri += Lraw * (total - cv) / (ev - cv);
            } else if (fr > 0.0) {
                Ltotal += Lraw * 0.1 * pow(fr, -0.9) * (qf - qa) / (ev - av);
            }
        } else if (o.regime == 4) {
            if (o.rel_spill) Lri += Lraw;
        }
        Lrs_x += Ltotal;
        Lfs_x += Ltotal;
        Lqr = 0.0;
        Lqf = 0.0;
    } else if (o.dam) {
        // q_r = max(min(W h sqrt(S) / n h^(2/3), rs / dt), 0), q_f likewise
        const double h = (double)a.rd[i], fdp = (double)a.fd[i];
        if (o.kr_pos) {
            if (o.kr_cap) {
                Lrs_x += Lqr / dt;
            } else {
                const double kr = (double)o.kr_raw;
                gn_x -= Lqr * kr / (double)v.river_manning[i];
                if (h > 0.0) Lrd_x += Lqr * (5.0 / 3.0) * kr / h;
            }
        }
        if (o.kf_pos) {
            if (o.kf_cap) {
                Lfs_x += Lqf / dt;
            } else {
                const double kf = (double)o.kf_raw;
                if (o.kf_area_raw > (REAL)0.0) {
                    const double Larea = Lqf * kf / (double)o.kf_area_raw;
                    Lfs_x += Larea / (double)v.river_length[i];
                    Lfd_x -= Larea * (double)v.river_width[i];
                }
                if (fdp > 0.0) Lfd_x += Lqf * (2.0 / 3.0) * kf / fdp;
            }
        }
        Lqr = 0.0;
        Lqf = 0.0;
    }
    // negative-flow limiter at 5 % of the storage
    double Lqr0, Lqf1, Ltot = 0.0;
    if (o.negative) {
        Lqr0 = Lqr * (double)o.limit;
        Lqf1 = Lqf * (double)o.limit;
        if (o.limited) {
            const double Llimit = Lqr * (double)o.qr0 + Lqf * (double)o.qf1;
            Ltot += Llimit * 0.05 / (double)o.total_neg;
            const double Ltn = -Llimit * (double)o.ratio / (double)o.total_neg;
            Lqr0 -= Ltn * dt;
            Lqf1 -= Ltn * dt;
        }
    } else {
        Lqr0 = Lqr;
        Lqf1 = Lqf;
    }
    const double Lqf0 = o.killed ? 0.0 : Lqf1;
    double Lro = 0.0, Lfo = 0.0, Lr_sifd = 0.0, Lf_sifd = 0.0, Lf_imp = 0.0, Lslope = 0.0, Lfslope = 0.0;
    double gn = 0.0;
    // (reciprocals formed once; the cube-root powers as powers of s^(-1/3))
    if (o.flood_cond) {
        const double fo = (double)a.q_f[i];
        const double fm = (double)v.flood_manning[i];
        const double rden = 1.0 / (double)o.den_f;
        const double Lnum = Lqf0 * rden, Lden = -Lqf0 * (double)o.qf0 * rden;
        Lfo += Lnum;
        Lf_imp += Lnum * g * dt * (double)o.fslope;
        Lfslope += Lnum * g * dt * (double)o.f_imp;
        const double cf = g * dt * fm * fm;
        const double fs_ = (double)o.f_sifd, rimp = 1.0 / (double)o.f_imp;
        const double r3 = rcbrt_pos_d(fs_);
        const double rfs = r3 * r3 * r3;   // 1 / fs
        const double rP = rfs * r3;        // 1 / (fs cbrt(fs)) = fs^(-4/3)
        const double sgn = fo > 0.0 ? 1.0 : (fo < 0.0 ? -1.0 : 0.0);
        Lfo += Lden * cf * sgn * (rP * rimp);
        Lf_sifd += Lden * (-cf * fabs(fo) * rimp) * (4.0 / 3.0) * (rP * rfs);
        Lf_imp += Lden * (-cf * fabs(fo) * (rP * rimp * rimp));
    }
    double Lupd_fcsa = (double)a.lout[ADJ_FCSA][i], Lfcsa = 0.0;
    if (o.flood_cond && o.f_imp_raw > (REAL)1e-6) {
        const double h = Lf_imp * 0.5 / (double)o.f_imp_raw;
        Lupd_fcsa += h * (double)o.f_csa_m;
        if (a.fcsa[i] > (REAL)1e-6) Lfcsa += h * (double)o.upd_fcsa;
    }
    if (o.river_cond) {
        const double rw = (double)v.river_width[i], n = (double)v.river_manning[i];
        const double rden = 1.0 / (double)o.den_r;
        const double Lnum = Lqr0 * rden, Lden = -Lqr0 * (double)o.qr0 * rden;
        const double rsf = (double)o.r_sifd;
        Lro += Lnum;
        Lr_sifd += Lnum * rw * g * dt * (double)o.slope;
        Lslope += Lnum * rw * g * dt * rsf;
        const double r3 = rcbrt_pos_d(rsf);
        const double rrs = r3 * r3 * r3;     // 1 / rsf
        const double Q = rrs * rrs * r3;     // 1 / (rsf^2 cbrt(rsf)) = rsf^(-7/3)
        const double unit = (double)o.unit_r, au = fabs(unit);
        const double sgn = unit > 0.0 ? 1.0 : (unit < 0.0 ? -1.0 : 0.0);
        gn += Lden * 2.0 * g * dt * n * au * Q;
        Lro += Lden * g * dt * n * n * sgn / rw * Q;
        Lr_sifd += Lden * g * dt * n * n * au * (-7.0 / 3.0) * (Q * rrs);
    }
    double Lupd_rcsd = (double)a.lout[ADJ_RCSD][i], Lupd_fcsd = (double)a.lout[ADJ_FCSD][i], Lrcsd = 0.0,
           Lfcsd = 0.0;
    if (o.f_sifd_raw > (REAL)1e-6) {
        const double h = Lf_sifd * 0.5 / (double)o.f_sifd_raw;
        Lupd_fcsd += h * (double)a.fcsd[i];
        Lfcsd += h * (double)o.upd_fcsd;
    }
    if (o.r_sifd_raw > (REAL)1e-6) {
        const double h = Lr_sifd * 0.5 / (double)o.r_sifd_raw;
        Lupd_rcsd += h * (double)a.rcsd[i];
        Lrcsd += h * (double)o.upd_rcsd;
    }
    double Lfs = 0.0, Lfd = 0.0, Lrd = 0.0, Lwse = 0.0, Lwse_dn = 0.0, Lmx = 0.0, Lre = 0.0;
    if (o.upd_fcsa_raw > (REAL)0.0) {
        Lfs += Lupd_fcsa / (double)v.river_length[i];
        Lfd -= Lupd_fcsa * (double)v.river_width[i];
    }
    if (o.fsd > (REAL)0.0) {
        if (o.mouth) Lwse += Lupd_fcsd;
        else Lmx += Lupd_fcsd;
    }
    if (o.mouth) {
        Lrd += Lupd_rcsd;
    } else {
        Lmx += Lupd_rcsd;
        Lre -= Lupd_rcsd;
    }
    if (o.slope_in) Lslope += Lfslope;
    double Lpd = 0.0;
    if (a.bif && (v.ev_flag ? v.ev_flag[i] != 0 : v.ev_ptr[i] != v.ev_ptr[i + 1])) {
        // the paths of i read its water surface, total storage and (levee
        // catchment) protected depth (B2)
        for (int32_t e = v.ev_ptr[i]; e < v.ev_ptr[i + 1]; ++e) {
            const int32_t code = v.ev_code[e];
            const int64_t p = code >> 1;
            Lwse += (double)((code & 1) ? a.p_w1[p] : a.p_w0[p]);
            Ltot += (double)((code & 1) ? a.p_t1[p] : a.p_t0[p]);
            if (a.lev) Lpd += (double)((code & 1) ? a.p_d1[p] : a.p_d0[p]);
        }
    }
    if (a.lev && v.is_levee[i]) {
        a.lin_ps[i] = (REAL)((double)a.lin_ps[i] + Ltot);
        a.lin_pd[i] = (REAL)Lpd;
    }
    if (RES) a.lin_ri[i] = (REAL)Lri;
    const double Lsd = Lslope / (double)v.dist[i];
    Lwse += Lsd;
    Lwse_dn -= Lsd;
    if (o.mx_own) Lwse += Lmx;
    else Lwse_dn += Lmx;
    Lrd += Lwse;
    Lre += Lwse;
    a.lin[ADJ_QR][i] = (REAL)Lro;
    a.lin[ADJ_QF][i] = (REAL)Lfo;
    a.lin[ADJ_RCSD][i] = (REAL)Lrcsd;
    a.lin[ADJ_FCSD][i] = (REAL)Lfcsd;
    a.lin[ADJ_FCSA][i] = (REAL)Lfcsa;
    a.lin[ADJ_RS][i] = (REAL)((double)a.lin[ADJ_RS][i] + Ltot + Lrs_x);
    a.lin[ADJ_FS][i] = (REAL)((double)a.lin[ADJ_FS][i] + Ltot + Lfs + Lfs_x);
    a.lin[ADJ_RD][i] = (REAL)(Lrd + Lrd_x);
    a.lin[ADJ_FD][i] = (REAL)(Lfd + Lfd_x);
    a.g_n[i] += gn + gn_x;
    a.g_h[i] += -Lre;
    a.d_wd[i] = o.mouth ? (REAL)0.0 : (REAL)Lwse_dn;
}

// A5: downstream water surface: wse_dn = river_depth[d] + elevation[d] -
// height[d] for every upstream catchment of d.
template <class REAL, class STO, bool CONTIG>
__global__ __launch_bounds__(kBlock) void k_adj_gather(View<REAL, STO> v, AdjView<REAL> a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= v.N) return;
    double s = 0.0;
    const int2 rg = v.up_range(k);
    for (int32_t p = rg.x; p < rg.y; ++p) {
        const int64_t u = CONTIG ? (int64_t)p : (int64_t)v.up_idx[p];
        if (CONTIG && u == k) continue;
        s += (double)a.d_wd[u];
    }
    a.lin[ADJ_RD][k] = (REAL)((double)a.lin[ADJ_RD][k] + s);
    a.g_h[k] -= s;
}

}  // namespace cmf
