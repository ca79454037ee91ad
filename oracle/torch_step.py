"""Plain PyTorch-op transcription of one sub-step of the base physics and the
bifurcation paths, for FP64 autograd (the checker of the hand-written adjoint).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows the same reference arithmetic as cmf_oracle.cpp, with vector ops:
  outflow    cmfgpu/phys/triton/outflow.py:128-245 (local-inertial river and
             floodplain update, opposite-direction kill, 5 % negative limiter)
  limiter    cmfgpu/phys/triton/outflow.py:346-365 (outgoing totals, storage
             limiter, limited flows) and the inflow sums :374-377
  storage    cmfgpu/phys/triton/storage.py:81-94 (update with negative carry)
  stage      cmfgpu/phys/triton/storage.py:105-181 (floodplain table scan,
             quadratic or linear inversion, river/flood split)
  paths      cmfgpu/phys/triton/bifurcation.py:13-104 (per-level flow between
             the end water surfaces, 5 % storage limiter, outgoing terms) and
             :106-151 (level flows and the path net limited by the ends); with
             levees cmfgpu/phys/triton/levee.py:565-684 (level 0 on the river
             surface, upper levels on the protected surface)
  levee      cmfgpu/phys/triton/levee.py:62-304 with the partition :18-59 (the
             floodplain scan for the storages at levee base and crown, cases
             2-4, river / flood / protected split)
  reservoir  cmfgpu/phys/triton/outflow.py:251-272 (kinematic flow upstream of
             dams), cmfgpu/phys/triton/reservoir.py:17-184 (release by storage
             regime from the inflow the previous sub-step gathered) and the
             reservoir inflow sum outflow.py:380-383
Branch picks follow the forward kernels (max/min pick the first operand only
when it is strictly larger/smaller), and every expression that is not taken
is evaluated on a harmless stand-in so that its gradient is zero, not NaN.
"""
from __future__ import annotations

import torch


def _pick(mask, a, b):
    return torch.where(mask, a, b)


def _safe(mask, x, fill=1.0):
    return torch.where(mask, x, torch.full_like(x, fill))


def _sqrt_floor(x, floor):
    """max(sqrt(x), floor), forward pick sqrt when sqrt(x) > floor."""
    with torch.no_grad():
        take = torch.sqrt(torch.clamp_min(x, 0.0)) > floor
    return _pick(take, torch.sqrt(_safe(take, x)), torch.full_like(x, floor))


def _maxz(x):
    """max(x, 0), picking x only when x > 0."""
    return _pick(x > 0, x, torch.zeros_like(x))


def _sqrt0(x):
    """sqrt(x) for x >= 0 with a zero (not infinite) gradient at x = 0."""
    pos = x > 0
    return _pick(pos, torch.sqrt(_safe(pos, x)), torch.zeros_like(x))


def _levee(prm, lidx, rs, fs, fdep_c, rdep_c, ffr_c):
    """Levee stage of the levee catchments `lidx` from their flood-stage state:
    river, flood and protected storage, river / flood / protected depth and
    flood fraction.  Every case is formed and the forward's pick selects."""
    rl, rw, rh = prm["river_length"][lidx], prm["river_width"][lidx], prm["river_height"][lidx]
    ca = prm["catchment_area"][lidx]
    crown, lfrac, base = prm["levee_crown_height"], prm["levee_fraction"], prm["levee_base_height"]
    table = prm["flood_depth_table"][lidx]
    L = table.shape[1]
    total = rs + fs
    rms = rl * rw * rh
    dwth_inc = (ca / rl) / L
    ld = lfrac * (ca / rl)
    with torch.no_grad():
        ilev = (lfrac * L).long()
    s_curr = rms
    dhgt_pre = torch.zeros_like(rms)
    dwth_pre = rw
    base_sto, fill_sto = rms, rms
    found_base = torch.zeros_like(rms, dtype=torch.bool)
    found_fill = torch.zeros_like(found_base)
    found_b = torch.zeros_like(found_base)
    dsto_fil_b = torch.zeros_like(rms)
    dwth_fil_b = torch.zeros_like(rms)
    ddph_fil_b = torch.zeros_like(rms)
    gradient_b = torch.zeros_like(rms)
    for k in range(L):
        depth = table[:, k]
        seg = _pick(depth - dhgt_pre > 1e-6, depth - dhgt_pre, torch.full_like(depth, 1e-6))
        s_next = s_curr + rl * (dwth_pre + 0.5 * dwth_inc) * seg
        cond_base = (base > dhgt_pre) & (base <= depth)
        rb = (base - dhgt_pre) / seg
        base_sto = _pick(cond_base, s_curr + rl * (dwth_pre + 0.5 * rb * dwth_inc) * (rb * seg), base_sto)
        found_base = found_base | cond_base
        cond_fill = (crown > dhgt_pre) & (crown <= depth)
        rf = (crown - dhgt_pre) / seg
        fill_sto = _pick(cond_fill, s_curr + rl * (dwth_pre + 0.5 * rf * dwth_inc) * (rf * seg), fill_sto)
        found_fill = found_fill | cond_fill
        s_top_loop = base_sto + (ld + rw) * (crown - base) * rl
        threshold = s_next + (ld + rw) * (crown - depth) * rl
        check = (ilev <= k) & ~found_b
        found_now = check & (total < threshold)
        lower = _pick(ilev == k, s_top_loop, dsto_fil_b)
        advance = check & ~found_now
        dsto_fil_b = _pick(advance, threshold, lower)
        dwth_fil_b = _pick(advance, dwth_inc * (k + 1) - ld, dwth_fil_b)
        ddph_fil_b = _pick(advance, depth - base, ddph_fil_b)
        gradient_b = _pick(found_now, seg / dwth_inc, gradient_b)
        found_b = found_b | found_now
        s_curr = s_next
        dhgt_pre = depth
        dwth_pre = dwth_pre + dwth_inc
    base_sto = _pick(found_base, base_sto,
                     _pick(base > dhgt_pre, s_curr + rl * dwth_pre * (base - dhgt_pre), rms))
    fill_sto = _pick(found_fill, fill_sto,
                     _pick(crown > dhgt_pre, s_curr + rl * dwth_pre * (crown - dhgt_pre), rms))
    s_top = base_sto + (ld + rw) * (crown - base) * rl
    case4 = total >= fill_sto
    case3 = ~case4 & (total >= s_top)
    case2 = ~case4 & ~case3 & (total >= base_sto)
    none = ~(case2 | case3 | case4)
    zero = torch.zeros_like(rms)
    # case 2
    f_dph_c2 = base + (total - base_sto) / (ld + rw) / rl
    r_sto_c2 = rms + rl * rw * f_dph_c2
    f_sto_c2 = _maxz(total - r_sto_c2)
    # case 3
    add_b = total - dsto_fil_b
    term_b = dwth_fil_b * dwth_fil_b + 2.0 * add_b / rl / (gradient_b + 1e-9)
    p_dph_found = base + ddph_fil_b + (-dwth_fil_b + _sqrt0(term_b)) * gradient_b
    p_dph_extra = base + ddph_fil_b + add_b / (dwth_fil_b * rl + 1e-9)
    p_dph_b = _pick(found_b, p_dph_found, p_dph_extra)
    f_frc_b = _pick(found_b, (dwth_fil_b + ld) / (dwth_inc * L), torch.ones_like(rms))
    r_sto_c3 = rms + rl * rw * crown
    f_sto_c3 = _maxz(s_top - r_sto_c3)
    # case 4
    f_sto_c4 = _maxz(s_top + (fdep_c - crown) * (ld + rw) * rl - rs)
    r_cand = _pick(case2, r_sto_c2, _pick(case3, r_sto_c3, rs))
    f_cand = _pick(case2, f_sto_c2, _pick(case3, f_sto_c3, _pick(case4, f_sto_c4, fs)))
    # partition (levee.py:18-59)
    r_hp = _pick(case4 | none, rs, _pick(r_cand < total, r_cand, total))
    remaining = _maxz(total - r_hp)
    f_cl = _maxz(f_cand)
    f_hp = _pick(case2, remaining, _pick(none, fs, _pick(f_cl < remaining, f_cl, remaining)))
    p_hp = _maxz(remaining - f_hp)
    r_dph = _pick(case2, r_sto_c2 / rl / rw, _pick(case3, r_sto_c3 / rl / rw, rdep_c))
    f_dph = _pick(case2, f_dph_c2, _pick(case3, crown.expand_as(rms), fdep_c))
    p_dph = _pick(case3, p_dph_b, _pick(case4, fdep_c, zero))
    with torch.no_grad():
        f_frc = _pick(case2, lfrac.expand_as(rms), _pick(case3, torch.clamp(f_frc_b, 0.0, 1.0), ffr_c))
    return r_hp, f_hp, p_hp, r_dph, f_dph, p_dph, f_frc


def _cbrt_sq(x):
    """(x * x)^(1/3) with a zero (not infinite) gradient at x = 0."""
    pos = x > 0
    xs = _safe(pos, x)
    return _pick(pos, torch.pow(xs * xs, 1.0 / 3.0), torch.zeros_like(x))


def _kinematic(prm, dn, c_elv, dist, r_man, f_man, r_wid, r_len, r_dep, f_dep, rs, fs, dt):
    """Kinematic river and floodplain flow of the catchments upstream of a
    dam (outflow.py:251-272)."""
    mk = prm.get("min_kinematic_slope", 1e-5)
    bed = (c_elv - c_elv[dn]) / dist
    bed = _pick(bed > mk, bed, torch.full_like(bed, mk))
    kin_riv = r_wid * r_dep * ((1.0 / r_man) * _sqrt0(bed) * _cbrt_sq(r_dep))
    kin_riv = _maxz(_pick(kin_riv < rs / dt, kin_riv, rs / dt))
    bed_f = _pick(bed < 0.005, bed, torch.full_like(bed, 0.005))
    area = _maxz(fs / r_len - f_dep * r_wid)
    kin_fld = area * ((1.0 / f_man) * _sqrt0(bed_f) * _cbrt_sq(f_dep))
    kin_fld = _maxz(_pick(kin_fld < fs / dt, kin_fld, fs / dt))
    return kin_riv, kin_fld


def _release(prm, ridx, rs, fs, inflow_state, runoff, dt):
    """Reservoir release by storage regime (reservoir.py:114-163): the
    conservation, adjustment and flood-control zones and the emergency
    spill, clamped to [0, S / dt]."""
    total = rs[ridx] + fs[ridx]
    inflow = inflow_state[ridx] + runoff[ridx]
    cv, ev, av = prm["conservation_volume"], prm["emergency_volume"], prm["adjustment_volume"]
    qn, qa, qf = prm["effective_normal_outflow"], prm["adjustment_outflow"], prm["flood_control_outflow"]
    r1 = total <= cv
    r2 = (total > cv) & (total <= av)
    r3 = (total > av) & (total <= ev)
    out1 = qn * _sqrt0(_safe(r1, total / cv))
    f2 = _safe(r2 & (total > cv), (total - cv) / (av - cv))
    out2 = qn + f2 ** 3 * (qa - qn)
    f3 = _safe(r3 & (total > av), (total - av) / (ev - av))
    tmp = qa + torch.pow(f3, 0.1) * (qf - qa)
    o_flood = qn + ((total - cv) / (ev - cv)) * (inflow - qn)
    out3 = _pick(inflow >= qf, _pick(o_flood > tmp, o_flood, tmp), tmp)
    out4 = _pick(inflow >= qf, inflow, qf)
    out = _pick(r1, out1, _pick(r2, out2, _pick(r3, out3, out4)))
    out = _maxz(out)
    cap = total / dt
    return _pick(out < cap, out, cap)


def _paths(st, prm, wse, tot, dt, g, pwse=None):
    """Per-level path flows limited at 5 % of the smaller end storage, and the
    outgoing terms at the upstream (pos) and downstream (neg) ends."""
    ci, di = prm["bifurcation_catchment_idx"], prm["bifurcation_downstream_idx"]
    bq, bcsd = st["bifurcation_outflow"], st["bifurcation_cross_section_depth"]
    bel, bw, bman = prm["bifurcation_elevation"], prm["bifurcation_width"], prm["bifurcation_manning"]
    w0, w1 = wse[ci], wse[di]
    max_w = _pick(w0 > w1, w0, w1)
    sx = (w0 - w1) / prm["bifurcation_length"]
    inside = (sx > -0.005) & (sx < 0.005)
    slope = _pick(inside, sx, torch.clamp(sx.detach(), -0.005, 0.005))
    upd = _maxz(max_w[:, None] - bel)
    a = _sqrt0(upd * bcsd)
    if pwse is None:
        b = _sqrt0(upd * 0.01)
        sifd = _pick(a > b, a, b)
    else:
        # levee variant: level 0 semi-implicit on the river surface (the
        # explicit depth where the product is not positive), upper levels
        # explicit on the protected surface
        p0, p1 = pwse[ci], pwse[di]
        max_pw = _pick(p0 > p1, p0, p1)
        upd = torch.cat([upd[:, :1], _maxz(max_pw[:, None] - bel[:, 1:])], dim=1)
        a = _sqrt0(upd * bcsd)
        semi0 = _pick(a[:, :1] <= 0, upd[:, :1], a[:, :1])
        sifd = torch.cat([semi0, upd[:, 1:]], dim=1)
    cond = sifd > 1e-5
    s_s = _safe(cond, sifd)
    unit = bq / bw
    num = bw * (unit + g * dt * s_s * slope[:, None])
    den = 1.0 + g * dt * (bman * bman) * torch.abs(unit) * (1.0 / (s_s * s_s * torch.pow(s_s, 1.0 / 3.0)))
    q = _pick(cond, num / den, torch.zeros_like(num))
    total = q.sum(dim=1)
    t0, t1 = tot[ci], tot[di]
    m = _pick(t0 < t1, t0, t1)
    nz = total != 0
    ratio = 0.05 * m / (torch.abs(_safe(nz, total)) * dt)
    limit = _pick(nz & (ratio < 1.0), ratio, torch.ones_like(ratio))
    qlim = q * limit[:, None]
    total_l = total * limit
    pos = _maxz(total_l) * dt
    neg = -_pick(total_l < 0, total_l, torch.zeros_like(total_l)) * dt
    return qlim, upd, pos, neg


def substep(st: dict, prm: dict, dt: float, gravity: float = 9.8) -> dict:
    """One sub-step.  st: river_outflow, flood_outflow, river_cross_section_depth,
    flood_cross_section_depth, flood_cross_section_area, river_storage,
    flood_storage, river_depth, flood_depth.  prm: river_manning, river_height
    (may require grad), flood_manning, river_width, river_length,
    catchment_elevation, catchment_area, downstream_distance, runoff,
    flood_depth_table [N, L], downstream_idx (int64).  With paths, prm also
    holds bifurcation_catchment_idx / bifurcation_downstream_idx (int64),
    bifurcation_length [P] and bifurcation_elevation / _width / _manning
    [P, LB], and st bifurcation_outflow / bifurcation_cross_section_depth.
    With levees, prm holds levee_catchment_idx (int64) and
    levee_crown_height / levee_fraction / levee_base_height [V], and st
    protected_storage and protected_depth."""
    dn = prm["downstream_idx"]
    n = dn.shape[0]
    idx = torch.arange(n, device=dn.device)
    mouth = dn == idx
    g = gravity
    r_out, f_out = st["river_outflow"], st["flood_outflow"]
    r_csd, f_csd, f_csa = (st["river_cross_section_depth"], st["flood_cross_section_depth"],
                           st["flood_cross_section_area"])
    rs, fs = st["river_storage"], st["flood_storage"]
    r_dep, f_dep = st["river_depth"], st["flood_depth"]
    r_man, f_man = prm["river_manning"], prm["flood_manning"]
    r_hgt, r_wid, r_len = prm["river_height"], prm["river_width"], prm["river_length"]
    c_elv, dist = prm["catchment_elevation"], prm["downstream_distance"]

    # ---------------------------------------------------------- outflow --
    river_elevation = c_elv - r_hgt
    wse = r_dep + river_elevation
    levee = "levee_catchment_idx" in prm
    ps = st["protected_storage"] if levee else torch.zeros_like(rs)
    tot = rs + fs + ps
    wse_dn = _pick(mouth, c_elv, r_dep[dn] + (c_elv[dn] - r_hgt[dn]))
    max_wse = _pick(wse > wse_dn, wse, wse_dn)
    slope = (wse - wse_dn) / dist
    inside = (slope > -0.005) & (slope < 0.005)
    fslope = _pick(inside, slope, torch.clamp(slope.detach(), -0.005, 0.005))
    upd_rcsd = _pick(mouth, r_dep, max_wse - river_elevation)
    r_sifd = _sqrt_floor(upd_rcsd * r_csd, 1e-6)
    fsurf = _pick(mouth, wse, max_wse)
    upd_fcsd = _maxz(fsurf - c_elv)
    f_sifd = _sqrt_floor(upd_fcsd * f_csd, 1e-6)
    upd_fcsa = _maxz(fs / r_len - f_dep * r_wid)

    river_cond = (r_sifd > 1e-5) & (upd_rcsd * r_wid > 1e-5)
    unit = r_out / r_wid
    rs_s = _safe(river_cond, r_sifd)
    num_r = r_wid * (unit + g * dt * rs_s * slope)
    den_r = 1.0 + g * dt * (r_man * r_man) * torch.abs(unit) * (1.0 / (rs_s * rs_s * torch.pow(rs_s, 1.0 / 3.0)))
    q_r = _pick(river_cond, num_r / den_r, torch.zeros_like(num_r))

    flood_cond = (f_sifd > 1e-5) & (upd_fcsa > 1e-5)
    with torch.no_grad():
        fcsa_take = f_csa > 1e-6
    f_csa_m = _pick(fcsa_take, f_csa, torch.full_like(f_csa, 1e-6))
    f_imp = _sqrt_floor(_safe(flood_cond, upd_fcsa) * f_csa_m, 1e-6)
    fs_s = _safe(flood_cond, f_sifd)
    num_f = f_out + g * dt * f_imp * fslope
    den_f = 1.0 + g * dt * (f_man * f_man) * torch.abs(f_out) * (1.0 / (fs_s * torch.pow(fs_s, 1.0 / 3.0))) / f_imp
    q_f = _pick(flood_cond, num_f / den_f, torch.zeros_like(num_f))

    q_f = _pick(q_r * q_f < 0, torch.zeros_like(q_f), q_f)
    negative = (q_r < 0) & ~mouth
    total_neg = _safe(negative, (-q_r - q_f) * dt)
    ratio = 0.05 * tot / total_neg
    limit = _pick(negative & (ratio < 1.0), ratio, torch.ones_like(ratio))
    q_r = q_r * limit
    q_f = q_f * limit
    reservoirs = "reservoir_catchment_idx" in prm
    if reservoirs:
        du = prm["is_dam_upstream"]
        kr, kf = _kinematic(prm, dn, c_elv, dist, r_man, f_man, r_wid, r_len, r_dep, f_dep, rs, fs, dt)
        q_r = _pick(du, kr, q_r)
        q_f = _pick(du, kf, q_f)
        ridx = prm["reservoir_catchment_idx"]
        out = _release(prm, ridx, rs, fs, st["reservoir_total_inflow"], prm["runoff"], dt)
        q_r = q_r.index_put((ridx,), out)
        q_f = q_f.index_put((ridx,), torch.zeros_like(out))

    paths = "bifurcation_catchment_idx" in prm
    if paths:
        pwse = None
        if levee:
            lidx = prm["levee_catchment_idx"]
            pw_l = c_elv[lidx] + st["protected_depth"][lidx]
            pwse = wse.index_put((lidx,), _pick(pw_l < wse[lidx], pw_l, wse[lidx]))
        qlim, b_upd, b_pos, b_neg = _paths(st, prm, wse, tot, dt, g, pwse)
        ci, di = prm["bifurcation_catchment_idx"], prm["bifurcation_downstream_idx"]

    # ------------------------------------------------- limiter and inflow --
    pos = (_maxz(q_r) + _maxz(q_f)) * dt
    neg_part = _pick(q_r < 0, q_r, torch.zeros_like(q_r)) + _pick(q_f < 0, q_f, torch.zeros_like(q_f))
    neg = _pick(mouth, torch.zeros_like(q_r), -neg_part * dt)
    nm = ~mouth
    og = pos.index_add(0, dn[nm], neg[nm])
    if paths:
        og = og.index_add(0, ci, b_pos).index_add(0, di, b_neg)
    rate = rs + fs
    og_ok = og > 1e-8
    ratio_l = rate / _safe(og_ok, og)
    lim = _pick(og_ok & (ratio_l < 1.0), ratio_l, torch.ones_like(ratio_l))
    l_dn = lim[dn]
    rout = q_r * _pick(q_r >= 0, lim, l_dn)
    fout = q_f * _pick(q_f >= 0, lim, l_dn)
    zero = torch.zeros_like(rout)
    rin = zero.index_add(0, dn[nm], rout[nm])
    fin = zero.index_add(0, dn[nm], fout[nm])
    gbif = zero
    if paths:
        l0, l1 = lim[ci], lim[di]
        raw = qlim.sum(dim=1)
        net = raw * _pick(raw >= 0, l0, l1)
        gbif = zero.index_add(0, ci, net).index_add(0, di, -net)
        b_next = qlim * _pick(qlim >= 0, l0[:, None], l1[:, None])

    # ---------------------------------------------------- storage update --
    sa = rs + (rin - rout) * dt
    carry = sa < 0
    sb = fs + _pick(carry, sa, zero) + (fin - fout - gbif) * dt
    rsu1 = _maxz(sa)
    rsu2 = _pick(sb < 0, _maxz(rsu1 + sb), rsu1)
    fsu1 = _maxz(sb)
    total_next = rsu2 + fsu1 + ps + prm["runoff"] * dt
    total = _maxz(total_next)

    # ------------------------------------------------------------ stage --
    table = prm["flood_depth_table"]
    L = table.shape[1]
    rms = r_len * r_wid * r_hgt
    ca = prm["catchment_area"]
    winc = (ca / r_len) / L
    ks = torch.arange(1, L + 1, dtype=total.dtype, device=total.device)
    W = r_wid[:, None] + ks[None, :] * winc[:, None]
    prev_W = torch.cat([r_wid[:, None], W[:, :-1]], dim=1)
    prev_H = torch.cat([torch.zeros_like(table[:, :1]), table[:, :-1]], dim=1)
    dS = r_len[:, None] * 0.5 * (prev_W + W) * (table - prev_H)
    S = rms[:, None] + torch.cumsum(dS, dim=1)
    wet = total > rms
    with torch.no_grad():
        level = (total[:, None] > S).long().cumprod(dim=1).sum(dim=1)
    S_ext = torch.cat([rms[:, None], S], dim=1)
    H_ext = torch.cat([torch.zeros_like(table[:, :1]), table], dim=1)
    pts = S_ext.gather(1, level[:, None])[:, 0]
    pfd = H_ext.gather(1, level[:, None])[:, 0]
    nfd = H_ext.gather(1, torch.clamp(level + 1, max=L)[:, None])[:, 0]
    ptw = r_wid + level.to(total.dtype) * winc
    top = level == L
    fd_top = pfd + (total - pts) / (ptw * r_len)
    mid = wet & ~top
    fg = _safe(mid, (nfd - pfd) / winc)
    A = ptw * ptw + 2.0 * (total - pts) / (fg * r_len)
    dw = torch.sqrt(_safe(mid, A)) - ptw
    fd_mid = pfd + dw * fg
    fdep = _pick(wet, _pick(top, fd_top, fd_mid), zero)
    cap = rms + r_len * r_wid * fdep
    rsf = _pick(wet & (cap < total), cap, total)
    rdep = rsf / (r_len * r_wid)
    fsf = _maxz(total - rsf)
    if levee:
        lidx = prm["levee_catchment_idx"]
        r_hp, f_hp, p_hp, r_dph, f_dph, p_dph, _ = _levee(prm, lidx, rsf[lidx], fsf[lidx], fdep[lidx], rdep[lidx],
                                                          zero[lidx])
        pdep = fdep.index_put((lidx,), p_dph)
        rsf = rsf.index_put((lidx,), r_hp)
        fsf = fsf.index_put((lidx,), f_hp)
        rdep = rdep.index_put((lidx,), r_dph)
        fdep = fdep.index_put((lidx,), f_dph)
        psf = zero.index_put((lidx,), p_hp)
    out = {"river_outflow": rout, "flood_outflow": fout,
           "river_cross_section_depth": upd_rcsd, "flood_cross_section_depth": upd_fcsd,
           "flood_cross_section_area": upd_fcsa, "river_storage": rsf, "flood_storage": fsf,
           "river_depth": rdep, "flood_depth": fdep}
    if paths:
        out["bifurcation_outflow"] = b_next
        out["bifurcation_cross_section_depth"] = b_upd
    if levee:
        out["protected_storage"] = psf
        out["protected_depth"] = pdep
    if reservoirs:
        # the limited inflow this sub-step gathers at each reservoir
        out["reservoir_total_inflow"] = _pick(prm["is_reservoir"], rin + fin, zero)
    return out
