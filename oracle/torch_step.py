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
             :106-151 (level flows and the path net limited by the ends)
Branch picks follow the forward kernels (max/min pick the first operand only
when it is strictly larger/smaller), and every expression that is not taken
is evaluated on a harmless stand-in so that its gradient is zero, not NaN.
No levee or reservoir.
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


def _paths(st, prm, wse, tot, dt, g):
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
    b = _sqrt0(upd * 0.01)
    sifd = _pick(a > b, a, b)
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
    [P, LB], and st bifurcation_outflow / bifurcation_cross_section_depth."""
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
    tot = rs + fs
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

    paths = "bifurcation_catchment_idx" in prm
    if paths:
        qlim, b_upd, b_pos, b_neg = _paths(st, prm, wse, tot, dt, g)
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
    total_next = rsu2 + fsu1 + prm["runoff"] * dt
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
    out = {"river_outflow": rout, "flood_outflow": fout,
           "river_cross_section_depth": upd_rcsd, "flood_cross_section_depth": upd_fcsd,
           "flood_cross_section_area": upd_fcsa, "river_storage": rsf, "flood_storage": fsf,
           "river_depth": rdep, "flood_depth": fdep}
    if paths:
        out["bifurcation_outflow"] = b_next
        out["bifurcation_cross_section_depth"] = b_upd
    return out
