#!/usr/bin/env python3
"""Benchmark of the HIP hydrodynamic hot path on one MI355X.

Workload (default): the 6-arcmin-scale synthetic network — N = 1,600,000
catchments, levees on 5 %, 40,000 bifurcation pathways (5 levels), 10 flood
levels — run with the plain method (one trial), adaptive CFL sub-steps under a
1800 s outer step, FP32 fluxes and FP64 storage.  One bench "step" is one
outer step; `value` counts catchment-sub-steps per second (every CFL sub-step
of every catchment).  Runoff is synthetic (lognormal, mean 1 mm/day) and
resident in HBM before timing starts.

The same JSON line carries:
  * batched: the 64-member run of the same network (independent runoff per
    member, each member with its own CFL sub-step), member-catchment-sub-steps/s,
    and its ratio to the plain rate (64 sequential plain runs);
  * roofline: the dominant kernel's algorithmic bytes per launch over its mean
    duration, timed with HIP events on the launch stream;
  * cpu_baseline: the CPU oracle (C++/OpenMP restatement of the reference
    step) on the host cores, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cama-flood-gpu_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "catchment-steps/s on 1 MI355X, plain vs 64-member batched; HBM % of peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic bytes per element of each fused kernel (f32 fluxes, f64
# storage), counting each field the kernel must read or write once; gathers
# from the downstream catchment are taken as cache hits.
def algorithmic_bytes(kernel: str, model, f_wet: float) -> float:
    N, T, P, V = model.N, model.T, model.P, model.V
    lev = V > 0
    bif = P > 0
    # the path pass that shares the outflow launch: ends 8, length 4, five
    # levels of manning/width/elevation/depth/flow read and depth/flow written,
    # path terms 12 (end-catchment state gathers taken as cache hits)
    paths = (8 + 4 + 5 * (4 * 3 + 8 + 8) + 12) * P * T
    # (the timed sub-step is not the last of its outer step: the water surfaces,
    # total storage, flood fraction and protected depth, written on the last
    # one only, are not counted)
    if kernel.startswith("k_outflow"):
        # writes: two flows, three cross-section fields and either the two
        # outgoing terms or, with provisional limiters (_prov), the limiter
        base = kernel.split("<")[0]
        prov = base.endswith("_prov")
        per = 84 + (24 if prov else 28) + (5 if lev else 0)
        # with _prov the path pass also writes its two terms at the ends' event slots
        return per * N * T + ((paths + (8 * P * T if prov else 0)) if base.startswith("k_outflow_paths") else 0)
    if kernel.startswith("k_limiter_fix"):
        # the flag bits (1/8 B per element), and per path end and trial its
        # record (16 B), storages (16 B), unlimited flows (8 B) and limiter
        # (4 B), the flows of its upstream catchments (8 B each) and per event
        # the path term (4 B); backflow catchments are not counted
        if not bif:
            return N * T / 8.0
        ends, ups = path_end_counts(model)
        return N * T / 8.0 + T * (ends * 44.0 + 8.0 * ups + 4.0 * 2 * P)
    if kernel.startswith("k_limiter"):
        per = 4 + 4 + 4 + 16 + 4 + (8 if bif else 0)
        return per * N * T
    if kernel.startswith("k_stage"):
        per = 12 + 12 + 4 + 24 + 4 + 12 + 40 + 44 * f_wet
        # path nets: events 8 + other end 4 + level sum 4 + limiter 4 at each
        # end, the five limited level flows read and written once
        return per * N * T + ((2 * (8 + 4 + 4 + 4) + 5 * 8) * P * T if bif else 0)
    if kernel.startswith("k_levee"):
        return (4 + 12 + 16 + 16 + 40 + 4 + 4 + 4 + 24 + 16) * V * T
    if kernel.startswith("k_bif_outflow"):
        return (8 + 4 + 8 + 16 + 5 * (4 * 3 + 8 + 8) + 8) * P * T
    if kernel.startswith("k_bif_inflow"):
        return (8 + 8 + 5 * 8 + 4) * P * T
    return 0.0


def path_end_counts(model):
    """(catchments that end a bifurcation path, their upstream catchments)."""
    import numpy as np
    hf = model.host.fields
    ci = np.asarray(hf["bifurcation_catchment_idx"])
    di = np.asarray(hf["bifurcation_downstream_idx"])
    ends = np.unique(np.concatenate([ci, di]))
    dn = np.asarray(hf["downstream_idx"])
    ups = np.bincount(dn[dn != np.arange(dn.size)], minlength=dn.size)[ends].sum()
    return int(ends.size), int(ups)


# schedule of the fused sub-step for every model the bench builds (--schedule)
SCHEDULE = "three_pass"


def scheduled(m):
    """The model with the bench's sub-step schedule applied."""
    if SCHEDULE != "three_pass":
        m.set_schedule(SCHEDULE)
    return m


def build_model(net, src, members: int, device: str, layout: str, precise: bool = False,
                precision: str = "float32"):
    return scheduled(_build_model(net, src, members, device, layout, precise, precision))


def _build_model(net, src, members: int, device: str, layout: str, precise: bool, precision: str):
    from cmfgpu_amd.models import CaMaFlood
    mods = ["base", "adaptive_time"]
    if net.num_paths:
        mods.append("bifurcation")
    if net.num_levees:
        mods.append("levee")
    kw = dict(opened_modules=tuple(mods), device=device, step_duration=1800.0, layout=layout,
              precise_surfaces=precise, precision=precision)
    if members <= 1:
        return CaMaFlood(src, **kw)
    import numpy as np
    from cmfgpu_amd import synthetic
    runoff = np.stack([synthetic.Runoff(src["catchment_area"], 4000 + m).next_day() for m in range(members)])
    return CaMaFlood(src, num_trials=members, batched={"runoff": runoff}, per_trial_time_step=True,
                     trial_layout="inner", **kw)


ERROR_FIELDS = ("river_outflow", "flood_outflow", "river_storage", "flood_storage", "river_depth",
                "flood_depth", "flood_fraction")


def fp64_errors(net, src, device: str, layout: str, days: int) -> dict:
    """Relative L2 error of each output field of the FP32 (FP64 storage) run
    against the (FP64, FP64) run of the same kernels, after `days` simulated
    days (48 outer steps each, fresh daily runoff) from the same state.  The
    FP64 profile is bit-identical to the FP64 CPU oracle (parity tests), so
    this is the error against an FP64 oracle run at full size.

    A third run, FP32 with water surfaces formed in double (precise_surfaces),
    gives the same errors under "precise_surfaces", and then that run and the
    FP32 run are timed from their final states (alternating blocks of 12
    outer steps, same launch shapes) for the cost of the flag."""
    import time
    import numpy as np
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.models import CaMaFlood
    mods = ["base", "adaptive_time"] + (["bifurcation"] if net.num_paths else []) \
        + (["levee"] if net.num_levees else [])
    models = [CaMaFlood(src, opened_modules=tuple(mods), device=device, step_duration=1800.0, layout=layout,
                        precision=p, precise_surfaces=x)
              for p, x in (("float32", False), ("float64", False), ("float32", True))]
    forcing = synthetic.Runoff(src["catchment_area"], 2003)
    count_mismatch = [0, 0]
    for _ in range(days):
        r = forcing.next_day()
        for m in models:
            m.set_inputs(r)
        for _ in range(48):
            a, b, c = (m.step_advance() for m in models)
            count_mismatch[0] += int(a != b)
            count_mismatch[1] += int(c != b)
    res = []
    for k in (0, 2):
        out = {}
        for name in ERROR_FIELDS:
            x, y = (models[j].get(name).astype(np.float64) for j in (k, 1))
            out[name] = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
        q, q64 = ((models[j].get("river_outflow").astype(np.float64) + models[j].get("flood_outflow"))
                  for j in (k, 1))
        res.append({"days": days, "outer_steps": 48 * days, "substep_count_mismatches": count_mismatch[k // 2],
                    "discharge_rel_l2": float(np.linalg.norm(q - q64) / np.linalg.norm(q64)),
                    "fields_rel_l2": {k2: float(f"{v:.4g}") for k2, v in out.items()}})
    models[1].close()
    rate = {0: [0, 0.0], 2: [0, 0.0]}
    for _ in range(4):
        for k in (0, 2):
            m = models[k]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(12):
                rate[k][0] += m.step_advance() * m.N
            torch.cuda.synchronize()
            rate[k][1] += time.perf_counter() - t0
    for k in (0, 2):
        models[k].close()
    out = res[0]
    out["precise_surfaces"] = dict(res[1], catchment_steps_per_s=rate[2][0] / rate[2][1],
                                   vs_reference_surfaces=round((rate[2][0] / rate[2][1]) / (rate[0][0] / rate[0][1]), 4))
    return out


def _rel_rows(x, y):
    """Relative L2 error of each member row of x against y."""
    import numpy as np
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return [float(f"{np.linalg.norm(x[t] - y[t]) / max(np.linalg.norm(y[t]), 1e-300):.4g}") for t in range(x.shape[0])]


def batched_line(args, net, src, device: str, plain_value: float, plain_ref_value: float, tuned) -> dict:
    """64 members of the same network (runoff seeds 4000 + m), each with its
    own CFL count, trial-inner, with precise surfaces (the headline profile)
    and with the reference's FP32 surfaces: one simulated day of warm-up
    (48 outer steps), then `--batched-steps` timed outer steps each.  Then the
    per-member error of one simulated day against the FP64 profile of the same
    batch, from the same initial state, for both surface modes."""
    import numpy as np
    import torch
    M = args.members
    out = {"members": M, "unit": "member-catchment-steps/s"}
    bsteps = args.batched_steps
    for precise in (True, False):
        bm = build_model(net, src, M, device, args.layout, precise=precise)
        if args.autotune:
            bm.autotune(use_cache=False)
        run_steps(bm, 48, per_trial=True)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        bwork = run_steps(bm, bsteps, per_trial=True)
        torch.cuda.synchronize()
        belapsed = time.perf_counter() - t2
        bvalue = bwork / belapsed
        bf_wet = float((bm.t["flood_depth"] > 0).float().mean().item())
        prov = path_end_counts(bm) if SCHEDULE == "three_pass" else None
        bgbs = batched_full_bytes(bm.N, M, bm.P, bf_wet, prov) * (bwork / (bm.N * M)) / belapsed / 1e9
        cgbs = fused_batched_bytes(bm.N, M, bm.V, bm.P, bf_wet) * (bwork / (bm.N * M)) / belapsed / 1e9
        base = plain_value if precise else plain_ref_value
        rec = {"value": bvalue, "steps": bsteps, "warmup": "one simulated day (48 outer steps)",
               "vs_plain": round(bvalue / base, 3), "GBps_algorithmic": round(bgbs, 1),
               "frac_of_peak": round(bgbs / HBM_PEAK_GBS, 4), "GBps_compulsory_fused": round(cgbs, 1),
               "frac_compulsory_fused": round(cgbs / HBM_PEAK_GBS, 4)}
        if precise:
            out.update(rec)
        else:
            out["reference_surfaces"] = rec
        log(f"batched ({'precise' if precise else 'reference'} surfaces): {M} members, {bsteps} outer steps, "
            f"{belapsed:.3f} s -> {bvalue:.4g} member-catchment-steps/s ({bvalue / base:.2f}x plain)")
        bm.close()
        del bm
        torch.cuda.empty_cache()
    if args.error_days > 0:
        models = {"fp64": build_model(net, src, M, device, args.layout, precision="float64"),
                  "precise": build_model(net, src, M, device, args.layout, precise=True),
                  "reference": build_model(net, src, M, device, args.layout, precise=False)}
        mismatch = {"precise": 0, "reference": 0}
        for _ in range(48 * args.error_days):
            n64 = None
            for name, m in models.items():
                m.step_advance()
                c = m.t["max_sub_steps"].cpu().numpy()
                if name == "fp64":
                    n64 = c
                else:
                    mismatch[name] += int((c != n64).sum())
        q = {k: m.get("river_outflow").astype(np.float64) + m.get("flood_outflow") for k, m in models.items()}
        for k in ("precise", "reference"):
            rows = _rel_rows(q[k], q["fp64"])
            rec = out if k == "precise" else out["reference_surfaces"]
            rec["discharge_rel_l2_per_member"] = rows
            rec["max_discharge_rel_l2"] = max(rows)
            rec["member_substep_count_mismatches"] = mismatch[k]
        out["error_days"] = args.error_days
        log(f"batched errors after {args.error_days} day(s): discharge up to {out['max_discharge_rel_l2']:.3g} "
            f"(precise), {out['reference_surfaces']['max_discharge_rel_l2']:.3g} (reference surfaces)")
        for m in models.values():
            m.close()
    return out


def basins_line(args, device: str) -> dict:
    """Uneven basins (log-uniform sizes 500..400,000, plain physics): every
    basin advances with its own CFL sub-step count and dt in one launch per
    pass (groups), eagerly (each sub-step launching only the blocks of basins
    still sub-stepping) and with each outer step's sub-step loop replayed from
    a captured HIP graph, against the same network with one shared sub-step
    count (the reference's semantics).  `value` counts the sub-steps each
    basin really takes."""
    import numpy as np
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.models import CaMaFlood
    t0 = time.perf_counter()
    bs = synthetic.build_basins(args.basins, scale=args.basins_scale)
    src = dict(bs.network.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2002).next_day()
    n = bs.network.num_catchments
    log(f"basins: {args.basins} basins, {n} catchments ({time.perf_counter() - t0:.1f} s)")
    sizes = np.diff(bs.basin_ptr)
    out = {"basins": args.basins, "catchments": int(n)}
    for grouped, graph in ((True, False), (True, True), (False, False)):
        m = scheduled(CaMaFlood(src, opened_modules=("base", "adaptive_time"), device=device, step_duration=1800.0,
                                groups=bs.group if grouped else None, use_graph=graph))
        for _ in range(max(args.warmup // 4, 2)):
            m.step_advance()
        torch.cuda.synchronize()
        steps = max(args.steps // 4, 4)
        work = 0
        t1 = time.perf_counter()
        for _ in range(steps):
            cnt = m.step_advance()
            if grouped:
                work += int((sizes * m.group_sub_steps()).sum())
            else:
                work += n * cnt
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        key = ("per_basin_dt_graph" if graph else "per_basin_dt") if grouped else "shared_dt"
        # whole-run HBM rate: the fused passes' algorithmic bytes per
        # catchment-sub-step times the measured rate (wall clock, CFL pass and
        # launch gaps included)
        f_wet = float((m.t["flood_depth"] > 0).float().mean().item())
        kern = (("k_outflow_prov", "k_limiter_fix", "k_stage_fused") if SCHEDULE == "three_pass"
                else ("k_outflow", "k_limiter_pair", "k_stage_fused"))
        per = sum(algorithmic_bytes(k, m, f_wet) for k in kern) / n
        gbs = work / el * per / 1e9
        cgbs = work / el * fused_substep_bytes(n, 0, 0, f_wet) / n / 1e9
        out[key] = {"outer_steps_per_s": steps / el, "catchment_steps_per_s": work / el,
                    "bytes_per_catchment_step": round(per, 1), "GBps": round(gbs, 1),
                    "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "f_wet": round(f_wet, 4),
                    "GBps_compulsory_fused": round(cgbs, 1), "frac_compulsory_fused": round(cgbs / HBM_PEAK_GBS, 4)}
        m.close()
        del m
        torch.cuda.empty_cache()
    out["value"] = out["per_basin_dt"]["catchment_steps_per_s"]
    out["unit"] = "catchment-steps/s"
    out["speedup_vs_shared_dt"] = round(out["per_basin_dt"]["outer_steps_per_s"] / out["shared_dt"]["outer_steps_per_s"], 3)
    log(f"basins: per-basin dt {out['per_basin_dt']['outer_steps_per_s']:.1f} outer steps/s "
        f"({out['per_basin_dt_graph']['outer_steps_per_s']:.1f} from graphs), shared dt "
        f"{out['shared_dt']['outer_steps_per_s']:.1f} ({out['speedup_vs_shared_dt']}x); HBM "
        f"{out['per_basin_dt']['GBps']:.0f} / {out['shared_dt']['GBps']:.0f} GB/s algorithmic")
    return out


def fused_substep_bytes(N: int, V: int, P: int, f_wet: float, LB: int = 5) -> float:
    """Compulsory HBM traffic of one fully fused sub-step (every field read
    or written once, gathers as cache hits): 128 B per catchment plus 44 B
    per wet catchment (area and floodplain row); levees add the protected
    storage and depth read and written (24 B per catchment) and 16 B of
    parameters per levee; paths add 16 B per catchment (water surface, total
    storage and limiter) and per path LB levels of Manning, width, elevation
    (read) and flow and depth (read and written), its length and two ends."""
    b = N * (128.0 + 44.0 * f_wet)
    if V:
        b += 24.0 * N + 16.0 * V
    if P:
        b += 16.0 * N + P * (LB * (4 + 4 + 4 + 8 + 8) + 4 + 8)
    return b


def fused_batched_bytes(N: int, M: int, V: int, P: int, f_wet: float, LB: int = 5) -> float:
    """The same for M members sharing the static parameters: per
    member-catchment runoff, state read and written and flood fraction (96 B),
    with levees the protected state (24 B), with paths the surface, total
    storage and limiter (16 B); per catchment the eight static fields (32 B)
    and, where any member is wet, area and floodplain row (44 B); per levee
    16 B of parameters; per path-member the level flows and depths read and
    written (LB x 16 B); per path the level parameters, length and ends."""
    b = N * M * (96.0 + (24.0 if V else 0.0) + (16.0 if P else 0.0)) + N * (32.0 + 44.0 * f_wet)
    if V:
        b += 16.0 * V
    if P:
        b += P * M * LB * 16.0 + P * (LB * 12.0 + 12.0)
    return b


def batched_full_bytes(N: int, M: int, P: int, f_wet: float, prov=None) -> float:
    """Algorithmic bytes of one fused sub-step of M trial-inner members with
    full physics, per-member data counted once per member and the shared
    parameters once per catchment (a sub-step after the first: the write-only
    fields are not written):
      per member-catchment 184 B: outflow reads flows, depths, river and
      flood storage, three cross-section fields (44 B) and writes two flows,
      three cross-section fields and two terms (28 B); the limiter reads two
      terms and two storages and writes the limiter (28 B); the stage reads
      own and upstream limiter and flows (24 B), three storages (24 B) and
      runoff (4 B) and writes two flows, two storages and two depths (32 B);
      per catchment 69 B + 44 f_wet: eight channel fields and the levee flag
      (33 B), upstream ranges and event records (16 B), channel geometry,
      levee row and event range (20 B), the table row and area where wet;
      per path-member 112 B: five levels of flow and depth (60 B), three path
      terms (12 B), the nets' rescale of the five flows (40 B);
      per path 72 B: five levels of Manning's n, width and elevation, length
      and the two ends.
    prov = (path ends, their upstream catchments): the provisional-limiter
    schedule, whose outflow writes the limiter (4 B) instead of the two terms
    and which has no limiter pass (152 B per member-catchment); its fix-up
    reads the flag bits (1/8 B per member-catchment), per path end its record
    (16 B) and per path end and member the storages, flows and limiter (28 B)
    and the flows of its upstream catchments (8 B each); the path pass writes
    its two terms at the ends' event slots and the fix-up reads them (16 B
    per path-member)."""
    if prov is None:
        return N * M * 184.0 + N * (69.0 + 44.0 * f_wet) + P * M * 112.0 + P * 72.0
    ends, ups = prov
    return (N * M * (152.0 + 0.125) + N * (69.0 + 44.0 * f_wet) + P * M * 128.0 + P * 72.0
            + ends * 16.0 + M * (ends * 28.0 + ups * 8.0))


def _reaches_below(src: dict, start: np.ndarray) -> np.ndarray:
    """Catchments on the river paths from `start` to their mouths."""
    import numpy as np
    dn = np.asarray(src["downstream_idx"], np.int64)
    below = np.zeros(dn.shape[0], bool)
    front = np.unique(np.asarray(start, np.int64))
    below[front] = True
    while front.size:
        nxt = dn[front]
        nxt = np.unique(nxt[~below[nxt]])
        below[nxt] = True
        front = nxt
    return below


def batched_full_line(args, device: str) -> dict:
    """Batched full physics: the 3-arcmin-scale network (6,000,000
    catchments, 150,000 bifurcation paths, levees on 5 %, 7,000 reservoirs)
    with `--full-members` runoff members (seeds 4000 + m), each with its own
    CFL count, trial-inner.  Member-catchment-steps/s over timed outer steps,
    the whole-run algorithmic HBM rate (batched_full_bytes), and after
    `--full-days` simulated days (0: none) the relative L2 error of each
    member's discharge, storages and flood depth against the FP64 profile of
    the same batch (bit-identical to the FP64 oracle)."""
    import numpy as np
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.models import CaMaFlood
    t0 = time.perf_counter()
    net = synthetic.build_config(4, scale=args.full_scale)
    src = dict(net.fields)
    M = args.full_members
    forcing = [synthetic.Runoff(src["catchment_area"], 4000 + m) for m in range(M)]
    runoff = np.stack([f.next_day() for f in forcing])
    N, P = net.num_catchments, net.num_paths
    log(f"batched full: {N} catchments, {P} paths, {net.num_levees} levees, {net.num_reservoirs} reservoirs, "
        f"{M} members ({time.perf_counter() - t0:.1f} s)")
    mods = ("base", "adaptive_time", "bifurcation", "levee", "reservoir")

    def build(precision, precise=False):
        return scheduled(CaMaFlood(src, opened_modules=mods, device=device, step_duration=1800.0, num_trials=M,
                                   batched={"runoff": runoff}, per_trial_time_step=True, trial_layout="inner",
                                   precision=precision, precise_surfaces=precise))

    out = {"workload": f"{N} catchments, {P} paths, {net.num_levees} levees, {net.num_reservoirs} reservoirs, "
                       f"{M} members with their own CFL counts, trial-inner", "members": M}
    bm = build("float32")
    if args.full_days > 0:
        # FP32 and FP64 profiles from the same state, and FP32 with precise
        # surfaces
        ref = build("float64")
        pr = build("float32", precise=True)
        # the FP64 profile again, from the initial river storage scaled by
        # 1 +- 1e-7 (one FP32 rounding): the scheme's own sensitivity
        pert = build("float64")
        sign = torch.from_numpy(np.random.default_rng(1).choice([-1.0, 1.0], pert.t["river_storage"].numel())).to(
            pert.t["river_storage"].device).reshape(pert.t["river_storage"].shape)
        pert.t["river_storage"].mul_(1.0 + 1e-7 * sign)
        del sign
        for day in range(args.full_days):
            if day:
                r = np.stack([f.next_day() for f in forcing])
                for m in (bm, ref, pr, pert):
                    m.set_inputs(r)
            for _ in range(48):
                bm.step_advance()
                ref.step_advance()
                pr.step_advance()
                pert.step_advance()
        errors = {}
        for name in ("river_storage", "flood_storage", "flood_depth"):
            x, y = bm.get(name).astype(np.float64), ref.get(name).astype(np.float64)
            errors[name] = [float(f"{np.linalg.norm(x[t] - y[t]) / max(np.linalg.norm(y[t]), 1e-300):.4g}")
                            for t in range(M)]
        q32, q64, qpr, qpt = ((m.get("river_outflow").astype(np.float64) + m.get("flood_outflow"))
                              for m in (bm, ref, pr, pert))
        ref.close()
        pr.close()
        pert.close()
        errors["discharge"] = [float(f"{np.linalg.norm(q32[t] - q64[t]) / np.linalg.norm(q64[t]):.4g}")
                               for t in range(M)]
        # the same outside the reaches below reservoirs: a release switches
        # regime at storage and inflow thresholds, and a regime that flips
        # between the two profiles changes the flow of the whole reach below
        reach = _reaches_below(src, src["reservoir_catchment_idx"])
        keep = ~reach
        # and without the catchments coupled to a reach within a sub-step:
        # those draining into it (their slope and backflow follow the reach's
        # water surface) and the other ends of paths touching it
        dn = np.asarray(src["downstream_idx"], np.int64)
        coupled = reach | reach[dn]
        ci, di = np.asarray(src["bifurcation_catchment_idx"]), np.asarray(src["bifurcation_downstream_idx"])
        coupled[ci[reach[di]]] = True
        coupled[di[reach[ci]]] = True
        keep2 = ~coupled
        errors["discharge_outside_reservoir_reaches"] = [
            float(f"{np.linalg.norm((q32[t] - q64[t])[keep]) / np.linalg.norm(q64[t][keep]):.4g}")
            for t in range(M)]
        precise = {"discharge": [float(f"{np.linalg.norm(qpr[t] - q64[t]) / np.linalg.norm(q64[t]):.4g}")
                                 for t in range(M)],
                   "discharge_outside_reservoir_reaches": [
                       float(f"{np.linalg.norm((qpr[t] - q64[t])[keep]) / np.linalg.norm(q64[t][keep]):.4g}")
                       for t in range(M)]}
        precise["discharge_outside_reservoir_coupled"] = [
            float(f"{np.linalg.norm((qpr[t] - q64[t])[keep2]) / np.linalg.norm(q64[t][keep2]):.4g}")
            for t in range(M)]
        precise["max_discharge_rel_l2_outside_reservoir_coupled"] = max(precise["discharge_outside_reservoir_coupled"])
        sens = {"discharge": [float(f"{np.linalg.norm(qpt[t] - q64[t]) / np.linalg.norm(q64[t]):.4g}")
                              for t in range(M)],
                "discharge_outside_reservoir_reaches": [
                    float(f"{np.linalg.norm((qpt[t] - q64[t])[keep]) / np.linalg.norm(q64[t][keep]):.4g}")
                    for t in range(M)]}
        sens["max_discharge_rel_l2"] = max(sens["discharge"])
        precise["max_discharge_rel_l2"] = max(precise["discharge"])
        precise["max_discharge_rel_l2_outside_reservoir_reaches"] = max(precise["discharge_outside_reservoir_reaches"])
        del q32, q64, qpr, qpt
        out.update({"error_days": args.full_days, "rel_l2_per_member": errors, "precise_surfaces": precise,
                    "fp64_perturbed_1e-7": sens, "reservoir_coupled_catchments": int(coupled.sum()),
                    "max_discharge_rel_l2": max(errors["discharge"]),
                    "reservoir_reach_catchments": int((~keep).sum()),
                    "max_discharge_rel_l2_outside_reservoir_reaches":
                        max(errors["discharge_outside_reservoir_reaches"])})
    # throughput
    run_steps(bm, 2, per_trial=True)
    steps = args.full_steps
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    work = run_steps(bm, steps, per_trial=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    subs = work / (N * M)                       # mean sub-steps per member over the timed steps
    f_wet = float((bm.t["flood_depth"] > 0).float().mean().item())
    prov = path_end_counts(bm) if SCHEDULE == "three_pass" else None
    gbs = batched_full_bytes(N, M, P, f_wet, prov) * subs / el / 1e9
    bm.close()
    out.update({"value": work / el, "unit": "member-catchment-steps/s", "steps": steps,
                "GBps_algorithmic": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                "f_wet": round(f_wet, 4)})
    msg = f"batched full: {work / el:.4g} member-catchment-steps/s, {gbs:.0f} GB/s algorithmic"
    if "max_discharge_rel_l2" in out:
        msg += (f"; discharge error after {args.full_days} day(s) up to {out['max_discharge_rel_l2']:.3g}, "
                f"{out['max_discharge_rel_l2_outside_reservoir_reaches']:.3g} outside reservoir reaches; precise "
                f"surfaces {out['precise_surfaces']['max_discharge_rel_l2']:.3g}, "
                f"{out['precise_surfaces']['max_discharge_rel_l2_outside_reservoir_reaches']:.3g} outside reaches, "
                f"{out['precise_surfaces']['max_discharge_rel_l2_outside_reservoir_coupled']:.3g} outside reaches and "
                f"their neighbours; FP64 perturbed by 1e-7: {out['fp64_perturbed_1e-7']['max_discharge_rel_l2']:.3g}")
    log(msg)
    return out


def graph_line(args, device: str) -> dict:
    """Launch-bound case: the 60,000-catchment plain network, sub-step loops
    launched eagerly vs replayed from captured HIP graphs."""
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.models import CaMaFlood
    net = synthetic.build_config(0)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2000).next_day()
    out = {"workload": "60,000-catchment plain network", "catchments": net.num_catchments}
    for use_graph in (False, True):
        m = scheduled(CaMaFlood(src, opened_modules=("base", "adaptive_time"), device=device, step_duration=1800.0,
                                use_graph=use_graph))
        for _ in range(8):
            m.step_advance()
        torch.cuda.synchronize()
        steps = max(args.steps, 24)
        work = 0
        t1 = time.perf_counter()
        for _ in range(steps):
            work += m.step_advance() * net.num_catchments
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        out["graph" if use_graph else "eager"] = {"outer_steps_per_s": steps / el, "catchment_steps_per_s": work / el}
        m.close()
    out["speedup"] = round(out["graph"]["outer_steps_per_s"] / out["eager"]["outer_steps_per_s"], 3)
    log(f"graph: {out['graph']['outer_steps_per_s']:.0f} vs eager {out['eager']['outer_steps_per_s']:.0f} "
        f"outer steps/s ({out['speedup']}x) on {net.num_catchments} catchments")
    return out


def _adjoint_run(args, net, src, device: str, precise: bool) -> dict:
    """One timed forward and backward of the calibration problem (see
    adjoint_line), with the reference's FP32 surfaces or precise ones."""
    import numpy as np
    import torch
    from cmfgpu_amd.adjoint import GaugeProblem
    from cmfgpu_amd.models import CaMaFlood
    model = CaMaFlood(src, opened_modules=("base", "adaptive_time", "bifurcation", "levee"), device=device,
                      step_duration=1800.0, precise_surfaces=precise)
    gauges = net.gauges
    steps = args.adjoint_steps
    K = 48
    rng = np.random.default_rng(2003)
    problem = GaugeProblem(model, gauges, np.ones((steps, gauges.shape[0])), checkpoint_every=K)
    n0, h0 = src["river_manning"].astype(np.float64), src["river_height"].astype(np.float64)
    problem.set_parameters(n0 * rng.lognormal(0.0, 0.2, n0.size), h0 * rng.lognormal(0.0, 0.2, h0.size))
    obs = []
    for _ in range(steps):
        model.step_advance()
        obs.append(problem._gauge_q().cpu().numpy())
    problem.set_parameters(n0, h0)
    problem._restore(problem.initial)
    warm = GaugeProblem(model, gauges, np.array(obs[:min(K, steps)]), checkpoint_every=K)
    warm.run()
    warm.gradient()
    del warm
    problem.obs = torch.as_tensor(np.array(obs), dtype=torch.float64, device=model.device)
    problem._restore(problem.initial)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = problem.run()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_n, g_h = problem.gradient()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    sub = sum(problem.counts)
    out = {"outer_steps": steps, "checkpoint_every": K, "substeps": sub, "loss": loss,
           "forward_ms_per_substep": t_fwd / sub * 1e3, "backward_ms_per_substep": t_bwd / sub * 1e3,
           "ratio": round(t_bwd / t_fwd, 3),
           "gradient_finite": bool(torch.isfinite(g_n).all().item() and torch.isfinite(g_h).all().item()),
           "gradient_nonzero_catchments": int(((g_n != 0) | (g_h != 0)).sum().item())}
    log(f"adjoint{' (precise surfaces)' if precise else ''}: {steps} outer steps ({sub} sub-steps), forward "
        f"{out['forward_ms_per_substep']:.3f} ms/sub-step, backward {out['backward_ms_per_substep']:.3f} "
        f"ms/sub-step ({out['ratio']}x)")
    del problem
    model.close()
    del model
    torch.cuda.empty_cache()
    return out


def adjoint_line(args, device: str) -> dict:
    """The calibration step on the 6-arcmin network with all its modules
    (base physics, 40,000 bifurcation paths, levees on 5 %): 2,000 gauges with
    upstream area > 5,000 km^2, observations from a forward run with Manning's
    n and river height scaled by lognormal(0, 0.2) factors, loss = MSE of log
    discharge over `--adjoint-steps` outer steps (720: 15 days), checkpoints
    every 48.  Forward time per sub-step against backward time (segment
    re-runs + reverse sweep) per sub-step, after one untimed forward and
    backward of one segment (first use of every kernel, and the tape of a
    full segment allocated once by the caching allocator); with the
    reference's FP32 surfaces, then with water surfaces formed in double."""
    from cmfgpu_amd import synthetic
    net = synthetic.build_config(3, scale=args.scale)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day()
    out = {"workload": f"{net.num_catchments} catchments, {net.num_paths} bifurcation paths, {net.num_levees} "
                       f"levees, {net.gauges.shape[0]} gauges, observations from lognormal(0, 0.2)-perturbed n "
                       f"and H"}
    out.update(_adjoint_run(args, net, src, device, precise=False))
    out["precise_surfaces"] = _adjoint_run(args, net, src, device, precise=True)
    return out


def forcing_line(model, reps: int = 20) -> dict:
    """Gridded runoff mapped onto the model's catchments (cmf_hip_map_runoff):
    a 0.1-degree grid (3600 x 1800 cells), 1-4 cells per catchment."""
    import numpy as np
    import torch
    from cmfgpu_amd.forcing import RunoffMapping
    rng = np.random.default_rng(12)
    n, cells = model.N, 3600 * 1800
    counts = rng.integers(1, 5, n)
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum(counts)
    # neighbouring catchments sit on neighbouring cells, as on a real map
    base = (np.sort(rng.integers(0, cells - 8, n))).astype(np.int64)
    indices = (np.repeat(base, counts) + (np.arange(indptr[-1]) - np.repeat(indptr[:-1], counts))).astype(np.int32)
    weights = rng.uniform(1e6, 1e8, indptr[-1]).astype(np.float32)
    rm = RunoffMapping(indptr, indices, weights, cells, 1e-3 / 86400.0).bind(model)
    grid = torch.rand(cells, device=model.device, dtype=torch.float32)
    rm.apply(grid)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        rm.apply(grid)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nnz = int(indptr[-1])
    algo = n * (8 + 4) + nnz * (4 + 4 + 4)
    out = {"catchments": n, "grid_cells": cells, "nonzeros": nnz, "ms": round(ms, 5),
           "algorithmic_bytes": algo, "GBps": round(algo / (ms * 1e-3) / 1e9, 1)}
    log(f"forcing: {nnz} weights onto {n} catchments in {ms * 1e3:.1f} us ({out['GBps']} GB/s)")
    return out


def run_steps(model, steps: int, per_trial: bool):
    """Advance `steps` outer steps; returns member-catchment-sub-steps done."""
    import torch
    work = 0
    for _ in range(steps):
        n = model.step_advance()
        if per_trial:
            counts = model.t["max_sub_steps"].cpu()
            work += int(counts.sum()) * model.N
        else:
            work += n * model.N * model.T
    torch.cuda.synchronize()
    return work


def kernel_times(model, reps: int):
    """Mean duration (ms) of each kernel of the fused sub-step, HIP events on the launch stream."""
    from cmfgpu_amd.autotune import kernel_times as kt
    return kt(model, reps)


def _oracle_from_model(model, net):
    """The CPU oracle on the GPU model's current state (reference order)."""
    import numpy as np
    from cmfgpu_amd.state import assemble
    from oracle.oracle import OracleModel
    src = dict(net.fields)
    src["runoff"] = model.get("runoff")
    host = assemble(src, model.opened_modules, cold_start=False, precise_surfaces=model.precise_surfaces)
    state = {k: v.copy() for k, v in host.fields.items()}
    for name in ("river_storage", "flood_storage", "protected_storage", "protected_depth", "river_depth",
                 "flood_depth", "river_outflow", "flood_outflow", "river_cross_section_depth",
                 "flood_cross_section_depth", "flood_cross_section_area", "flood_fraction",
                 "bifurcation_outflow", "bifurcation_cross_section_depth"):
        if name in state and name in model.t:
            state[name] = np.ascontiguousarray(model.get(name), dtype=state[name].dtype)
    return OracleModel(state, flags=host.flags, num_flood_levels=host.num_flood_levels,
                       num_bifurcation_levels=host.num_bifurcation_levels)


def _timed_reps(orc, n, outer_steps: int, reps: int = 3):
    """Median catchment-sub-steps/s of `reps` runs of `outer_steps` outer steps."""
    import statistics
    rates = []
    subs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        k = int(orc.advance(1800.0, outer_steps).sum())
        rates.append(n * k / (time.perf_counter() - t0))
        subs.append(k)
    return statistics.median(rates), subs


def _cpu_config(config: int, threads: int, all_steps: int, one_steps: int) -> dict:
    """A benchmark configuration on the CPU oracle: one simulated day of
    warm-up (48 outer steps, all threads), then the median of 3 runs of
    `all_steps` outer steps on all threads and of 3 runs of `one_steps` on
    one thread."""
    import numpy as np
    from cmfgpu_amd import synthetic
    from oracle.oracle import OracleModel, set_threads
    from cmfgpu_amd.state import assemble
    net = synthetic.build_config(config)
    src = dict(net.fields)
    rho = synthetic.CONFIGS[config]["rho"]
    forcing = synthetic.Runoff(src["catchment_area"], 2000 + config, rho=rho)
    src["runoff"] = forcing.next_day()
    mods = ("base", "adaptive_time") + (("bifurcation",) if net.num_paths else ()) + \
        (("levee",) if net.num_levees else ())
    host = assemble(src, mods)
    orc = OracleModel({k: v.copy() for k, v in host.fields.items()}, flags=host.flags,
                      num_flood_levels=host.num_flood_levels, num_bifurcation_levels=host.num_bifurcation_levels)
    prev = set_threads(threads)
    t0 = time.perf_counter()
    orc.advance(1800.0, 48)
    warm = time.perf_counter() - t0
    orc.state["runoff"][:] = forcing.next_day()
    rate_all, subs_all = _timed_reps(orc, net.num_catchments, all_steps)
    set_threads(1)
    rate_one, subs_one = _timed_reps(orc, net.num_catchments, one_steps)
    set_threads(prev)
    return {"catchments": net.num_catchments, "paths": net.num_paths, "levees": net.num_levees,
            "warmup": f"one simulated day (48 outer steps, {warm:.1f} s)",
            "all_threads": {"value": rate_all, "threads": threads,
                            "sample": f"median of 3 runs of {all_steps} outer steps (sub-steps {subs_all})"},
            "one_thread": {"value": rate_one,
                           "sample": f"median of 3 runs of {one_steps} outer step(s) (sub-steps {subs_one})"},
            "unit": "catchment-steps/s"}


def cpu_baseline(model, net, budget_s: float):
    """The CPU oracle (C++/OpenMP restatement of the reference step) timed on
    the host cores.  `value`: the same workload as the GPU line, from the GPU
    run's state, median of 3 runs sized to about budget_s / 3 each; one
    thread likewise on one outer step.  `configs`: the 60,000-catchment plain
    configuration and the 270,000-catchment full-physics one, each after a
    simulated day of warm-up, medians of 3 runs on all threads and on one."""
    from oracle.oracle import set_threads
    orc = _oracle_from_model(model, net)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    prev = set_threads(threads)
    t0 = time.perf_counter()
    k0 = int(orc.advance(1800.0, 1)[0])
    per_step = time.perf_counter() - t0
    steps = max(1, int(budget_s / 3 / per_step))
    rate, subs = _timed_reps(orc, net.num_catchments, steps)
    set_threads(1)
    rate1, subs1 = _timed_reps(orc, net.num_catchments, 1)
    set_threads(prev)
    out = {"value": rate, "unit": "catchment-steps/s", "cores": threads, "kind": "port",
           "sample": f"median of 3 runs of {steps} outer steps (sub-steps {subs}) of the same network from "
                     f"the GPU run's state, CPU oracle (C++ restatement, OpenMP, {threads} threads), after one "
                     f"untimed outer step ({k0} sub-steps)",
           "one_thread": {"value": rate1, "unit": "catchment-steps/s",
                          "sample": f"median of 3 runs of 1 outer step (sub-steps {subs1}) on one thread"}}
    out["configs"] = {"60k_plain": _cpu_config(0, threads, 12, 2),
                      "270k_full_physics": _cpu_config(1, threads, 6, 1)}
    return out


def read_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE + WRITE_SIZE runs, calibrated), with where that
    figure comes from: the file and the source commit it records."""
    path = ROOT / "profiles" / "hbm_traffic.txt"
    if not path.exists():
        return None, None
    commit = None
    value = None
    for line in path.read_text().splitlines():
        if line.startswith("# source commit:"):
            commit = line.split(":", 1)[1].strip()
        parts = line.split()
        if (value is None and len(parts) >= 2 and not line.startswith("#")
                and kernel.split("<")[0] == parts[0]):
            try:
                value = float(parts[1])        # the kernel's own row, not a "(first sub-step)" one
            except ValueError:
                value = None
    return value, {"file": "profiles/hbm_traffic.txt", "measured_at_commit": commit,
                   "how": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of the same workload, "
                          "scaled by a device copy of known size"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48, help="timed outer steps (1800 s each)")
    ap.add_argument("--warmup", type=int, default=48, help="untimed outer steps before timing")
    ap.add_argument("--members", type=int, default=64, help="members of the batched line (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=12, help="timed outer steps of the batched run")
    ap.add_argument("--scale", type=float, default=1.0, help="network size factor (1.0 = 1.6M catchments)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--basins", type=int, default=400, help="uneven basins of the grouped line (0: skip)")
    ap.add_argument("--basins-scale", type=float, default=1.0, help="basin size factor (1.0: about 20M catchments)")
    ap.add_argument("--full-members", type=int, default=16,
                    help="members of the batched full-physics line on 6M catchments (0: skip)")
    ap.add_argument("--full-scale", type=float, default=1.0, help="size factor of that network")
    ap.add_argument("--full-steps", type=int, default=4, help="timed outer steps of that line")
    ap.add_argument("--full-days", type=int, default=1, help="simulated days of its FP32-vs-FP64 error run")
    ap.add_argument("--autotune", type=int, default=1, help="pick kernel block sizes by timing first (0: 256)")
    ap.add_argument("--adjoint-steps", type=int, default=720, help="outer steps of the adjoint line (0: skip)")
    ap.add_argument("--graph", type=int, default=1, help="eager vs HIP-graph line on the small network (0: skip)")
    ap.add_argument("--error-days", type=int, default=1,
                    help="simulated days of the FP32-vs-FP64 error run (0: skip)")
    ap.add_argument("--layout", default="auto", choices=("auto", "basin", "level", "network", "keyed"),
                    help="device catchment order")
    ap.add_argument("--reference-surfaces", type=int, default=1,
                    help="also time the plain run with the reference's FP32 surfaces (0: skip)")
    ap.add_argument("--schedule", default="three_pass", choices=("three_pass", "limiter_pass", "block_local"),
                    help="schedule of the fused sub-step (same results; for A/B timing)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()

    global SCHEDULE
    SCHEDULE = args.schedule
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU basin sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cmfgpu_amd import synthetic
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    device = "cuda:0"

    if args.calibrate_copy_mb > 0:
        n = args.calibrate_copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=device)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mb} MiB read + {args.calibrate_copy_mb} MiB written, x3")
        del a, b

    t0 = time.perf_counter()
    net = synthetic.build_config(3, scale=args.scale)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day()
    log(f"network: {net.num_catchments} catchments, {net.num_paths} paths, {net.num_levees} levees "
        f"({time.perf_counter() - t0:.1f} s)")

    # ---------------------------------------------------------- plain --
    # the headline run forms water surfaces in double (precise_surfaces), the
    # FP32 profile that meets a 1e-5 discharge error against FP64; the same
    # run with the reference's FP32 surfaces is timed beside it
    model = build_model(net, src, 1, device, args.layout, precise=True)
    layout_used = model.layout
    tuned = model.autotune(use_cache=False) if args.autotune else None
    if tuned:
        log("autotune: " + ", ".join(f"{k}={v['block']}" for k, v in sorted(tuned.items())))
    run_steps(model, args.warmup, per_trial=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    work = run_steps(model, args.steps, per_trial=False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    value = work / elapsed
    substeps = work // model.N
    log(f"plain: {args.steps} outer steps, {substeps} sub-steps, {elapsed:.3f} s -> {value:.4g} catchment-steps/s")
    reference_surfaces = None
    ref_model = build_model(net, src, 1, device, args.layout, precise=False) if args.reference_surfaces else None
    if ref_model is not None:
        if tuned:
            from cmfgpu_amd.autotune import get_block_sizes, set_block_sizes
            set_block_sizes(ref_model, get_block_sizes(model))
        run_steps(ref_model, args.warmup, per_trial=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rwork = run_steps(ref_model, args.steps, per_trial=False)
        torch.cuda.synchronize()
        relapsed = time.perf_counter() - t1
        reference_surfaces = {"value": rwork / relapsed, "unit": "catchment-steps/s",
                              "ms_per_step": relapsed / args.steps * 1e3, "substeps_timed": int(rwork // model.N)}
        log(f"plain, reference surfaces: {rwork / relapsed:.4g} catchment-steps/s")
        ref_model.close()
        del ref_model

    f_wet = float((model.t["flood_depth"] > 0).float().mean().item())
    ktimes = kernel_times(model, args.kernel_reps)
    dominant = max(ktimes, key=ktimes.get)
    bytes_dom = algorithmic_bytes(dominant, model, f_wet)
    achieved = bytes_dom / (ktimes[dominant] * 1e-3) / 1e9
    substep_ms = sum(ktimes.values())
    bytes_all = sum(algorithmic_bytes(k, model, f_wet) for k in ktimes)
    log("kernels (ms): " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(ktimes.items(), key=lambda x: -x[1])))
    traffic, traffic_source = read_traffic(dominant)
    # the whole sub-step against the compulsory traffic of a fully fused one
    # (wall clock of the timed run, CFL pass and host sync included)
    fused_b = fused_substep_bytes(model.N, model.V, model.P, f_wet)
    us_sub = elapsed / substeps * 1e6
    fused_gbs = fused_b / (us_sub * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_source,
                "kernel": dominant,
                "algorithmic_bytes": int(bytes_dom), "kernel_ms": round(ktimes[dominant], 5),
                "substep_kernels_ms": round(substep_ms, 5),
                "substep_frac": round(bytes_all / (substep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "fused_substep": {"compulsory_bytes": int(fused_b), "us_per_substep": round(us_sub, 2),
                                  "GBps": round(fused_gbs, 1), "frac": round(fused_gbs / HBM_PEAK_GBS, 4)},
                "f_wet": round(f_wet, 4)}

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(model, net, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.4g} catchment-steps/s on {cpu['cores']} threads")
    forcing = forcing_line(model)
    model.close()
    del model
    torch.cuda.empty_cache()

    # -------------------------------------------------------- batched --
    batched = None
    if args.members > 1:
        batched = batched_line(args, net, src, device, value,
                               reference_surfaces["value"] if reference_surfaces else value, tuned)

    basins = None
    if args.basins > 0:
        basins = basins_line(args, device)
    graph = graph_line(args, device) if args.graph else None
    batched_full = batched_full_line(args, device) if args.full_members > 0 else None
    adjoint = adjoint_line(args, device) if args.adjoint_steps > 0 else None

    errors = None
    if args.error_days > 0:
        errors = fp64_errors(net, src, device, args.layout, args.error_days)
        log(f"FP32 vs FP64 after {args.error_days} day(s): discharge rel L2 {errors['discharge_rel_l2']:.3g}; "
            f"precise surfaces {errors['precise_surfaces']['discharge_rel_l2']:.3g} at "
            f"{errors['precise_surfaces']['vs_reference_surfaces']:.3f}x the rate")

    line = {
        "metric": METRIC, "value": value, "unit": "catchment-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 storage)", "data": "synthetic",
        "config": {"workload": "6-arcmin synthetic network, plain method, levees + bifurcation, adaptive CFL, "
                               "water surfaces formed in double (precise_surfaces)",
                   "catchments": net.num_catchments, "bifurcation_paths": net.num_paths,
                   "levees": net.num_levees, "flood_levels": 10, "outer_step_s": 1800.0,
                   "substeps_timed": int(substeps), "layout": layout_used, "schedule": SCHEDULE, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "reference_surfaces": reference_surfaces,
        "batched": batched, "basins": basins, "graph": graph,
        "adjoint": adjoint, "forcing": forcing, "batched_full": batched_full,
        "block_sizes": {k: v["block"] for k, v in tuned.items()} if tuned else None,
        "error_vs_fp64": errors,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
