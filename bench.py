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
# from the downstream catchment are taken as cache hits.  See DESIGN.md.
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
        per = 84 + 28 + (5 if lev else 0)
        return per * N * T + (paths if kernel.startswith("k_outflow_paths") else 0)
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


def build_model(net, src, members: int, device: str, layout: str):
    from cmfgpu_amd.models import CaMaFlood
    mods = ["base", "adaptive_time"]
    if net.num_paths:
        mods.append("bifurcation")
    if net.num_levees:
        mods.append("levee")
    if members <= 1:
        return CaMaFlood(src, opened_modules=tuple(mods), device=device, step_duration=1800.0, layout=layout)
    import numpy as np
    from cmfgpu_amd import synthetic
    runoff = np.stack([synthetic.Runoff(src["catchment_area"], 4000 + m).next_day() for m in range(members)])
    return CaMaFlood(src, opened_modules=tuple(mods), device=device, step_duration=1800.0,
                     num_trials=members, batched={"runoff": runoff}, per_trial_time_step=True,
                     trial_layout="inner", layout=layout)


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
        m = CaMaFlood(src, opened_modules=("base", "adaptive_time"), device=device, step_duration=1800.0,
                      groups=bs.group if grouped else None, use_graph=graph)
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
        per = sum(algorithmic_bytes(k, m, f_wet) for k in ("k_outflow", "k_limiter_pair", "k_stage_fused")) / n
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


def batched_full_bytes(N: int, M: int, P: int, f_wet: float) -> float:
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
      and the two ends."""
    return N * M * 184.0 + N * (69.0 + 44.0 * f_wet) + P * M * 112.0 + P * 72.0


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
        return CaMaFlood(src, opened_modules=mods, device=device, step_duration=1800.0, num_trials=M,
                         batched={"runoff": runoff}, per_trial_time_step=True, trial_layout="inner",
                         precision=precision, precise_surfaces=precise)

    out = {"workload": f"{N} catchments, {P} paths, {net.num_levees} levees, {net.num_reservoirs} reservoirs, "
                       f"{M} members with their own CFL counts, trial-inner", "members": M}
    bm = build("float32")
    if args.full_days > 0:
        # FP32 and FP64 profiles from the same state, and FP32 with precise
        # surfaces
        ref = build("float64")
        pr = build("float32", precise=True)
        for day in range(args.full_days):
            if day:
                r = np.stack([f.next_day() for f in forcing])
                for m in (bm, ref, pr):
                    m.set_inputs(r)
            for _ in range(48):
                bm.step_advance()
                ref.step_advance()
                pr.step_advance()
        errors = {}
        for name in ("river_storage", "flood_storage", "flood_depth"):
            x, y = bm.get(name).astype(np.float64), ref.get(name).astype(np.float64)
            errors[name] = [float(f"{np.linalg.norm(x[t] - y[t]) / max(np.linalg.norm(y[t]), 1e-300):.4g}")
                            for t in range(M)]
        q32, q64, qpr = ((m.get("river_outflow").astype(np.float64) + m.get("flood_outflow")) for m in (bm, ref, pr))
        ref.close()
        pr.close()
        errors["discharge"] = [float(f"{np.linalg.norm(q32[t] - q64[t]) / np.linalg.norm(q64[t]):.4g}")
                               for t in range(M)]
        # the same outside the reaches below reservoirs: a release switches
        # regime at storage and inflow thresholds, and a regime that flips
        # between the two profiles changes the flow of the whole reach below
        keep = ~_reaches_below(src, src["reservoir_catchment_idx"])
        errors["discharge_outside_reservoir_reaches"] = [
            float(f"{np.linalg.norm((q32[t] - q64[t])[keep]) / np.linalg.norm(q64[t][keep]):.4g}")
            for t in range(M)]
        precise = {"discharge": [float(f"{np.linalg.norm(qpr[t] - q64[t]) / np.linalg.norm(q64[t]):.4g}")
                                 for t in range(M)],
                   "discharge_outside_reservoir_reaches": [
                       float(f"{np.linalg.norm((qpr[t] - q64[t])[keep]) / np.linalg.norm(q64[t][keep]):.4g}")
                       for t in range(M)]}
        precise["max_discharge_rel_l2"] = max(precise["discharge"])
        precise["max_discharge_rel_l2_outside_reservoir_reaches"] = max(precise["discharge_outside_reservoir_reaches"])
        del q32, q64, qpr
        out.update({"error_days": args.full_days, "rel_l2_per_member": errors, "precise_surfaces": precise,
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
    gbs = batched_full_bytes(N, M, P, f_wet) * subs / el / 1e9
    bm.close()
    out.update({"value": work / el, "unit": "member-catchment-steps/s", "steps": steps,
                "GBps_algorithmic": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                "f_wet": round(f_wet, 4)})
    msg = f"batched full: {work / el:.4g} member-catchment-steps/s, {gbs:.0f} GB/s algorithmic"
    if "max_discharge_rel_l2" in out:
        msg += (f"; discharge error after {args.full_days} day(s) up to {out['max_discharge_rel_l2']:.3g}, "
                f"{out['max_discharge_rel_l2_outside_reservoir_reaches']:.3g} outside reservoir reaches; precise "
                f"surfaces {out['precise_surfaces']['max_discharge_rel_l2']:.3g}, "
                f"{out['precise_surfaces']['max_discharge_rel_l2_outside_reservoir_reaches']:.3g}")
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
        m = CaMaFlood(src, opened_modules=("base", "adaptive_time"), device=device, step_duration=1800.0,
                      use_graph=use_graph)
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


def adjoint_line(args, device: str) -> dict:
    """Cost of the adjoint on the 6-arcmin network with all its modules: the
    base physics, 40,000 bifurcation paths and levees on 5 % of the
    catchments: 2,000 gauges with upstream area > 5,000 km^2,
    `--adjoint-steps` outer steps, checkpoints every 48.  Forward time per
    sub-step against backward time (segment re-run + reverse sweep) per
    sub-step."""
    import numpy as np
    import torch
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.adjoint import GaugeProblem
    from cmfgpu_amd.models import CaMaFlood
    net = synthetic.build_config(3, scale=args.scale)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day()
    model = CaMaFlood(src, opened_modules=("base", "adaptive_time", "bifurcation", "levee"), device=device,
                      step_duration=1800.0)
    gauges = net.gauges
    steps = args.adjoint_steps
    problem = GaugeProblem(model, gauges, np.ones((steps, gauges.shape[0])), checkpoint_every=48)
    # warm-up: the same forward and backward once (first use of every kernel,
    # and the tape of a full checkpoint segment, tens of GB at this size, is
    # allocated once by the caching allocator, as in a calibration loop that
    # takes many gradients); then the timed pass from the same initial state
    problem.run()
    problem.gradient()
    problem._restore(problem.initial)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    problem.run()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_n, g_h = problem.gradient()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    sub = sum(problem.counts)
    out = {"workload": f"{net.num_catchments} catchments, {net.num_paths} bifurcation paths, {net.num_levees} "
                       f"levees, {gauges.shape[0]} gauges",
           "outer_steps": steps, "substeps": sub, "forward_ms_per_substep": t_fwd / sub * 1e3,
           "backward_ms_per_substep": t_bwd / sub * 1e3, "ratio": round(t_bwd / t_fwd, 3),
           "gradient_finite": bool(torch.isfinite(g_n).all().item() and torch.isfinite(g_h).all().item())}
    log(f"adjoint: forward {out['forward_ms_per_substep']:.3f} ms/sub-step, backward "
        f"{out['backward_ms_per_substep']:.3f} ms/sub-step ({out['ratio']}x)")
    model.close()
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


def cpu_baseline(model, net, budget_s: float):
    """CPU oracle on the GPU model's current state; bounded sample of outer steps."""
    import numpy as np
    from cmfgpu_amd.state import assemble
    from oracle.oracle import OracleModel
    src = dict(net.fields)
    src["runoff"] = model.get("runoff")
    host = assemble(src, model.opened_modules, cold_start=False)
    state = {k: v.copy() for k, v in host.fields.items()}
    for name in ("river_storage", "flood_storage", "protected_storage", "protected_depth", "river_depth",
                 "flood_depth", "river_outflow", "flood_outflow", "river_cross_section_depth",
                 "flood_cross_section_depth", "flood_cross_section_area", "flood_fraction",
                 "bifurcation_outflow", "bifurcation_cross_section_depth"):
        if name in state and name in model.t:
            state[name] = np.ascontiguousarray(model.get(name), dtype=state[name].dtype)
    orc = OracleModel(state, flags=host.flags, num_flood_levels=host.num_flood_levels,
                      num_bifurcation_levels=host.num_bifurcation_levels)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    substeps = 0
    steps = 0
    t0 = time.perf_counter()
    while True:
        substeps += int(orc.advance(1800.0, 1)[0])
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    # one thread: one outer step more from where the sample ended
    from oracle.oracle import set_threads
    prev = set_threads(1)
    t1 = time.perf_counter()
    one_sub = int(orc.advance(1800.0, 1)[0])
    dt1 = time.perf_counter() - t1
    set_threads(prev)
    return {"value": net.num_catchments * substeps / dt, "unit": "catchment-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} outer steps ({substeps} CFL sub-steps) of the same network from the GPU "
                      f"run's state, CPU oracle (C++ restatement, OpenMP, {threads} threads), {dt:.1f} s",
            "one_thread": {"value": net.num_catchments * one_sub / dt1, "unit": "catchment-steps/s",
                           "sample": f"the next outer step ({one_sub} CFL sub-steps) on one thread, {dt1:.1f} s"}}


def read_traffic(kernel: str):
    path = ROOT / "profiles" / "hbm_traffic.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 2 and not line.startswith("#") and kernel.split("<")[0] == parts[0]:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48, help="timed outer steps (1800 s each)")
    ap.add_argument("--warmup", type=int, default=48, help="untimed outer steps before timing")
    ap.add_argument("--members", type=int, default=64, help="members of the batched line (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=0, help="timed outer steps of the batched run")
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
    ap.add_argument("--adjoint-steps", type=int, default=48, help="outer steps of the adjoint line (0: skip)")
    ap.add_argument("--graph", type=int, default=1, help="eager vs HIP-graph line on the small network (0: skip)")
    ap.add_argument("--error-days", type=int, default=1,
                    help="simulated days of the FP32-vs-FP64 error run (0: skip)")
    ap.add_argument("--layout", default="basin", choices=("basin", "level", "network"),
                    help="device catchment order")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()

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
    model = build_model(net, src, 1, device, args.layout)
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

    f_wet = float((model.t["flood_depth"] > 0).float().mean().item())
    ktimes = kernel_times(model, args.kernel_reps)
    dominant = max(ktimes, key=ktimes.get)
    bytes_dom = algorithmic_bytes(dominant, model, f_wet)
    achieved = bytes_dom / (ktimes[dominant] * 1e-3) / 1e9
    substep_ms = sum(ktimes.values())
    bytes_all = sum(algorithmic_bytes(k, model, f_wet) for k in ktimes)
    log("kernels (ms): " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(ktimes.items(), key=lambda x: -x[1])))
    traffic = read_traffic(dominant)
    # the whole sub-step against the compulsory traffic of a fully fused one
    # (wall clock of the timed run, CFL pass and host sync included)
    fused_b = fused_substep_bytes(model.N, model.V, model.P, f_wet)
    us_sub = elapsed / substeps * 1e6
    fused_gbs = fused_b / (us_sub * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dominant,
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
        M = args.members
        bm = build_model(net, src, M, device, args.layout)
        if args.autotune:
            bm.autotune(use_cache=False)
        run_steps(bm, max(args.warmup // 4, 1), per_trial=True)
        bsteps = args.batched_steps or max(args.steps // 4, 2)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        bwork = run_steps(bm, bsteps, per_trial=True)
        torch.cuda.synchronize()
        belapsed = time.perf_counter() - t2
        bvalue = bwork / belapsed
        # whole-run algorithmic rate, shared parameters counted once per catchment
        bf_wet = float((bm.t["flood_depth"] > 0).float().mean().item())
        bgbs = batched_full_bytes(bm.N, M, bm.P, bf_wet) * (bwork / (bm.N * M)) / belapsed / 1e9
        cgbs = fused_batched_bytes(bm.N, M, bm.V, bm.P, bf_wet) * (bwork / (bm.N * M)) / belapsed / 1e9
        batched = {"members": M, "value": bvalue, "unit": "member-catchment-steps/s", "steps": bsteps,
                   "vs_plain": round(bvalue / value, 3), "GBps_algorithmic": round(bgbs, 1),
                   "frac_of_peak": round(bgbs / HBM_PEAK_GBS, 4),
                   "GBps_compulsory_fused": round(cgbs, 1), "frac_compulsory_fused": round(cgbs / HBM_PEAK_GBS, 4)}
        log(f"batched: {M} members, {bsteps} outer steps, {belapsed:.3f} s -> {bvalue:.4g} "
            f"member-catchment-steps/s ({bvalue / value:.2f}x plain), {bgbs:.0f} GB/s algorithmic")
        bm.close()

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
        "config": {"workload": "6-arcmin synthetic network, plain method, levees + bifurcation, adaptive CFL",
                   "catchments": net.num_catchments, "bifurcation_paths": net.num_paths,
                   "levees": net.num_levees, "flood_levels": 10, "outer_step_s": 1800.0,
                   "substeps_timed": int(substeps), "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "basins": basins, "graph": graph,
        "adjoint": adjoint, "forcing": forcing, "batched_full": batched_full,
        "block_sizes": {k: v["block"] for k, v in tuned.items()} if tuned else None,
        "error_vs_fp64": errors,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
