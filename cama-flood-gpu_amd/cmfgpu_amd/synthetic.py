"""Seeded synthetic river networks and runoff forcing.

The reference has no synthetic generator (it reads parameters.nc built from
CaMa/MERIT maps); this module builds networks of the shape the benchmark
configurations describe, then orders them with the reference's own ordering
pipeline (network.finalize_connectivity), so that the arrays are in the order
a parameters.nc written by MERITMap would hold.

Seeds: network = 1000 + config, forcing = 2000 + config, basin b = 3000 + b,
member m = 4000 + m.

Builder choices (the configurations leave them open):
  * the downstream link of catchment g (generation order, outlets first) is
    g - k with k ~ Geometric(0.01) folded into [1, min(2000, g)];
  * downstream_distance = river_length; the elevation profile (slope
    U(1e-4, 1e-3) times river length, outlets at 0 m) is applied to
    catchment_elevation, the bank-top elevation water-surface slopes are
    taken from; the bed lies a bank depth below it.  Applied to the bed
    instead, bank tops rise downstream wherever the downstream channel is
    deeper, water ponds behind those steps, and FP32 and FP64 runs of such a
    network drift apart by several percent of discharge within a day;
  * levee crown U(1, 5) m above bank; where the derived levee base is not
    below it, the crown is lifted to base + U(1, 5) m;
  * bifurcation length = mean river length of the two ends, width
    lognormal(median 100 m, sigma 0.5), Manning (0.03, 0.1, 0.1, 0.1, 0.1);
  * reservoir capacity lognormal(median 5e7 m^3, sigma 1.0), conservation
    0.40 and emergency 0.95 of capacity, normal outflow = mean inflow at
    1 mm/day over the upstream area, flood-control outflow 3x that.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import network as netmod

NUM_FLOOD_LEVELS = 10
NUM_BIFURCATION_LEVELS = 5
BIFURCATION_LEVEL_HEIGHTS = np.array([0.0, 0.5, 1.0, 2.0, 3.0], dtype=np.float64)
BIFURCATION_MANNING = np.array([0.03, 0.1, 0.1, 0.1, 0.1], dtype=np.float32)

# (catchments, levee fraction, bifurcation paths, reservoirs, gauges, days)
CONFIGS = {
    0: dict(n=60_000, levee=0.0, paths=0, reservoirs=0, gauges=0, days=30, rho=None),
    1: dict(n=270_000, levee=0.05, paths=12_000, reservoirs=0, gauges=0, days=365, rho=0.8),
    2: dict(n=None, levee=0.0, paths=0, reservoirs=0, gauges=0, days=90, rho=None,
            basins=400, basin_min=500, basin_max=400_000),
    3: dict(n=1_600_000, levee=0.05, paths=40_000, reservoirs=0, gauges=2_000, days=15, rho=None),
    4: dict(n=6_000_000, levee=0.05, paths=150_000, reservoirs=7_000, gauges=0, days=30, rho=None,
            members=16),
}


def _pointer_jump_sum(parent: np.ndarray, value: np.ndarray) -> np.ndarray:
    """Sum of `value` along the path from each node to its root (parent < 0)."""
    acc = value.astype(np.float64).copy()
    anc = parent.astype(np.int64).copy()
    while True:
        m = anc >= 0
        if not m.any():
            return acc
        idx = np.flatnonzero(m)
        a = anc[idx]
        acc[idx] = acc[idx] + acc[a]
        anc[idx] = anc[a]


def _pointer_jump_depth(parent: np.ndarray) -> np.ndarray:
    return _pointer_jump_sum(parent, (parent >= 0).astype(np.float64)).astype(np.int64)


def interp_levee_base(table: np.ndarray, levee_idx: np.ndarray, fraction: np.ndarray) -> np.ndarray:
    """LeveeModule.levee_base_height (cmfgpu/modules/levee.py:140-217), float32.

    Linear interpolation of the flood-depth table at position fraction * L,
    with virtual level 0 at depth 0.
    """
    L = table.shape[-1]
    rows = table[levee_idx].astype(np.float32)
    position = np.clip(fraction.astype(np.float32) * np.float32(L), np.float32(0.0), np.float32(L))
    lower = np.floor(position).astype(np.int32)
    upper = lower + 1
    frac = (position - lower.astype(np.float32)).astype(np.float32)
    lower_gather = np.clip(lower - 1, 0, None)
    lower_val = np.take_along_axis(rows, lower_gather[:, None], axis=1)[:, 0]
    lower_val = np.where(lower == 0, np.float32(0.0), lower_val).astype(np.float32)
    upper_idx = np.minimum(upper - 1, L - 1)
    upper_val = np.take_along_axis(rows, upper_idx[:, None], axis=1)[:, 0]
    return (lower_val + frac * (upper_val - lower_val)).astype(np.float32)


@dataclass
class Network:
    """Parameters and initial state of one synthetic network, reference order.

    ``fields`` maps reference module field names to numpy arrays:
    catchment fields [N], flood_depth_table [N, L], bifurcation fields [P] or
    [P, Lb], levee fields [V], reservoir fields [R].
    """

    name: str
    fields: dict = field(default_factory=dict)
    levels: np.ndarray = None          # hop distance to the mouth
    upstream_area: np.ndarray = None   # m^2
    gauges: np.ndarray = None          # catchment indices

    @property
    def num_catchments(self) -> int:
        return int(self.fields["downstream_idx"].shape[0])

    @property
    def num_paths(self) -> int:
        f = self.fields.get("bifurcation_catchment_idx")
        return 0 if f is None else int(f.shape[0])

    @property
    def num_levees(self) -> int:
        f = self.fields.get("levee_catchment_idx")
        return 0 if f is None else int(f.shape[0])

    @property
    def num_reservoirs(self) -> int:
        f = self.fields.get("reservoir_catchment_idx")
        return 0 if f is None else int(f.shape[0])


def generate_network(n: int, seed: int, *, levee: float = 0.0, paths: int = 0,
                     reservoirs: int = 0, gauges: int = 0, name: str = "",
                     downstream: np.ndarray | None = None) -> Network:
    """The seeded random forest above, or with `downstream` a prescribed
    topology in generation order (entry g: the catchment g drains into, an
    index below g, or -1 for an outlet; n is then its length), with the same
    parameter draws."""
    rng = np.random.default_rng(seed)
    g = np.arange(n, dtype=np.int64)
    outlet = rng.random(n) < 0.01
    outlet[0] = True
    k = rng.geometric(0.01, size=n).astype(np.int64)
    window = np.maximum(np.minimum(2000, g), 1)
    k = ((k - 1) % window) + 1
    dn_gen = g - k
    dn_gen[outlet] = -1
    if downstream is not None:
        dn_gen = np.asarray(downstream, dtype=np.int64).copy()
        if dn_gen.shape != (n,) or np.any(dn_gen >= g) or np.any(dn_gen < -1):
            raise ValueError("downstream: n entries, each -1 or an index below its own")
        outlet = dn_gen < 0

    length = rng.lognormal(np.log(15_000.0), 0.5, n)
    width = rng.lognormal(np.log(60.0), 0.8, n)
    height = rng.lognormal(np.log(3.0), 0.5, n)
    area = rng.lognormal(np.log(600e6), 0.7, n)
    slope = np.maximum(rng.uniform(1e-4, 1e-3, n), 1e-5)
    layers = rng.uniform(0.0, 2.0, (n, NUM_FLOOD_LEVELS))
    top = _pointer_jump_sum(dn_gen, np.where(outlet, 0.0, slope * length))

    ids = rng.permutation(n).astype(np.int64) * 7 + 11
    down_ids = np.where(outlet, -1, ids[np.maximum(dn_gen, 0)])
    conn = netmod.finalize_connectivity(ids, down_ids)
    o = conn.order

    f = {}
    f["catchment_id"] = conn.catchment_id
    f["downstream_id"] = conn.downstream_id
    f["downstream_idx"] = conn.downstream_idx
    f["catchment_basin_id"] = conn.catchment_basin_id
    f["river_length"] = length[o].astype(np.float32)
    f["river_width"] = width[o].astype(np.float32)
    f["river_height"] = height[o].astype(np.float32)
    f["catchment_area"] = area[o].astype(np.float32)
    f["downstream_distance"] = f["river_length"].copy()
    f["catchment_elevation"] = top[o].astype(np.float32)
    f["river_manning"] = np.full(n, 0.03, np.float32)
    f["flood_manning"] = np.full(n, 0.10, np.float32)
    f["flood_depth_table"] = np.cumsum(layers[o], axis=1).astype(np.float32)
    # initial state: half bankfull in the channel
    rl, rw, rh = f["river_length"], f["river_width"], f["river_height"]
    f["river_storage"] = 0.5 * (rl.astype(np.float64) * rw.astype(np.float64) * rh.astype(np.float64))
    f["river_depth"] = (np.float32(0.5) * rh).astype(np.float32)

    net = Network(name=name or f"synthetic-n{n}-s{seed}", fields=f)
    didx = conn.downstream_idx
    dn_internal = np.where(didx == np.arange(n), -1, didx)
    net.levels = _pointer_jump_depth(dn_internal)
    net.upstream_area = netmod.accumulate_downstream(didx, f["catchment_area"].astype(np.float64))

    if levee > 0:
        V = int(round(levee * n))
        lidx = np.sort(rng.choice(n, size=V, replace=False)).astype(np.int32)
        frac = rng.uniform(0.1, 0.5, V).astype(np.float32)
        crown = rng.uniform(1.0, 5.0, V).astype(np.float32)
        base = interp_levee_base(f["flood_depth_table"], lidx, frac)
        lift = rng.uniform(1.0, 5.0, V).astype(np.float32)
        crown = np.where(base >= crown, base + lift, crown).astype(np.float32)
        f["levee_catchment_idx"] = lidx
        f["levee_catchment_id"] = conn.catchment_id[lidx]
        f["levee_fraction"] = frac
        f["levee_crown_height"] = crown
        f["levee_base_height"] = base

    if paths > 0:
        lev = net.levels
        order_by_level = np.argsort(lev, kind="stable")
        lev_sorted = lev[order_by_level]
        ci = rng.integers(0, n, size=paths)
        lo = np.searchsorted(lev_sorted, lev[ci] - 5, side="left")
        hi = np.searchsorted(lev_sorted, lev[ci] + 5, side="right")
        pick = lo + np.floor(rng.random(paths) * (hi - lo)).astype(np.int64)
        di = order_by_level[np.minimum(pick, n - 1)]
        same = di == ci
        alt = lo + ((pick - lo + 1) % np.maximum(hi - lo, 1))
        di = np.where(same, order_by_level[np.minimum(alt, n - 1)], di)
        keep = di != ci
        ci, di = ci[keep].astype(np.int32), di[keep].astype(np.int32)
        P = ci.shape[0]
        bed_ci = (f["catchment_elevation"][ci].astype(np.float64) - f["river_height"][ci].astype(np.float64))
        f["bifurcation_catchment_idx"] = ci
        f["bifurcation_downstream_idx"] = di
        f["bifurcation_catchment_id"] = conn.catchment_id[ci]
        f["bifurcation_downstream_id"] = conn.catchment_id[di]
        f["bifurcation_path_id"] = np.arange(P, dtype=np.int64)
        f["bifurcation_length"] = (0.5 * (f["river_length"][ci].astype(np.float64)
                                          + f["river_length"][di].astype(np.float64))).astype(np.float32)
        f["bifurcation_width"] = rng.lognormal(np.log(100.0), 0.5, (P, NUM_BIFURCATION_LEVELS)).astype(np.float32)
        f["bifurcation_elevation"] = (bed_ci[:, None] + BIFURCATION_LEVEL_HEIGHTS[None, :]).astype(np.float32)
        f["bifurcation_manning"] = np.broadcast_to(BIFURCATION_MANNING, (P, NUM_BIFURCATION_LEVELS)).copy()

    if reservoirs > 0:
        ua = net.upstream_area
        candidates = np.flatnonzero((ua > np.median(ua)) & (didx != np.arange(n)))
        R = min(reservoirs, candidates.shape[0])
        ridx = np.sort(rng.choice(candidates, size=R, replace=False)).astype(np.int32)
        cap = rng.lognormal(np.log(5e7), 1.0, R)
        qmean = ua[ridx] * 1e-3 / 86400.0
        f["reservoir_catchment_idx"] = ridx
        f["reservoir_catchment_id"] = conn.catchment_id[ridx]
        f["reservoir_id"] = np.arange(R, dtype=np.int64)
        f["reservoir_capacity"] = cap.astype(np.float32)
        f["conservation_volume"] = (0.40 * cap).astype(np.float32)
        f["emergency_volume"] = (0.95 * cap).astype(np.float32)
        f["normal_outflow"] = qmean.astype(np.float32)
        f["flood_control_outflow"] = (3.0 * qmean).astype(np.float32)
        f["reservoir_area"] = (cap / 10.0).astype(np.float32)

    if gauges > 0:
        cand = np.flatnonzero(net.upstream_area > 5_000e6)
        G = min(gauges, cand.shape[0])
        net.gauges = np.sort(rng.choice(cand, size=G, replace=False)).astype(np.int64)
    return net


def build_config(config: int, scale: float = 1.0) -> Network:
    """Network of one benchmark configuration (basins of config 2: see build_basins)."""
    c = CONFIGS[config]
    if c["n"] is None:
        raise ValueError("configuration 2 is a set of basins; use build_basins")
    n = max(int(round(c["n"] * scale)), 200)
    return generate_network(
        n, 1000 + config,
        levee=c["levee"],
        paths=int(round(c["paths"] * scale)),
        reservoirs=int(round(c["reservoirs"] * scale)),
        gauges=int(round(c["gauges"] * scale)),
        name=f"config{config}" + ("" if scale == 1.0 else f"@{scale:g}"),
    )


def basin_sizes(count: int = 400, lo: int = 500, hi: int = 400_000, seed: int = 3000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


@dataclass
class BasinSet:
    """Independent basins concatenated into one network (reference order of
    each basin, basins one after another); ``group`` labels each catchment
    with its basin and ``basin_ptr`` bounds the basins."""

    network: Network
    group: np.ndarray
    basin_ptr: np.ndarray


def build_basins(count: int = 400, lo: int = 500, hi: int = 400_000, scale: float = 1.0) -> BasinSet:
    """The uneven-basin workload: `count` basins with sizes log-uniform in
    [lo, hi] (seed 3000), basin b generated as the plain network with seed
    3000 + b.  `scale` multiplies every size (at least 50 catchments)."""
    sizes = np.maximum(np.round(basin_sizes(count, lo, hi) * scale), 50).astype(np.int64)
    parts = [generate_network(int(sizes[b]), 3000 + b) for b in range(count)]
    ptr = np.zeros(count + 1, np.int64)
    ptr[1:] = np.cumsum([p.num_catchments for p in parts])
    f = {}
    for name in parts[0].fields:
        arrs = []
        for b, p in enumerate(parts):
            a = p.fields[name]
            if name == "downstream_idx":
                a = (a.astype(np.int64) + ptr[b]).astype(np.int32)
            elif name in ("catchment_id", "downstream_id", "catchment_basin_id"):
                a = a.astype(np.int64) + ptr[b] * 8 + 16 * b
            arrs.append(a)
        f[name] = np.concatenate(arrs)
    net = Network(name=f"basins{count}" + ("" if scale == 1.0 else f"@{scale:g}"), fields=f)
    net.levels = np.concatenate([p.levels for p in parts])
    net.upstream_area = np.concatenate([p.upstream_area for p in parts])
    group = np.repeat(np.arange(count, dtype=np.int32), np.diff(ptr))
    return BasinSet(network=net, group=group, basin_ptr=ptr)


class Runoff:
    """Daily runoff forcing (m^3/s per catchment, float32).

    Depth per day is lognormal with mean 1 mm/day and spatial sigma 1.0
    (mu = -sigma^2 / 2); with `rho` the standard normal field follows
    z_t = rho z_{t-1} + sqrt(1 - rho^2) e_t.
    """

    def __init__(self, catchment_area: np.ndarray, seed: int, rho: float | None = None,
                 sigma: float = 1.0):
        self.area = catchment_area.astype(np.float64)
        self.rng = np.random.default_rng(seed)
        self.rho = rho
        self.sigma = sigma
        self.z = None

    def next_day(self) -> np.ndarray:
        n = self.area.shape[0]
        e = self.rng.standard_normal(n)
        if self.rho is None or self.z is None:
            self.z = e
        else:
            self.z = self.rho * self.z + np.sqrt(1.0 - self.rho ** 2) * e
        depth_mm = np.exp(-0.5 * self.sigma ** 2 + self.sigma * self.z)
        return (depth_mm * 1e-3 / 86400.0 * self.area).astype(np.float32)
