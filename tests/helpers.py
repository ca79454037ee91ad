"""Shared test helpers: small networks, oracle and device runs."""
from __future__ import annotations

import numpy as np

from cmfgpu_amd import synthetic
from cmfgpu_amd.state import assemble
from oracle.oracle import OracleModel

FULL = ("base", "adaptive_time", "bifurcation", "levee", "reservoir")
STATE_FIELDS = ("river_outflow", "flood_outflow", "river_storage", "flood_storage", "protected_storage",
                "river_depth", "flood_depth", "protected_depth", "flood_fraction",
                "river_cross_section_depth", "flood_cross_section_depth", "flood_cross_section_area",
                "river_inflow", "flood_inflow")


def small_source(n=3000, seed=7, levee=0.05, paths=60, reservoirs=10, runoff_seed=2000):
    net = synthetic.generate_network(n, seed, levee=levee, paths=paths, reservoirs=reservoirs)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], runoff_seed).next_day()
    return net, src


def oracle_for(src, modules, precision="float32", log_buffer_size=0, **kw):
    host = assemble(src, modules, precision=precision, log_buffer_size=log_buffer_size, **kw)
    state = {k: v.copy() for k, v in host.fields.items()}
    return OracleModel(state, flags=host.flags, num_flood_levels=host.num_flood_levels,
                       num_bifurcation_levels=host.num_bifurcation_levels, log_buffer_size=log_buffer_size,
                       precision=(0 if host.sto == np.float64 else 2) if host.real == np.float32 else 1,
                       gravity=host.gravity, adaptive_time_factor=host.adaptive_time_factor,
                       min_kinematic_slope=host.min_kinematic_slope), host


def mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if not bad.any():
        return None
    i = np.flatnonzero(bad.ravel())[0]
    return f"{int(bad.sum())} elements differ, first at {i}: {a.ravel()[i]!r} vs {b.ravel()[i]!r}"
