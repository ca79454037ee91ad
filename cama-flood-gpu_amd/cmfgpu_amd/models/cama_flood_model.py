"""CaMaFlood: the model driver (cmfgpu/models/cama_flood_model.py:46-173).

Holds every module field as a device tensor, owns the libcmfhip network
context, and advances outer steps either through the fused C-ABI schedule
(``mode="fused"``, one call per outer step) or through the reference's own
per-op sequence (``mode="ops"``, the zero-argument registered ops called in
the order of step_advance, cama_flood_model.py:104-173).

Device layout.  With ``layout="basin"`` catchments are
renumbered breadth-first from the mouths inside each basin, basin after basin
(network.basin_level_order), so that the upstream catchments of every
catchment are one contiguous range next to it; ``layout="level"`` uses the
breadth-first levels of the whole forest (network.level_order); with
``layout="network"`` the reference's own order is kept; ``layout="keyed"``
uses the forest-wide levels with each level's sibling groups arranged by the
work they carry (levee catchments, bifurcation path ends, both, neither:
network.keyed_level_order), so that the levee partition and the path nets run
in few waves.  ``layout="auto"`` (the default) is "keyed" for one trial with
levees or bifurcation paths and no groups, "basin" otherwise.  With trials,
``trial_layout="inner"`` stores element (trial t, catchment i) at i*T + t
(all members of a catchment adjacent), ``"major"`` at t*N + i (the
reference's [T, N]).  ``get``/``set_inputs`` always speak the reference's
order and shapes.
"""
from __future__ import annotations

import ctypes
from datetime import datetime, timedelta
from pathlib import Path

import numpy as np
import torch

from .. import _lib
from .. import network as netmod
from ..phys import (compute_adaptive_time_step, compute_bifurcation_inflow, compute_bifurcation_outflow,
                    compute_flood_stage, compute_flood_stage_log, compute_inflow,
                    compute_levee_bifurcation_outflow, compute_levee_stage, compute_levee_stage_log,
                    compute_outflow, compute_reservoir_outflow)
from ..phys.registry import bind
from ..state import assemble

CATCHMENT_SHARED = ("river_width", "river_length", "river_height", "catchment_elevation", "catchment_area",
                    "downstream_distance", "river_manning", "flood_manning", "flood_depth_table", "runoff",
                    "is_levee", "is_reservoir", "is_dam_related", "is_dam_upstream")
CATCHMENT_STATE = ("river_storage", "flood_storage", "protected_storage", "protected_depth", "river_depth",
                   "flood_depth", "river_outflow", "flood_outflow", "river_cross_section_depth",
                   "flood_cross_section_depth", "flood_cross_section_area", "flood_fraction", "river_inflow",
                   "flood_inflow", "outgoing_storage", "total_storage", "water_surface_elevation",
                   "protected_water_surface_elevation", "global_bifurcation_outflow", "limit_rate",
                   "reservoir_total_inflow")
PATH_SHARED = ("bifurcation_manning", "bifurcation_width", "bifurcation_length", "bifurcation_elevation")
PATH_STATE = ("bifurcation_outflow", "bifurcation_cross_section_depth")
LEVEE_PARAMS = ("levee_base_height", "levee_crown_height", "levee_fraction")
RESERVOIR_PARAMS = ("conservation_volume", "emergency_volume", "adjustment_volume", "effective_normal_outflow",
                    "adjustment_outflow", "flood_control_outflow")
INDEX_FIELDS = ("downstream_idx", "bifurcation_catchment_idx", "bifurcation_downstream_idx",
                "levee_catchment_idx", "reservoir_catchment_idx")
LOG_PTR_NAMES = {name: k for k, name in enumerate(_lib.LOG_FIELDS)}


def _torch_dtype(a: np.ndarray):
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
            np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
            np.dtype(np.bool_): torch.bool}[a.dtype]


def _group_basins(order, groups: np.ndarray, dn: np.ndarray):
    """Order the basins of a basin layout group by group (ascending label,
    basins keeping their order inside a group).  Returns the new permutation,
    the group boundaries in internal order and the sorted labels.  Every
    catchment of a basin must carry its basin's label."""
    n = dn.shape[0]
    if groups.shape != (n,):
        raise ValueError(f"groups must have one label per catchment ({n}), got shape {groups.shape}")
    if np.any(groups[dn] != groups):
        bad = int(np.flatnonzero(groups[dn] != groups)[0])
        raise ValueError(f"catchment {bad} drains into another group: a group must be a union of basins")
    bptr = order.basin_ptr
    mouths = order.perm[bptr[1:] - 1]
    blabel = groups[mouths]
    labels, gidx = np.unique(blabel, return_inverse=True)
    border = np.argsort(gidx, kind="stable")
    sizes = np.diff(bptr)
    perm = np.concatenate([order.perm[bptr[b]:bptr[b + 1]] for b in border]).astype(np.int64)
    gsize = np.bincount(gidx, weights=sizes, minlength=labels.shape[0]).astype(np.int64)
    group_ptr = np.zeros(labels.shape[0] + 1, np.int64)
    group_ptr[1:] = np.cumsum(gsize)
    return perm, group_ptr, labels



def _path_order(ci: np.ndarray, di: np.ndarray) -> np.ndarray:
    """Device order of the bifurcation paths (entry q: the input path placed
    q-th), from their ends' device positions ci, di (input order).

    Paths are taken by ascending device position of their upstream end, so
    that the path pass and the path ends of a wave touch consecutive paths,
    subject to one rule: the paths of every catchment keep their input order
    among themselves, since each catchment sums its paths' terms in ascending
    path order (the reference's order).  That is a linear extension of the
    per-catchment chains (edges run from lower to higher input index, so the
    graph has no cycle), formed with a heap keyed by the upstream end."""
    import heapq
    P = int(ci.shape[0])
    if P == 0:
        return np.zeros(0, np.int64)
    c = np.concatenate([ci, di]).astype(np.int64)
    p = np.concatenate([np.arange(P), np.arange(P)])
    o = np.lexsort((p, c))
    c, p = c[o], p[o]
    same = c[1:] == c[:-1]
    src, dst = p[:-1][same], p[1:][same]
    indeg = np.bincount(dst, minlength=P)
    so = np.argsort(src, kind="stable")
    src, dst = src[so], dst[so]
    start = np.searchsorted(src, np.arange(P + 1))
    key = ci.astype(np.int64)
    heap = [(int(key[q]), q) for q in np.flatnonzero(indeg == 0).tolist()]
    heapq.heapify(heap)
    order = []
    while heap:
        _, q = heapq.heappop(heap)
        order.append(q)
        for r in dst[start[q]:start[q + 1]].tolist():
            indeg[r] -= 1
            if indeg[r] == 0:
                heapq.heappush(heap, (int(key[r]), r))
    if len(order) != P:
        raise RuntimeError("bifurcation path order: the per-catchment chains do not cover every path")
    return np.asarray(order, np.int64)


class CaMaFlood:
    def __init__(self, src: dict, *, opened_modules=("base", "adaptive_time"), precision="float32",
                 device="cuda:0", num_trials: int = 1, batched: dict | None = None,
                 per_trial_time_step: bool = False, layout: str = "auto", trial_layout: str = "inner",
                 log_buffer_size: int = 20000, step_duration: float = 1800.0, num_sub_steps: int = 1,
                 start_time: datetime = datetime(2000, 1, 1), log_path: str | Path | None = None,
                 mode: str = "fused", gravity: float = 9.8, adaptive_time_factor: float = 0.7,
                 min_kinematic_slope: float = 1e-5, groups=None, use_graph: bool = False,
                 precise_surfaces: bool = False, mixed_precision: bool = True):
        if mode not in ("fused", "ops"):
            raise ValueError("mode must be 'fused' or 'ops'")
        if layout not in ("auto", "basin", "level", "network", "keyed"):
            raise ValueError("layout must be 'auto', 'basin', 'level', 'network' or 'keyed'")
        if trial_layout not in ("inner", "major"):
            raise ValueError("trial_layout must be 'inner' or 'major'")
        if num_trials < 1:
            raise ValueError("num_trials must be >= 1")
        opened = tuple(dict.fromkeys(("base",) + tuple(opened_modules)))
        if "log" in opened and num_trials > 1:
            raise ValueError("the log module does not support trials (cama_flood_model.py:65-67)")
        if mode == "ops" and per_trial_time_step:
            raise ValueError("per-trial time steps need the fused schedule")
        if groups is not None and (mode != "fused" or num_trials != 1 or layout not in ("basin", "auto")):
            raise ValueError("groups need the fused schedule, one trial and the basin layout")
        if groups is not None and log_path is not None:
            raise ValueError("the text log is global; read per-group sums with log_sums()")
        self.opened_modules = opened
        self.mode = mode
        if use_graph and mode != "fused":
            raise ValueError("graph replay needs the fused schedule")
        self.use_graph = bool(use_graph)
        self.device = torch.device(device)
        self.T = int(num_trials)
        self.per_trial = bool(per_trial_time_step)
        self.trial_layout = trial_layout if self.T > 1 else "major"
        self.step_duration = float(step_duration)
        self.num_sub_steps = int(num_sub_steps)
        self.current_time = start_time
        self.log_path = Path(log_path) if log_path is not None else None
        self.host = assemble(src, opened, precision=precision, log_buffer_size=log_buffer_size,
                             gravity=gravity, adaptive_time_factor=adaptive_time_factor,
                             min_kinematic_slope=min_kinematic_slope, precise_surfaces=precise_surfaces,
                             mixed_precision=mixed_precision)
        self.precise_surfaces = bool(precise_surfaces)
        self.flags = self.host.flags | (_lib.PER_TRIAL_DT if self.per_trial else 0)
        self.real = self.host.real
        hf = self.host.fields
        self.N = self.host.num_catchments
        self.P = int(hf["bifurcation_catchment_idx"].shape[0]) if "bifurcation_catchment_idx" in hf else 0
        self.Lb = self.host.num_bifurcation_levels
        self.V = int(hf["levee_catchment_idx"].shape[0]) if "levee_catchment_idx" in hf else 0
        self.R = int(hf["reservoir_catchment_idx"].shape[0]) if "reservoir_catchment_idx" in hf else 0
        self.L = self.host.num_flood_levels
        if layout == "auto":
            layout = "keyed" if (groups is None and self.T == 1 and (self.V or self.P)) else "basin"

        dn = hf["downstream_idx"]
        self.G = 0
        self.group_labels = None
        group_ptr = None
        if layout == "basin":
            self.order = netmod.basin_level_order(dn)
            self.perm = self.order.perm.astype(np.int64)
            if groups is not None:
                self.perm, group_ptr, self.group_labels = _group_basins(self.order, np.asarray(groups), dn)
                self.G = int(group_ptr.shape[0] - 1)
                self.inv = np.empty_like(self.perm)
                self.inv[self.perm] = np.arange(self.N, dtype=np.int64)
                dn_int = self.inv[dn[self.perm]].astype(np.int32)
            else:
                self.inv = self.order.inv.astype(np.int64)
                dn_int = self.order.dn.astype(np.int32)
        elif layout == "level":
            self.order = netmod.level_order(dn)
            self.perm = self.order.perm.astype(np.int64)
            self.inv = self.order.inv.astype(np.int64)
            dn_int = self.order.dn.astype(np.int32)
        elif layout == "keyed":
            key = np.zeros(self.N, np.uint8)
            if self.V:
                key[np.asarray(hf["levee_catchment_idx"], np.int64)] |= 1
            if self.P:
                key[np.asarray(hf["bifurcation_catchment_idx"], np.int64)] |= 2
                key[np.asarray(hf["bifurcation_downstream_idx"], np.int64)] |= 2
            self.order = netmod.keyed_level_order(dn, key)
            self.perm = self.order.perm.astype(np.int64)
            self.inv = self.order.inv.astype(np.int64)
            dn_int = self.order.dn.astype(np.int32)
        else:
            netmod.check_flow_direction(dn)
            self.order = None
            self.perm = np.arange(self.N, dtype=np.int64)
            self.inv = self.perm.copy()
            dn_int = dn.astype(np.int32)
        self.layout = layout

        batched = dict(batched or {})
        if (self.V and "levee_base_height" not in batched
                and batched.keys() & {"levee_fraction", "flood_depth_table"}):
            # the base height is derived from the fraction over the floodplain
            # table (levee.py:140-217), so a trial with its own fraction or
            # table gets its own base height
            batched["levee_base_height"] = self._trial_levee_base(batched)
        self.batched_names = set(batched)
        bits = 0
        for name in batched:
            if name not in _lib.BATCHED_BITS:
                raise ValueError(f"{name} cannot be batched")
            bits |= _lib.BATCHED_BITS[name]
        self.batched_bits = bits

        self.t = {}
        self.t["downstream_idx"] = torch.from_numpy(dn_int).to(self.device)
        for name in CATCHMENT_SHARED:
            if name in hf:
                src_arr = batched.get(name, hf[name])
                self.t[name] = self._to_catchment(name, np.asarray(src_arr, dtype=hf[name].dtype),
                                                  per_trial=name in batched)
        for name in CATCHMENT_STATE:
            if name in hf:
                self.t[name] = self._to_catchment(name, hf[name], per_trial=True, replicate=True)
        if "bifurcation_catchment_idx" in hf:
            # (with the module open and no paths the arrays are empty, still bound)
            ci = self.inv[hf["bifurcation_catchment_idx"]].astype(np.int32)
            di = self.inv[hf["bifurcation_downstream_idx"]].astype(np.int32)
            self.pperm = _path_order(ci, di)
            self.pinv = np.empty(self.P, np.int64)
            self.pinv[self.pperm] = np.arange(self.P)
            ci = np.ascontiguousarray(ci[self.pperm])
            di = np.ascontiguousarray(di[self.pperm])
            self.t["bifurcation_catchment_idx"] = torch.from_numpy(ci).to(self.device)
            self.t["bifurcation_downstream_idx"] = torch.from_numpy(di).to(self.device)
            for name in PATH_SHARED:
                src_arr = batched.get(name, hf[name])
                self.t[name] = self._to_path(np.asarray(src_arr, dtype=hf[name].dtype), per_trial=name in batched)
            for name in PATH_STATE:
                self.t[name] = self._to_path(hf[name], per_trial=True, replicate=True)
        else:
            ci = di = np.zeros(0, np.int32)
        if self.V:
            # levee rows in the device order of their catchments (row r holds
            # the input's levee lperm[r]): a wave of levee catchments then
            # reads consecutive rows of the levee parameters
            lidx_in = self.inv[hf["levee_catchment_idx"]].astype(np.int64)
            self.lperm = np.argsort(lidx_in, kind="stable")
            self.linv = np.empty(self.V, np.int64)
            self.linv[self.lperm] = np.arange(self.V)
            lidx_dev = lidx_in[self.lperm].astype(np.int32)
            self.t["levee_catchment_idx"] = torch.from_numpy(lidx_dev).to(self.device)
            levee_index = np.full(self.N, -1, np.int32)
            levee_index[lidx_dev] = np.arange(self.V, dtype=np.int32)
            self.t["levee_index"] = torch.from_numpy(levee_index).to(self.device)
            for name in LEVEE_PARAMS:
                src_arr = np.asarray(batched.get(name, hf[name]), dtype=hf[name].dtype)[..., self.lperm]
                self.t[name] = torch.from_numpy(np.ascontiguousarray(src_arr)).to(self.device)
            if batched.keys() & {"levee_crown_height", "levee_base_height", "levee_fraction"}:
                # per-trial levee parameters obey the module's rules in every trial
                rows = {k: np.broadcast_to(np.asarray(batched.get(k, hf[k])), (self.T, self.V))
                        for k in ("levee_crown_height", "levee_base_height", "levee_fraction")}
                frac, crown, base = rows["levee_fraction"], rows["levee_crown_height"], rows["levee_base_height"]
                if (~np.isfinite(frac) | (frac < 0) | (frac >= 1)).any():
                    raise ValueError("levee_fraction must be finite and lie within [0, 1) in every trial")
                bad = int((~np.isfinite(crown) | ~np.isfinite(base) | (base >= crown)).sum())
                if bad:
                    raise ValueError(f"{bad} trial levees have non-finite heights or a base height greater than or "
                                     f"equal to the crown height")
        if self.R:
            self.t["reservoir_catchment_idx"] = torch.from_numpy(self.inv[hf["reservoir_catchment_idx"]].astype(np.int32)).to(self.device)
            for name in RESERVOIR_PARAMS:
                self.t[name] = torch.from_numpy(np.ascontiguousarray(hf[name])).to(self.device)
        if "log_sums" in hf:
            ls = hf["log_sums"]
            if self.G:
                ls = np.zeros((ls.shape[0], self.G * ls.shape[1]), ls.dtype)
            self.t["log_sums"] = torch.from_numpy(ls.copy()).to(self.device)
        tdt = torch.float32 if self.real == np.float32 else torch.float64
        slots = self.G if self.G else (self.T if self.per_trial else 1)
        self.t["time_step"] = torch.zeros(slots, dtype=tdt, device=self.device)
        self.t["outer_time_step"] = torch.zeros(1, dtype=tdt, device=self.device)
        self.t["current_step"] = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.t["max_sub_steps"] = torch.zeros(slots, dtype=torch.int32, device=self.device)

        # network context (host indices in device order)
        h = ctypes.c_void_p()
        gp = group_ptr if group_ptr is not None else np.zeros(1, np.int64)
        _lib.check(_lib.lib().cmf_network_create_grouped(
            self.N, dn_int.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), self.P,
            ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), di.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            self.T, self.G, gp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(h)),
            "cmf_network_create_grouped")
        self.network_handle = h
        self.contiguous = bool(_lib.lib().cmf_network_is_contiguous(h))
        self._table = self._build_table()
        self.step_count = 0
        self.last_sub_steps = 0
        self._perm_dev = None
        # with the log module: keep each outer step's sums in last_log_sums
        self.keep_log_sums = False
        self.last_log_sums = None

    def _trial_levee_base(self, batched: dict) -> np.ndarray:
        """[T, V] levee base heights, each trial's from its own fraction and
        floodplain table (LeveeModule._interp_lookup, levee.py:172-217)."""
        from ..modules.levee import interp_lookup
        hf = self.host.fields
        real = hf["levee_fraction"].dtype.type
        lidx = hf["levee_catchment_idx"]
        frac = np.broadcast_to(np.asarray(batched.get("levee_fraction", hf["levee_fraction"]), dtype=real),
                               (self.T, self.V))
        table = np.asarray(batched.get("flood_depth_table", hf["flood_depth_table"]),
                           dtype=hf["flood_depth_table"].dtype)
        L = table.shape[-1]
        rows = []
        for t in range(self.T):
            tab = table[t] if table.ndim == 3 else table
            rows.append(interp_lookup(tab, lidx, (frac[t] * real(L)).astype(real)))
        return np.stack(rows).astype(real)

    # ------------------------------------------------------------ layout --
    def _to_catchment(self, name, arr, per_trial=False, replicate=False):
        """Host array in network order -> device tensor in the device layout.

        replicate: one [N, ...] array copied to every trial (on the device).
        per_trial without replicate: a [T, N, ...] array of per-trial values.
        """
        a = np.asarray(arr)
        if per_trial and self.T > 1 and not replicate:
            if a.shape[0] != self.T:
                raise ValueError(f"{name}: expected a leading trial dimension of {self.T}")
            x = torch.from_numpy(np.ascontiguousarray(a[:, self.perm])).to(self.device)
            return x.transpose(0, 1).contiguous() if self.trial_layout == "inner" else x
        x = torch.from_numpy(np.ascontiguousarray(a[self.perm])).to(self.device)
        if per_trial and self.T > 1:
            if self.trial_layout == "inner":
                return x.unsqueeze(1).expand(x.shape[0], self.T, *x.shape[1:]).contiguous()
            return x.unsqueeze(0).expand(self.T, *x.shape).contiguous()
        return x

    def _to_path(self, arr, per_trial=False, replicate=False):
        """Host array in the input's path order ([P, ...], or [T, P, ...] per
        trial) -> device tensor with the paths in the device order (pperm)."""
        a = np.asarray(arr)
        if per_trial and self.T > 1 and not replicate:
            x = torch.from_numpy(np.ascontiguousarray(a[:, self.pperm])).to(self.device)
            return x.transpose(0, 1).contiguous() if self.trial_layout == "inner" else x
        x = torch.from_numpy(np.ascontiguousarray(a[self.pperm])).to(self.device)
        if per_trial and self.T > 1:
            if self.trial_layout == "inner":
                return x.unsqueeze(1).expand(x.shape[0], self.T, *x.shape[1:]).contiguous()
            return x.unsqueeze(0).expand(self.T, *x.shape).contiguous()
        return x

    def _strides(self):
        if self.T > 1 and self.trial_layout == "inner":
            return 1, self.T, 1, self.T
        return self.N, 1, self.P, 1

    def _build_table(self) -> _lib.CmfFields:
        f = _lib.CmfFields()
        f.num_catchments = self.N
        f.num_trials = self.T
        f.num_flood_levels = self.L
        f.num_bifurcation_paths = self.P
        f.num_bifurcation_levels = self.Lb
        f.num_levees = self.V
        f.num_reservoirs = self.R
        f.log_buffer_size = self.host.log_buffer_size
        f.trial_stride, f.catchment_stride, f.path_trial_stride, f.path_stride = self._strides()
        if self.real == np.float64:
            f.precision = _lib.PRECISION_F64_F64
        else:
            f.precision = _lib.PRECISION_F32_F64 if self.host.sto == np.float64 else _lib.PRECISION_F32_F32
        f.flags = self.flags
        f.batched = self.batched_bits
        f.gravity = self.host.gravity
        f.adaptive_time_factor = self.host.adaptive_time_factor
        f.min_kinematic_slope = self.host.min_kinematic_slope
        for name in _lib.POINTER_FIELDS:
            tensor = self.t.get(name)
            setattr(f, name, ctypes.c_void_p(tensor.data_ptr()) if tensor is not None else None)
        if "log_sums" in self.t:
            ls = self.t["log_sums"]
            row = ls.stride(0) * ls.element_size()
            for k in range(len(_lib.LOG_FIELDS)):
                f.log_sums[k] = ls.data_ptr() + k * row
        return f

    def field_table_ref(self):
        return ctypes.byref(self._table)

    # ---------------------------------------------------------- fields ----
    def get(self, name: str) -> np.ndarray:
        """A field in the reference's order and shape ([T, ...] for per-trial fields when T > 1)."""
        x = self.t[name].cpu().numpy()
        if name in INDEX_FIELDS:
            if name == "downstream_idx":
                return self.host.fields["downstream_idx"].copy()
            if name == "levee_catchment_idx":
                x = x[self.linv]
            elif name in ("bifurcation_catchment_idx", "bifurcation_downstream_idx"):
                x = x[self.pinv]
            return self.perm[x].astype(np.int32)
        if name in LEVEE_PARAMS:
            return np.ascontiguousarray(x[..., self.linv])
        if name in CATCHMENT_STATE or name in CATCHMENT_SHARED:
            per_trial = name in CATCHMENT_STATE or name in self.batched_names
            if per_trial and self.T > 1:
                if self.trial_layout == "inner":
                    x = np.swapaxes(x, 0, 1)
                return x[:, self.inv]
            return x[self.inv]
        if name in PATH_STATE or name in PATH_SHARED:
            per_trial = name in PATH_STATE or name in self.batched_names
            if per_trial and self.T > 1:
                if self.trial_layout == "inner":
                    x = np.swapaxes(x, 0, 1)
                return np.ascontiguousarray(x[:, self.pinv])
            return np.ascontiguousarray(x[self.pinv])
        return x

    def log_sums(self) -> np.ndarray:
        """The water-balance sums, [11, log_buffer_size] (with groups [11, G, log_buffer_size])."""
        x = self.t["log_sums"].cpu().numpy()
        if self.G:
            return x.reshape(x.shape[0], self.G, -1)
        return x

    def enable_outflow_mean(self) -> None:
        """Accumulate the time integral of total outflow (river + flood) over
        every sub-step on the device (the reference's "mean" statistics of
        total_outflow); read it with outflow_mean()."""
        self.t["acc_outflow"] = torch.zeros(self.N * self.T, dtype=torch.float64, device=self.device)
        self._acc_seconds = 0.0
        _lib.check(_lib.lib().cmf_network_set_outflow_accumulator(self.network_handle,
                                                                  _lib.ptr(self.t["acc_outflow"])),
                   "cmf_network_set_outflow_accumulator")

    def outflow_mean(self, reset: bool = True) -> np.ndarray:
        """Time-mean total outflow since the last reset (m^3/s), reference order
        ([T, N] with trials)."""
        acc = self.t["acc_outflow"]
        secs = self._acc_seconds
        x = (acc / secs).cpu().numpy() if secs > 0 else np.zeros(acc.shape[0])
        if reset:
            acc.zero_()
            self._acc_seconds = 0.0
        if self.T > 1:
            x = x.reshape(self.N, self.T).T if self.trial_layout == "inner" else x.reshape(self.T, self.N)
            return x[:, self.inv]
        return x[self.inv]

    def autotune(self, **kw) -> dict:
        """Choose the block size of each fused-schedule kernel by timing (cmfgpu_amd.autotune.tune)."""
        from ..autotune import tune
        return tune(self, **kw)

    def set_schedule(self, name: str) -> None:
        """Schedule of the fused sub-step, all with the same results:
        "three_pass" (the default): an outflow pass that also leaves each
        storage limiter its own outgoing term fixes, a fix-up pass over the
        path ends and the catchments with backflow from upstream, and the
        stage pass; "limiter_pass": outflow, limiter and stage passes
        over the whole network; "block_local" (one trial, breadth-first
        layout, no reservoirs): one pass in which each block of 512
        catchments completes the limiters and stages it holds whole, then
        list passes over the rest; measured slower on MI355X, kept as an
        option."""
        if name not in _lib.SCHEDULES:
            raise ValueError(f"schedule must be one of {sorted(_lib.SCHEDULES)}")
        _lib.check(_lib.lib().cmf_network_set_schedule(self.network_handle, _lib.SCHEDULES[name]),
                   "cmf_network_set_schedule")

    def local_schedule_counts(self) -> tuple[int, int, int]:
        """(block size, catchments of the limiter list, of the stage list) of
        the block-local schedule; block size 0 where it does not apply."""
        b, l, s = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().cmf_network_local_counts(self.network_handle, ctypes.byref(b), ctypes.byref(l),
                                                       ctypes.byref(s)), "cmf_network_local_counts")
        return int(b.value), int(l.value), int(s.value)

    def group_sub_steps(self) -> np.ndarray:
        """Sub-step count of each group in the last outer step (group order = sorted labels)."""
        return self.t["max_sub_steps"].cpu().numpy()

    def set_inputs(self, runoff) -> None:
        """Stage runoff forcing (cama_flood_model.py:74-102): [N], or [T, N] when runoff is batched.

        A torch tensor is permuted into the device layout on its own device
        (an index gather on the GPU when it lives there, no host round trip);
        a numpy array is permuted on the host and copied once."""
        per_trial = "runoff" in self.batched_names
        expect = (self.T, self.N) if (per_trial and self.T > 1) else (self.N,)
        if isinstance(runoff, torch.Tensor):
            if tuple(runoff.shape) != expect:
                raise ValueError(f"runoff must have shape {expect}, got {tuple(runoff.shape)}")
            dst = self.t["runoff"]
            r = runoff.detach().to(device=dst.device, dtype=dst.dtype)
            if self._perm_dev is None:
                self._perm_dev = torch.as_tensor(self.perm, device=dst.device)
            if per_trial and self.T > 1:
                r = r.index_select(1, self._perm_dev)
                if self.trial_layout == "inner":
                    r = r.t()
            else:
                r = r.index_select(0, self._perm_dev)
            dst.copy_(r)
            return
        r = np.asarray(runoff).astype(self.real, copy=False)
        if r.shape != expect:
            raise ValueError(f"runoff must have shape {expect}, got {r.shape}")
        if per_trial and self.T > 1:
            r = r[:, self.perm]
            if self.trial_layout == "inner":
                r = r.T
        else:
            r = r[self.perm]
        self.t["runoff"].copy_(torch.from_numpy(np.ascontiguousarray(r)))

    def set_runoff_device(self, runoff_internal: torch.Tensor) -> None:
        """Copy runoff already in the device layout (for benchmarks that keep forcing on the GPU)."""
        self.t["runoff"].copy_(runoff_internal)

    def check_spec(self, spec) -> None:
        enabled = set()
        for feat, (kind, key) in spec.feature_sources.items():
            if kind == "module_enabled" and key in self.opened_modules:
                enabled.add(feat)
            elif kind == "module_flag":
                base = self.host.modules["base"]
                if getattr(base, key):
                    enabled.add(feat)
        missing = []
        for name in spec.required_fields(enabled):
            if name in LOG_PTR_NAMES:
                if "log_sums" not in self.t:
                    missing.append(name)
            elif name not in self.t:
                missing.append(name)
        if missing:
            raise ValueError(f"{spec.name}: model has no field(s) {missing}")

    # ------------------------------------------------------------ step ----
    def step_advance(self) -> int:
        """Advance one outer step; returns the number of sub-steps used."""
        if self.mode == "fused":
            count = ctypes.c_int32(0)
            fn = _lib.lib().cmf_hip_advance_graph if self.use_graph else _lib.lib().cmf_hip_advance
            _lib.check(fn(self.network_handle, self.field_table_ref(), self.step_duration,
                          self.num_sub_steps, ctypes.byref(count), _lib.current_stream()),
                       "cmf_hip_advance_graph" if self.use_graph else "cmf_hip_advance")
            n = int(count.value)
        else:
            n = self._step_ops()
        self.last_sub_steps = n
        log = self.host.modules.get("log")
        if log is not None and self.G:
            # per-group sums of this outer step: [11, G, log_buffer_size]
            self.last_group_log = self.log_sums()
            self.t["log_sums"].zero_()
        elif log is not None:
            log.set_time(self.step_duration / n, n, self.current_time)
            if self.log_path is not None or self.keep_log_sums:
                sums = self.log_sums()
                if self.log_path is not None:
                    log.write_step(self.log_path, sums)
                if self.keep_log_sums:
                    # this outer step's sums, one column per sub-step: [11, n]
                    self.last_log_sums = sums[:, :n].copy()
            self.t["log_sums"].zero_()
        if "acc_outflow" in self.t:
            self._acc_seconds += self.step_duration
        self.current_time = self.current_time + timedelta(seconds=self.step_duration)
        self.step_count += 1
        return n

    def _step_ops(self) -> int:
        """The reference's sub-step sequence, op by op (cama_flood_model.py:104-173)."""
        has = set(self.opened_modules)
        with torch.inference_mode(), bind(self):
            self.t["outer_time_step"].fill_(self.step_duration)
            if "adaptive_time" in has:
                self.t["max_sub_steps"].zero_()
                compute_adaptive_time_step()
                count = int(self.t["max_sub_steps"].max().item())
            else:
                count = self.num_sub_steps
            self.t["time_step"].fill_(self.step_duration / count)
            log = "log" in has
            for k in range(count):
                if log:
                    self.t["current_step"].fill_(k)
                compute_outflow()
                if "reservoir" in has:
                    compute_reservoir_outflow()
                if "bifurcation" in has:
                    if "levee" in has:
                        compute_levee_bifurcation_outflow()
                    else:
                        compute_bifurcation_outflow()
                compute_inflow()
                if "bifurcation" in has:
                    compute_bifurcation_inflow()
                if log:
                    compute_flood_stage_log()
                else:
                    compute_flood_stage()
                if "levee" in has:
                    if log:
                        compute_levee_stage_log()
                    else:
                        compute_levee_stage()
            if not log:
                self.t["current_step"].fill_(count - 1)
        return count

    def close(self) -> None:
        if getattr(self, "network_handle", None):
            _lib.lib().cmf_network_destroy(self.network_handle)
            self.network_handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
