"""Gradients of a gauge-discharge loss with respect to river Manning's n and
river (bank) height, by the hand-written adjoint of the fused sub-step
(cmf_hip_substep_adjoint, csrc/cmf_adjoint.hip).

Loss over S outer steps and G gauge catchments, Q = river + flood outflow of
the gauge after each outer step:

    L = 1/(S G) * sum_{s,g} (log(max(Q_sg, 0) + eps) - log(max(Q_obs_sg, 0) + eps))^2

The forward run advances the model with its own adaptive CFL sub-steps and
records the sub-step count of every outer step; the backward holds those
counts (dt = outer step / count) fixed.  The forward keeps a checkpoint of the
state every K outer steps; the backward walks the segments from the last one,
re-runs each segment sub-step by sub-step from its checkpoint keeping every
sub-step's input state on the device, then sweeps the segment in reverse.
The initial state and every other parameter are held fixed.

``GaugeLoss.apply(manning, height, problem)`` wraps this in a
torch.autograd.Function.  The base physics, the bifurcation paths (whose
per-level flows and depths join the state), the levee stage (whose protected
storage and depth join the state) and the reservoirs (whose gathered inflow
joins the state) are differentiated; trials and grouped networks are
refused.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

STATE = ("river_outflow", "flood_outflow", "river_cross_section_depth", "flood_cross_section_depth",
         "flood_cross_section_area", "river_storage", "flood_storage", "river_depth", "flood_depth")
PATH_STATE = ("bifurcation_outflow", "bifurcation_cross_section_depth")
LEVEE_STATE = ("protected_storage", "protected_depth")
RES_STATE = ("reservoir_total_inflow",)
QR, QF = 0, 1


INTERMEDIATES = ("qraw_r", "qraw_f", "lim")
# tape slot layout: REAL rows (state without storages, then intermediates), STO rows
TAPE_REAL = ("river_outflow", "flood_outflow", "river_cross_section_depth", "flood_cross_section_depth",
             "flood_cross_section_area", "river_depth", "flood_depth") + INTERMEDIATES
TAPE_STO = ("river_storage", "flood_storage")


class CmfAdjoint(ctypes.Structure):
    """Mirror of struct cmf_adjoint."""

    _fields_ = [(name, ctypes.c_void_p) for name in STATE + INTERMEDIATES] + [
        ("adj_out", ctypes.c_void_p * len(STATE)),
        ("adj_in", ctypes.c_void_p * len(STATE)),
        ("grad_river_manning", ctypes.c_void_p),
        ("grad_river_height", ctypes.c_void_p),
        ("time_step", ctypes.c_double),
        ("mode", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ] + [(name, ctypes.c_void_p) for name in PATH_STATE] + [
        ("adj_out_path", ctypes.c_void_p * 2),
        ("adj_in_path", ctypes.c_void_p * 2),
        ("path_terms", ctypes.c_void_p),
    ] + [(name, ctypes.c_void_p) for name in LEVEE_STATE] + [
        ("adj_out_levee", ctypes.c_void_p * 2),
        ("adj_in_levee", ctypes.c_void_p * 2),
        ("reservoir_total_inflow", ctypes.c_void_p),
        ("adj_out_reservoir", ctypes.c_void_p),
        ("adj_in_reservoir", ctypes.c_void_p),
    ]


ADJ_DEFER, ADJ_PENDING, ADJ_FLUSH = 1, 2, 4


class CmfTape(ctypes.Structure):
    """Mirror of struct cmf_tape."""

    _fields_ = [(name, ctypes.c_void_p) for name in STATE + INTERMEDIATES + PATH_STATE + ("path_terms",)
                + LEVEE_STATE + RES_STATE]


_SIG_DONE = False


def _declare():
    global _SIG_DONE
    if _SIG_DONE:
        return
    lib = _lib.lib()
    lib.cmf_hip_substep_adjoint.restype = ctypes.c_int
    lib.cmf_hip_substep_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.CmfFields),
                                            ctypes.POINTER(CmfAdjoint), ctypes.c_void_p]
    lib.cmf_hip_substep_taped.restype = ctypes.c_int
    lib.cmf_hip_substep_taped.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.CmfFields), ctypes.c_int32,
                                          ctypes.c_int32, ctypes.POINTER(CmfTape), ctypes.c_void_p]
    lib.cmf_hip_advance_taped.restype = ctypes.c_int
    lib.cmf_hip_advance_taped.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.CmfFields), ctypes.c_double,
                                          ctypes.c_int32, ctypes.POINTER(CmfTape), ctypes.c_void_p]
    lib.cmf_hip_adjoint_steps.restype = ctypes.c_int
    lib.cmf_hip_adjoint_steps.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.CmfFields),
                                          ctypes.POINTER(CmfAdjoint), ctypes.c_int32, ctypes.c_void_p]
    _SIG_DONE = True


class GaugeProblem:
    """A calibration problem on a built model: gauges (catchment indices in
    the model's input order, the reference's order that ``get`` returns), the
    observed discharge [S, G], the runoff of each outer step (one array, or a
    callable step -> array) and the checkpoint interval K."""

    def __init__(self, model, gauges, observed, runoff=None, checkpoint_every: int = 48, eps: float = 1e-3):
        extra = set(model.opened_modules) - {"base", "adaptive_time", "bifurcation", "levee", "reservoir"}
        if extra:
            raise ValueError(f"the adjoint covers the base physics, bifurcation, levee and reservoir; modules "
                             f"{sorted(extra)} are open")
        if model.T != 1 or model.G or model.batched_names or model.mode != "fused":
            raise ValueError("the adjoint needs one trial, an ungrouped network, no per-trial parameters and "
                             "the fused schedule")
        # the adjoint takes the protected storage of a catchment without a
        # levee as zero, which is what every stage pass leaves there
        ps = model.t.get("protected_storage")
        if ps is not None:
            lev = model.t.get("is_levee") if "levee" in model.opened_modules else None
            outside = ps if lev is None else ps[~lev.bool()]
            if bool((outside != 0).any()):
                raise ValueError("the adjoint needs zero protected storage outside levee catchments in the "
                                 "initial state")
        self.model = model
        self.paths = "bifurcation" in model.opened_modules and model.P > 0
        self.levees = "levee" in model.opened_modules
        self.reservoirs = "reservoir" in model.opened_modules
        self.gauges = np.asarray(gauges, dtype=np.int64)
        self.obs = torch.as_tensor(np.asarray(observed, dtype=np.float64), device=model.device)
        if self.obs.ndim != 2 or self.obs.shape[1] != self.gauges.shape[0]:
            raise ValueError("observed must be [steps, gauges]")
        self.steps = int(self.obs.shape[0])
        self.K = int(checkpoint_every)
        self.eps = float(eps)
        self.runoff = runoff
        self.g_dev = torch.as_tensor(model.inv[self.gauges], device=model.device)
        self.initial = self._snapshot()
        self.counts: list[int] = []
        self.checkpoints: dict[int, list] = {}
        self.last_loss = None
        _declare()

    # ------------------------------------------------------------ state --
    def _state(self):
        names = (STATE + (PATH_STATE if self.paths else ()) + (LEVEE_STATE if self.levees else ())
                 + (RES_STATE if self.reservoirs else ()))
        return [self.model.t[name] for name in names]

    def _snapshot(self):
        return [t.clone() for t in self._state()]

    def _restore(self, snap):
        for t, s in zip(self._state(), snap):
            t.copy_(s)

    def _apply_runoff(self, step: int):
        if self.runoff is None:
            return
        r = self.runoff(step) if callable(self.runoff) else self.runoff
        if r is not None:
            self.model.set_inputs(r)

    def set_parameters(self, manning=None, height=None):
        """Per-catchment river Manning's n and river height, in the model's
        input order (the reference's catchment order, as ``get`` returns)."""
        m = self.model
        perm = torch.as_tensor(m.perm, device=m.device)
        for name, x in (("river_manning", manning), ("river_height", height)):
            if x is not None:
                x = torch.as_tensor(x).to(m.device)
                m.t[name].copy_(x[perm].to(m.t[name].dtype))

    def parameters(self):
        m = self.model
        return m.get("river_manning"), m.get("river_height")

    # ------------------------------------------------------------- loss --
    def _gauge_q(self):
        t = self.model.t
        return (t["river_outflow"][self.g_dev].double() + t["flood_outflow"][self.g_dev].double())

    def _loss_term(self, step, q):
        qc = q.clamp_min(0.0)
        return ((torch.log(qc + self.eps) - self._log_obs(step)) ** 2).sum()

    def _log_obs(self, step):
        return torch.log(self.obs[step].clamp_min(0.0) + self.eps)

    def _loss_seeds(self, q):
        """dL/dQ of every outer step and gauge at once ([S, G], from the
        forward's gauge discharge q)."""
        qc = q.clamp_min(0.0)
        r = torch.log(qc + self.eps) - torch.log(self.obs.clamp_min(0.0) + self.eps)
        return torch.where(q > 0, 2.0 * r / (qc + self.eps), torch.zeros_like(q)) / (self.steps * self.obs.shape[1])

    def _substep(self, k: int, last: bool):
        m = self.model
        _lib.check(_lib.lib().cmf_hip_substep(m.network_handle, m.field_table_ref(), k, 1 if last else 0,
                                              _lib.current_stream()), "cmf_hip_substep")

    def _step_fixed(self, count: int):
        m = self.model
        m.t["time_step"].fill_(m.step_duration / count)
        for k in range(count):
            self._substep(k, k == count - 1)

    def run(self, fixed_counts=None) -> float:
        """Forward from the initial state; returns the loss.  With
        `fixed_counts` the sub-step counts are prescribed (dt held)."""
        self._restore(self.initial)
        self.counts = []
        self.checkpoints = {}
        total = torch.zeros((), dtype=torch.float64, device=self.model.device)
        # the gauge discharge after every outer step, for the backward's seeds
        self.q_hist = torch.empty(self.obs.shape, dtype=torch.float64, device=self.model.device)
        for step in range(self.steps):
            if step % self.K == 0:
                self.checkpoints[step] = self._snapshot()
            self._apply_runoff(step)
            if fixed_counts is None:
                n = self.model.step_advance()
            else:
                n = int(fixed_counts[step])
                self._step_fixed(n)
            self.counts.append(n)
            q = self._gauge_q()
            self.q_hist[step] = q
            total += self._loss_term(step, q)
        self.final = self._snapshot()
        self.last_loss = float(total.item()) / (self.steps * self.obs.shape[1])
        return self.last_loss

    def gradient(self):
        """dL/d(river_manning), dL/d(river_height) (float64 device tensors, in
        the model's input order) for the last `run`."""
        if not self.counts:
            raise RuntimeError("run() first")
        m = self.model
        dev = m.device
        n = m.N
        real_t = m.t["river_outflow"].dtype
        sto_t = m.t["river_storage"].dtype
        lam = [[torch.zeros(n, dtype=real_t, device=dev) for _ in STATE] for _ in range(2)]
        pe = m.P * m.t["bifurcation_outflow"].shape[-1] if self.paths else 0
        lam_p = [[torch.zeros(pe, dtype=real_t, device=dev) for _ in PATH_STATE] for _ in range(2)]
        lam_l = [[torch.zeros(n if self.levees else 0, dtype=real_t, device=dev) for _ in LEVEE_STATE]
                 for _ in range(2)]
        lam_r = [torch.zeros(n if self.reservoirs else 0, dtype=real_t, device=dev) for _ in range(2)]
        g_n = torch.zeros(n, dtype=torch.float64, device=dev)
        g_h = torch.zeros(n, dtype=torch.float64, device=dev)
        starts = sorted(self.checkpoints)
        slots = max(sum(self.counts[s0:min(s0 + self.K, self.steps)]) for s0 in starts)
        tape_r = torch.empty((slots, len(TAPE_REAL), n), dtype=real_t, device=dev)
        tape_s = torch.empty((slots, len(TAPE_STO), n), dtype=sto_t, device=dev)
        tape_p = torch.empty((slots, len(PATH_STATE), pe), dtype=real_t, device=dev) if self.paths else None
        # the path intermediates of each sub-step: limited level flows, level
        # sums, outgoing terms at the two ends
        tape_pt = torch.empty((slots, pe + 3 * m.P), dtype=real_t, device=dev) if self.paths else None
        # with levees the input protected storage (STO) and depth (REAL)
        tape_ls = torch.empty((slots, n), dtype=sto_t, device=dev) if self.levees else None
        tape_ld = torch.empty((slots, n), dtype=real_t, device=dev) if self.levees else None
        # with reservoirs the inflow each gathered in the previous sub-step (STO)
        tape_ri = torch.empty((slots, n), dtype=sto_t, device=dev) if self.reservoirs else None
        tapes, adjs = [], []
        for k in range(slots):
            ptrs = {name: tape_r[k, x].data_ptr() for x, name in enumerate(TAPE_REAL)}
            ptrs.update({name: tape_s[k, x].data_ptr() for x, name in enumerate(TAPE_STO)})
            if self.paths:
                ptrs.update({name: tape_p[k, x].data_ptr() for x, name in enumerate(PATH_STATE)})
                ptrs["path_terms"] = tape_pt[k].data_ptr()
            if self.levees:
                ptrs["protected_storage"] = tape_ls[k].data_ptr()
                ptrs["protected_depth"] = tape_ld[k].data_ptr()
            if self.reservoirs:
                ptrs["reservoir_total_inflow"] = tape_ri[k].data_ptr()
            tp = CmfTape(**ptrs)
            pair = []
            for parity in range(2):
                a = CmfAdjoint(**ptrs)
                for x in range(len(STATE)):
                    a.adj_out[x] = lam[parity][x].data_ptr()
                    a.adj_in[x] = lam[1 - parity][x].data_ptr()
                if self.paths:
                    for x in range(len(PATH_STATE)):
                        a.adj_out_path[x] = lam_p[parity][x].data_ptr()
                        a.adj_in_path[x] = lam_p[1 - parity][x].data_ptr()
                if self.levees:
                    for x in range(len(LEVEE_STATE)):
                        a.adj_out_levee[x] = lam_l[parity][x].data_ptr()
                        a.adj_in_levee[x] = lam_l[1 - parity][x].data_ptr()
                if self.reservoirs:
                    a.adj_out_reservoir = lam_r[parity].data_ptr()
                    a.adj_in_reservoir = lam_r[1 - parity].data_ptr()
                a.grad_river_manning = g_n.data_ptr()
                a.grad_river_height = g_h.data_ptr()
                pair.append(a)
            tapes.append(tp)
            adjs.append(pair)
        lib = _lib.lib()
        # the taped re-runs reproduce the forward bit for bit, so the seeds of
        # every outer step come from the forward's gauge discharge at once
        seeds = self._loss_seeds(self.q_hist).to(real_t)
        parity = 0
        first_done = False
        for s0 in reversed(starts):
            s1 = min(s0 + self.K, self.steps)
            self._restore(self.checkpoints[s0])
            marks = []
            slot = 0
            for step in range(s0, s1):
                self._apply_runoff(step)
                cnt = self.counts[step]
                arr = (CmfTape * cnt)(*tapes[slot:slot + cnt])
                _lib.check(lib.cmf_hip_advance_taped(m.network_handle, m.field_table_ref(), m.step_duration, cnt,
                                                     arr, _lib.current_stream()), "cmf_hip_advance_taped")
                slot += cnt
                marks.append((step, slot))
            for step, end in reversed(marks):
                cnt = self.counts[step]
                seed = seeds[step]
                lam[parity][QR].index_add_(0, self.g_dev, seed)
                lam[parity][QF].index_add_(0, self.g_dev, seed)
                self._apply_runoff(step)
                dt = m.step_duration / cnt
                # sub-step k of this outer step runs after the (cnt - 1 - k) later ones
                arr = (CmfAdjoint * cnt)(*[adjs[end - cnt + k][(parity + cnt - 1 - k) % 2] for k in range(cnt)])
                for k in range(cnt):
                    arr[k].time_step = dt
                    # the gather of each call is folded into the next one
                    arr[k].mode = ADJ_DEFER | (ADJ_PENDING if (first_done or k < cnt - 1) else 0)
                first_done = True
                _lib.check(lib.cmf_hip_adjoint_steps(m.network_handle, m.field_table_ref(), arr, cnt,
                                                     _lib.current_stream()), "cmf_hip_adjoint_steps")
                parity = (parity + cnt) % 2
        # the deferred gather of the last call (the first sub-step)
        fl = CmfAdjoint()
        ctypes.memmove(ctypes.byref(fl), ctypes.byref(adjs[0][(parity + 1) % 2]), ctypes.sizeof(CmfAdjoint))
        fl.mode = ADJ_FLUSH
        _lib.check(lib.cmf_hip_substep_adjoint(m.network_handle, m.field_table_ref(), ctypes.byref(fl),
                                               _lib.current_stream()), "cmf_hip_substep_adjoint")
        del tape_r, tape_s, tape_p, tape_pt, tape_ls, tape_ld, tape_ri
        self._restore(self.final)
        inv = torch.as_tensor(m.inv, device=dev)
        return g_n[inv], g_h[inv]


class GaugeLoss(torch.autograd.Function):
    """loss = GaugeLoss.apply(manning, height, problem): the gauge-discharge
    loss of `problem` (a float64 scalar on the inputs' device) as a function
    of per-catchment river Manning's n and river height, both in the model's
    input order; backward runs the adjoint and returns each gradient in its
    input's dtype and on its input's device."""

    @staticmethod
    def forward(ctx, manning, height, problem):
        problem.set_parameters(manning.detach(), height.detach())
        loss = problem.run()
        ctx.problem = problem
        ctx.meta = ((manning.dtype, manning.device), (height.dtype, height.device))
        return torch.tensor(loss, dtype=torch.float64, device=manning.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_n, g_h = ctx.problem.gradient()
        (dn, devn), (dh, devh) = ctx.meta
        gn = (grad_out.to(g_n.device) * g_n).to(device=devn, dtype=dn)
        gh = (grad_out.to(g_h.device) * g_h).to(device=devh, dtype=dh)
        return gn, gh, None
