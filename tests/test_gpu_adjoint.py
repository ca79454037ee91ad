"""Hand-written adjoint of the fused sub-step (cmf_hip_substep_adjoint) against
FP64 torch autograd through a plain PyTorch-op transcription of the reference
step (oracle/torch_step.py) and against central finite differences.

Problem: plain network (3,000 catchments; the autograd check on a
20,000-catchment network), 16 gauges (largest upstream area), loss = mean
squared log-discharge error over 8 outer steps against a run whose Manning's
n and river height were scaled by lognormal(0, 0.2) factors; checkpoints
every 3 outer steps (so the backward re-runs segments).  The same checks run
with bifurcation paths open (100 paths on 3,000 catchments, 600 on 20,000),
with levees on 30 % of the catchments (their initial river storage scaled
by U(1, 100), so that every levee case occurs), with both, and with both plus
reservoirs (one per 50 catchments, each started at U(0.1, 1.1) of its
capacity so that every release regime occurs); the central differences then
take half their parameters at path ends, levees, reservoirs and the
catchments draining into them.

Tolerances: FP64 adjoint vs FP64 autograd, relative L2 of each gradient 1e-6
(the two sum floats in different orders); central differences (h = 1e-3
relative, sub-step counts held) on 64 random parameters (32 n, 32 H, among
those whose gradient exceeds 1e-4 of the largest): relative 2e-2 for every
component where the loss is smooth inside +-h (the forward and backward
one-sided quotients agree within 10 %; where a branch of the step, such as a
wet/dry or floodplain-level switch, flips inside +-h the quotient is no
derivative), at most 8 of the 64 at such kinks, median relative error 1e-3
over all 64; the FP32 adjoint against the FP64 one, relative L2 5e-2.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PLAIN = ("base", "adaptive_time")
STEPS = 8


KINDS = {"plain": (0, 0.0, 0), "paths": (1, 0.0, 0), "levees": (0, 0.3, 0), "both": (1, 0.3, 0),
         "all": (1, 0.3, 1)}


def _setup(precision, n=3000, kind="plain", precise=False):
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.adjoint import GaugeProblem
    from cmfgpu_amd.models import CaMaFlood
    has_paths, levee, has_res = KINDS[kind]
    paths = (n // 30 if n <= 3000 else n * 3 // 100) if has_paths else 0
    net = synthetic.generate_network(n, 1003, paths=paths, levee=levee, reservoirs=n // 50 if has_res else 0)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    if levee:
        rs = np.array(src["river_storage"], dtype=np.float64)
        lid = np.asarray(src["levee_catchment_idx"])
        rs[lid] *= np.random.default_rng(3).uniform(1.0, 100.0, lid.size)
        src["river_storage"] = rs.astype(np.asarray(src["river_storage"]).dtype)
    if has_res:
        rs = np.array(src["river_storage"], dtype=np.float64)
        rid = np.asarray(src["reservoir_catchment_idx"])
        cap = np.asarray(src["reservoir_capacity"], dtype=np.float64)
        rs[rid] = cap * np.random.default_rng(5).uniform(0.1, 1.1, rid.size)
        src["river_storage"] = rs.astype(np.asarray(src["river_storage"]).dtype)
    gauges = np.argsort(net.upstream_area)[-16:]
    mods = (PLAIN + (("bifurcation",) if paths else ()) + (("levee",) if levee else ())
            + (("reservoir",) if has_res else ()))
    model = CaMaFlood(src, opened_modules=mods, device="cuda:0", step_duration=1800.0, precision=precision,
                      precise_surfaces=precise)
    rng = np.random.default_rng(7)
    n0, h0 = src["river_manning"].astype(np.float64), src["river_height"].astype(np.float64)
    n_true = n0 * rng.lognormal(0.0, 0.2, n0.shape[0])
    h_true = h0 * rng.lognormal(0.0, 0.2, h0.shape[0])
    problem = GaugeProblem(model, gauges, np.ones((STEPS, 16)), checkpoint_every=3)
    # observations: the discharge of a run with the perturbed parameters
    problem.set_parameters(n_true, h_true)
    obs = []
    for _ in range(STEPS):
        model.step_advance()
        obs.append(problem._gauge_q().cpu().numpy())
    problem.set_parameters(n0, h0)
    problem._restore(problem.initial)
    problem.obs = torch.as_tensor(np.array(obs), dtype=torch.float64, device=model.device)
    return net, src, model, problem, gauges


def _torch_gradient(model, problem, gauges, n0, h0):
    from oracle.torch_step import substep
    problem._restore(problem.initial)
    st = {k: torch.tensor(model.get(k), dtype=torch.float64) for k in (
        "river_outflow", "flood_outflow", "river_cross_section_depth", "flood_cross_section_depth",
        "flood_cross_section_area", "river_storage", "flood_storage", "river_depth", "flood_depth")}
    prm = {k: torch.tensor(model.get(k), dtype=torch.float64) for k in (
        "flood_manning", "river_width", "river_length", "catchment_elevation", "catchment_area",
        "downstream_distance", "runoff", "flood_depth_table")}
    prm["downstream_idx"] = torch.tensor(model.get("downstream_idx"), dtype=torch.int64)
    if problem.reservoirs:
        st["reservoir_total_inflow"] = torch.tensor(model.get("reservoir_total_inflow"), dtype=torch.float64)
        for k in ("conservation_volume", "emergency_volume", "adjustment_volume", "effective_normal_outflow",
                  "adjustment_outflow", "flood_control_outflow"):
            prm[k] = torch.tensor(model.get(k), dtype=torch.float64)
        prm["reservoir_catchment_idx"] = torch.tensor(model.get("reservoir_catchment_idx"), dtype=torch.int64)
        for k in ("is_dam_upstream", "is_reservoir"):
            prm[k] = torch.tensor(model.get(k), dtype=torch.bool)
    if problem.levees:
        for k in ("protected_storage", "protected_depth"):
            st[k] = torch.tensor(model.get(k), dtype=torch.float64)
        for k in ("levee_crown_height", "levee_fraction", "levee_base_height"):
            prm[k] = torch.tensor(model.get(k), dtype=torch.float64)
        prm["levee_catchment_idx"] = torch.tensor(model.get("levee_catchment_idx"), dtype=torch.int64)
    if problem.paths:
        for k in ("bifurcation_outflow", "bifurcation_cross_section_depth"):
            st[k] = torch.tensor(model.get(k), dtype=torch.float64)
        for k in ("bifurcation_length", "bifurcation_elevation", "bifurcation_width", "bifurcation_manning"):
            prm[k] = torch.tensor(model.get(k), dtype=torch.float64)
        for k in ("bifurcation_catchment_idx", "bifurcation_downstream_idx"):
            prm[k] = torch.tensor(model.get(k), dtype=torch.int64)
    n = torch.tensor(n0, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(h0, dtype=torch.float64, requires_grad=True)
    prm["river_manning"], prm["river_height"] = n, h
    obs = problem.obs.cpu()
    g = torch.tensor(gauges)
    loss = torch.zeros((), dtype=torch.float64)
    for step, cnt in enumerate(problem.counts):
        dt = 1800.0 / cnt
        for _ in range(cnt):
            st = substep(st, prm, dt)
        q = st["river_outflow"][g] + st["flood_outflow"][g]
        qc = torch.where(q > 0, q, torch.zeros_like(q))
        lo = torch.log(obs[step].clamp_min(0.0) + problem.eps)
        loss = loss + ((torch.log(qc + problem.eps) - lo) ** 2).sum()
    loss = loss / (len(problem.counts) * g.shape[0])
    loss.backward()
    return float(loss), n.grad.numpy(), h.grad.numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("kind", list(KINDS))
def test_adjoint_matches_fp64_autograd(kind):
    net, src, model, problem, gauges = _setup("float64", n=20000, kind=kind)
    loss = problem.run()
    assert loss > 0 and len(problem.checkpoints) == 3
    g_n, g_h = (x.cpu().numpy() for x in problem.gradient())
    n0, h0 = (x.astype(np.float64) for x in problem.parameters())
    t_loss, t_n, t_h = _torch_gradient(model, problem, gauges, n0, h0)
    assert abs(t_loss - loss) <= 1e-9 * abs(loss)
    assert np.count_nonzero(t_n) > 50
    print(f"adjoint vs FP64 autograd: manning {_rel(g_n, t_n):.2e}, height {_rel(g_h, t_h):.2e}", flush=True)
    assert _rel(g_n, t_n) <= 1e-6, _rel(g_n, t_n)
    assert _rel(g_h, t_h) <= 1e-6, _rel(g_h, t_h)
    model.close()


def _pick_checks(rng, g, ends, count):
    """`count` indices whose gradient exceeds 1e-4 of the largest; with
    `ends` given (path ends, levee catchments), half of them among those."""
    cand = np.flatnonzero(np.abs(g) > 1e-4 * np.abs(g).max())
    if ends is None:
        return list(rng.choice(cand, min(count, cand.size), replace=False))
    at_ends = np.intersect1d(cand, ends)
    first = rng.choice(at_ends, min(count // 2, at_ends.size), replace=False)
    rest = np.setdiff1d(cand, first)
    return list(first) + list(rng.choice(rest, count - first.size, replace=False))


@pytest.mark.parametrize("kind", list(KINDS))
def test_adjoint_against_central_differences(kind):
    net, src, model, problem, gauges = _setup("float64", kind=kind)
    problem.run()
    counts = list(problem.counts)
    g_n, g_h = (x.cpu().numpy() for x in problem.gradient())
    n0, h0 = (x.astype(np.float64) for x in problem.parameters())
    rng = np.random.default_rng(11)
    ends = None
    if kind != "plain":
        ends = np.zeros(0, np.int64)
        if "bifurcation_catchment_idx" in src:
            ends = np.union1d(src["bifurcation_catchment_idx"], src["bifurcation_downstream_idx"])
        if "levee_catchment_idx" in src:
            ends = np.union1d(ends, src["levee_catchment_idx"])
        if "reservoir_catchment_idx" in src:
            rid = np.asarray(src["reservoir_catchment_idx"])
            ends = np.union1d(ends, rid)
            ends = np.union1d(ends, np.flatnonzero(np.isin(np.asarray(src["downstream_idx"]), rid)))
    checks = [("n", i) for i in _pick_checks(rng, g_n, ends, 32)] + \
             [("h", i) for i in _pick_checks(rng, g_h, ends, 32)]
    assert len(checks) == 64, len(checks)
    loss0 = problem.run(fixed_counts=counts)
    rel, kink = [], []
    for kind, i in checks:
        base = n0 if kind == "n" else h0
        step = 1e-3 * base[i]
        vals = []
        for sgn in (1.0, -1.0):
            x = base.copy()
            x[i] += sgn * step
            if kind == "n":
                problem.set_parameters(x, h0)
            else:
                problem.set_parameters(n0, x)
            vals.append(problem.run(fixed_counts=counts))
        problem.set_parameters(n0, h0)
        fd = (vals[0] - vals[1]) / (2.0 * step)
        ad = (g_n if kind == "n" else g_h)[i]
        # one-sided quotients: disagreeing by more than 10 % marks a kink of
        # the loss inside +-h (a branch of the step switching there)
        fwd, bwd = (vals[0] - loss0) / step, (loss0 - vals[1]) / step
        kink.append(abs(fwd - bwd) > 0.1 * max(abs(fwd), abs(bwd)))
        rel.append(abs(fd - ad) / abs(ad))
        print(f"central difference {kind}[{i}]: fd {fd:.6e} adjoint {ad:.6e} rel {rel[-1]:.2e}"
              f"{' (kink inside +-h)' if kink[-1] else ''}", flush=True)
    rel, kink = np.array(rel), np.array(kink)
    print(f"central differences: median rel {np.median(rel):.2e}, max over smooth points "
          f"{rel[~kink].max():.2e}, {int(kink.sum())} of {rel.size} at kinks", flush=True)
    assert np.all(rel[~kink] <= 2e-2), rel[~kink]
    assert int(kink.sum()) <= 8, int(kink.sum())
    assert np.median(rel) <= 1e-3, np.median(rel)
    model.close()


@pytest.mark.parametrize("kind", ["plain", "all"])
def test_fp32_adjoint_close_to_fp64(kind):
    grads = {}
    for prec in ("float64", "float32"):
        net, src, model, problem, gauges = _setup(prec, kind=kind)
        problem.run()
        grads[prec] = [x.cpu().numpy() for x in problem.gradient()]
        model.close()
    for a, b in zip(grads["float32"], grads["float64"]):
        print(f"FP32 adjoint vs FP64 adjoint: {_rel(a, b):.2e}", flush=True)
        assert _rel(a, b) <= 5e-2, _rel(a, b)


@pytest.mark.parametrize("kind", ["plain", "both"])
def test_precise_fp32_adjoint_close_to_fp64(kind):
    """The FP32 adjoint with water surfaces formed in double (precise_surfaces)
    against the FP64 one: as close as the FP32 adjoint with the reference's
    FP32 surfaces (within three times its error), and within the FP32 bound
    above.  Over these 8 outer steps both errors come from the FP32 arithmetic
    of the reverse sweep itself (about 3e-6 on the plain network), not from
    the forward trajectory that precise surfaces bring closer to FP64."""
    grads = {}
    for name, prec, precise in (("fp64", "float64", False), ("fp32", "float32", False),
                                ("precise", "float32", True)):
        net, src, model, problem, gauges = _setup(prec, kind=kind, precise=precise)
        problem.run()
        grads[name] = [x.cpu().numpy() for x in problem.gradient()]
        model.close()
    for k, label in ((0, "n"), (1, "H")):
        e32 = _rel(grads["fp32"][k], grads["fp64"][k])
        ep = _rel(grads["precise"][k], grads["fp64"][k])
        print(f"{kind} d/d{label}: FP32 {e32:.2e}, FP32 precise surfaces {ep:.2e} against FP64", flush=True)
        assert ep <= 3.0 * max(e32, 1e-6) and ep <= 5e-2, (e32, ep)


@pytest.mark.parametrize("precision", ["float32", "float64", "float32-precise"])
def test_recomputed_intermediates_match_the_tape(precision):
    """One sub-step's adjoint from the taped forward intermediates (unlimited
    flows, limiters, path terms) and again with them recomputed from the saved
    state (passes A0, A1, B0): bit-identical adjoints and gradients, so the
    reverse re-forms the forward's values exactly (network with paths and
    levees; "float32-precise": FP32 with water surfaces formed in double,
    which the reverse re-forms the same way)."""
    import ctypes
    from cmfgpu_amd import _lib
    from cmfgpu_amd.adjoint import (CmfAdjoint, CmfTape, LEVEE_STATE, PATH_STATE, STATE, TAPE_REAL, TAPE_STO,
                                    _declare)
    precise = precision.endswith("-precise")
    net, src, model, problem, gauges = _setup(precision.split("-")[0], kind="both", precise=precise)
    assert model.precise_surfaces == precise
    for _ in range(3):
        model.step_advance()
    _declare()
    lib = _lib.lib()
    dev, n, P = model.device, model.N, model.P
    real_t, sto_t = model.t["river_outflow"].dtype, model.t["river_storage"].dtype
    pe = P * model.t["bifurcation_outflow"].shape[-1]
    tr = torch.empty((len(TAPE_REAL), n), dtype=real_t, device=dev)
    ts = torch.empty((len(TAPE_STO), n), dtype=sto_t, device=dev)
    tp = torch.empty((len(PATH_STATE), pe), dtype=real_t, device=dev)
    tt = torch.empty(pe + 3 * P, dtype=real_t, device=dev)
    ptrs = {name: tr[x].data_ptr() for x, name in enumerate(TAPE_REAL)}
    ptrs.update({name: ts[x].data_ptr() for x, name in enumerate(TAPE_STO)})
    ptrs.update({name: tp[x].data_ptr() for x, name in enumerate(PATH_STATE)})
    ptrs["path_terms"] = tt.data_ptr()
    tls = torch.empty(n, dtype=sto_t, device=dev)
    tld = torch.empty(n, dtype=real_t, device=dev)
    ptrs["protected_storage"], ptrs["protected_depth"] = tls.data_ptr(), tld.data_ptr()
    count = 6
    model.t["time_step"].fill_(model.step_duration / count)
    _lib.check(lib.cmf_hip_substep_taped(model.network_handle, model.field_table_ref(), 2, 0,
                                         ctypes.byref(CmfTape(**ptrs)), _lib.current_stream()), "taped")
    gen = torch.Generator(device="cpu").manual_seed(5)
    lout = [torch.randn(n, generator=gen, dtype=torch.float64).to(dev, real_t) for _ in STATE]
    lout_p = [torch.randn(pe, generator=gen, dtype=torch.float64).to(dev, real_t) for _ in PATH_STATE]
    lout_l = [torch.randn(n, generator=gen, dtype=torch.float64).to(dev, real_t) for _ in LEVEE_STATE]
    results = []
    for taped in (True, False):
        lin = [torch.zeros(n, dtype=real_t, device=dev) for _ in STATE]
        lin_p = [torch.zeros(pe, dtype=real_t, device=dev) for _ in PATH_STATE]
        lin_l = [torch.zeros(n, dtype=real_t, device=dev) for _ in LEVEE_STATE]
        g_n = torch.zeros(n, dtype=torch.float64, device=dev)
        g_h = torch.zeros(n, dtype=torch.float64, device=dev)
        p = dict(ptrs)
        if not taped:
            for k in ("qraw_r", "qraw_f", "lim", "path_terms"):
                p[k] = None
        a = CmfAdjoint(**p)
        for x in range(len(STATE)):
            a.adj_out[x] = lout[x].data_ptr()
            a.adj_in[x] = lin[x].data_ptr()
        for x in range(len(PATH_STATE)):
            a.adj_out_path[x] = lout_p[x].data_ptr()
            a.adj_in_path[x] = lin_p[x].data_ptr()
        for x in range(len(LEVEE_STATE)):
            a.adj_out_levee[x] = lout_l[x].data_ptr()
            a.adj_in_levee[x] = lin_l[x].data_ptr()
        a.grad_river_manning, a.grad_river_height = g_n.data_ptr(), g_h.data_ptr()
        a.time_step = model.step_duration / count
        _lib.check(lib.cmf_hip_substep_adjoint(model.network_handle, model.field_table_ref(), ctypes.byref(a),
                                               _lib.current_stream()), "adjoint")
        results.append([x.cpu().numpy() for x in lin + lin_p + lin_l + [g_n, g_h]])
    for x, y in zip(*results):
        assert np.array_equal(x, y)
    assert np.count_nonzero(results[0][len(STATE)]) > 10   # the path flows' adjoint
    model.close()


@pytest.mark.parametrize("kind", ["plain", "all"])
def test_gauge_loss_function_matches_the_problem_gradient(kind):
    """GaugeLoss.apply, the torch.autograd.Function the adjoint is delivered
    through: on float32 parameter tensors in the model's input order (the
    basin layout, where the device order is another permutation), the loss
    is problem.run()'s and torch.autograd.grad returns problem.gradient(),
    as float32 tensors on the parameters' device, scaled by the incoming
    gradient."""
    from cmfgpu_amd.adjoint import GaugeLoss
    net, src, model, problem, gauges = _setup("float32", kind=kind)
    assert not np.array_equal(model.perm, np.arange(model.N))
    n0 = torch.tensor(src["river_manning"], dtype=torch.float32, device="cuda:0", requires_grad=True)
    h0 = torch.tensor(src["river_height"], dtype=torch.float32, device="cuda:0", requires_grad=True)
    loss = GaugeLoss.apply(n0, h0, problem)
    assert loss.dtype == torch.float64 and loss.device == n0.device
    gn, gh = torch.autograd.grad(3.0 * loss, (n0, h0))
    assert gn.dtype == torch.float32 and gh.dtype == torch.float32 and gn.device == n0.device
    ref_loss = problem.run()
    ref_n, ref_h = problem.gradient()
    assert float(loss) == ref_loss
    want_n = (3.0 * ref_n).to(torch.float32)
    want_h = (3.0 * ref_h).to(torch.float32)
    assert torch.equal(gn, want_n) and torch.equal(gh, want_h)
    assert float(want_n.abs().max()) > 0 and float(want_h.abs().max()) > 0
    # the same parameters permuted into device order would give another loss:
    # the function takes them in input order
    p = torch.as_tensor(model.perm, device="cuda:0")
    other = GaugeLoss.apply(n0.detach()[p], h0.detach()[p], problem)
    assert float(other) != float(loss)
    model.close()
