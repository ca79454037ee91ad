"""Hand-written adjoint of the fused sub-step (cmf_hip_substep_adjoint) against
FP64 torch autograd through a plain PyTorch-op transcription of the reference
step (oracle/torch_step.py) and against central finite differences.

Problem: plain network (3,000 catchments; the autograd check on a
20,000-catchment network), 16 gauges (largest upstream area), loss = mean
squared log-discharge error over 8 outer steps against a run whose Manning's
n and river height were scaled by lognormal(0, 0.2) factors; checkpoints
every 3 outer steps (so the backward re-runs segments).

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


def _setup(precision, n=3000):
    from cmfgpu_amd import synthetic
    from cmfgpu_amd.adjoint import GaugeProblem
    from cmfgpu_amd.models import CaMaFlood
    net = synthetic.generate_network(n, 1003)
    src = dict(net.fields)
    src["runoff"] = synthetic.Runoff(src["catchment_area"], 2003).next_day() * np.float32(4.0)
    gauges = np.argsort(net.upstream_area)[-16:]
    model = CaMaFlood(src, opened_modules=PLAIN, device="cuda:0", step_duration=1800.0, precision=precision)
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


def test_adjoint_matches_fp64_autograd():
    net, src, model, problem, gauges = _setup("float64", n=20000)
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


def test_adjoint_against_central_differences():
    net, src, model, problem, gauges = _setup("float64")
    problem.run()
    counts = list(problem.counts)
    g_n, g_h = (x.cpu().numpy() for x in problem.gradient())
    n0, h0 = (x.astype(np.float64) for x in problem.parameters())
    rng = np.random.default_rng(11)
    cand_n = np.flatnonzero(np.abs(g_n) > 1e-4 * np.abs(g_n).max())
    cand_h = np.flatnonzero(np.abs(g_h) > 1e-4 * np.abs(g_h).max())
    checks = [("n", i) for i in rng.choice(cand_n, min(32, cand_n.size), replace=False)] + \
             [("h", i) for i in rng.choice(cand_h, min(32, cand_h.size), replace=False)]
    assert len(checks) == 64, (cand_n.size, cand_h.size)
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


def test_fp32_adjoint_close_to_fp64():
    grads = {}
    for prec in ("float64", "float32"):
        net, src, model, problem, gauges = _setup(prec)
        problem.run()
        grads[prec] = [x.cpu().numpy() for x in problem.gradient()]
        model.close()
    for a, b in zip(grads["float32"], grads["float64"]):
        print(f"FP32 adjoint vs FP64 adjoint: {_rel(a, b):.2e}", flush=True)
        assert _rel(a, b) <= 5e-2, _rel(a, b)
