"""The whole step (INT8-split rho1, INT8-split H1 adjoint, far field) captured once in a HIP graph and
replayed: every replay, with eager work and synchronisations between replays, must give the eager
results bit for bit (the INT8 adjoint sums in 64-bit fixed point, so its atomics commute); the FP64
adjoint, whose FP64 atomics land in a run-dependent order, to 1e-13."""
import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import api, synth

pytestmark = pytest.mark.gpu


def test_step_graph_replays_match_eager(cuda):
    prob = synth.make_rho1_problem(16, 12.0, 150, (30, 200), (20, 230), config_k=12)
    fp = synth.make_far_problem(16, 12.0, 5000, 6, config_k=12, xyz=prob.atom_xyz)
    plan = api.Rho1Plan.from_problem(prob)
    fplan = api.FarPlan.from_problem(fp)
    finp = api.FarInputs(fp)
    wave = api.to_device(prob.wave)
    P1 = api.to_device(synth.make_p1(prob.atom_xyz, prob.basis, 3012))
    g = api.to_device(np.random.default_rng(1).standard_normal(prob.n_full_points))
    rho = torch.empty(prob.n_full_points, dtype=torch.float64, device="cuda")
    H = torch.empty(prob.n_basis * prob.n_basis, dtype=torch.float64, device="cuda")
    H64 = torch.empty(prob.n_basis * prob.n_basis, dtype=torch.float64, device="cuda")
    dv = torch.empty(fp.n_full_points, dtype=torch.float64, device="cuda")

    def step():
        plan.forward(wave, P1, precision=api.OZAKI, out=rho)
        H.zero_()
        plan.adjoint(wave, g, H=H, precision=api.OZAKI)
        H64.zero_()
        plan.adjoint(wave, g, H=H64, precision=api.FP64)
        dv.zero_()
        fplan.forward(dv, *finp.args(), variant=api.FAR_HORNER)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()  # first calls build the per-plan device state outside the capture
        step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    ref = [rho.clone(), H.clone(), dv.clone(), H64.clone()]
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for rep in range(4):
        rho.fill_(np.nan)
        H.fill_(np.nan)
        H64.fill_(np.nan)
        dv.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(rho, ref[0]), f"replay {rep}: rho1 differs"
        assert torch.equal(dv, ref[2]), f"replay {rep}: far field differs"
        assert torch.equal(H, ref[1]), f"replay {rep}: INT8 adjoint differs"
        hmax = float(ref[3].abs().max())
        herr = float((H64 - ref[3]).abs().max())
        print(f"[graph replay {rep}] FP64 H1 max|diff| / max {herr / hmax:.2e}", flush=True)
        assert herr <= 1e-13 * hmax
        step()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(rho, ref[0])


def test_rho1_precisions_graph_replays_match_eager(cuda):
    """Every rho1 forward of the precision study (FP64 full product and symmetric half, INT8 split,
    FP32-MFMA, BF16x3) captured in one HIP graph at its default launch (persistent, work-claiming
    counters cleared by a fill kernel, per-call P' conversion) and replayed four times with eager
    work in between: each replay gives the eager bits."""
    prob = synth.make_rho1_problem(16, 12.0, 150, (30, 200), (20, 230), config_k=16)
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    P1 = api.to_device(synth.make_p1(prob.atom_xyz, prob.basis, 3016))
    precs = (api.FP64, api.FP64_SYM, api.OZAKI, api.FP32, api.BF16X3)
    outs = [torch.empty(prob.n_full_points, dtype=torch.float64, device="cuda") for _ in precs]

    def run():
        for p, o in zip(precs, outs):
            plan.forward(wave, P1, precision=p, out=o)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()  # conversions and work lists made outside the capture
        run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run()
    for rep in range(4):
        for o in outs:
            o.fill_(np.nan)
        plan.forward(wave, P1, precision=api.FP64, out=outs[0])  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for p, o, r in zip(precs, outs, ref):
            assert torch.equal(o, r), f"replay {rep}: precision {p} differs"
