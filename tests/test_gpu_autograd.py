"""autograd wrappers: gradients against FP64 central finite differences (step 1e-6), the CPU oracle of
the first-order H integration and the closed form dV/dM_lm = Y_lm / r^(l+1)."""
import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import api, autograd, multipole_tables as mt, synth

pytestmark = pytest.mark.gpu


def test_rho1_grad_is_h1_and_matches_fd(cuda, oracle):
    prob = synth.make_rho1_problem(16, 12.0, 120, (50, 150), (30, 224), config_k=4)
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    rng = np.random.default_rng(5)
    P = api.to_device(synth.make_p1(prob.atom_xyz, prob.basis, 3004)).requires_grad_(True)
    v = api.to_device(rng.standard_normal(prob.n_full_points))
    L = torch.dot(v, autograd.rho1(P, plan, wave))
    L.backward()
    G = P.grad.detach()
    # finite differences along a random direction (L is linear in P1: exact up to rounding)
    dP = api.to_device(rng.standard_normal(prob.n_basis * prob.n_basis))
    h = 1e-6
    with torch.no_grad():
        Lp = torch.dot(v, plan.forward(wave, P.detach() + h * dP))
        Lm = torch.dot(v, plan.forward(wave, P.detach() - h * dP))
    fd = float((Lp - Lm) / (2 * h))
    an = float(torch.dot(G, dP))
    print(f"[rho1 autograd] fd {fd:.10e} analytic {an:.10e}", flush=True)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    # oracle: evaluate_first_order_H with partition_tab = 1, v_hartree_gradient = v
    from tests.test_gpu_adjoint import _oracle_h1
    ref = _oracle_h1(oracle, prob, np.asfortranarray(v.cpu().numpy()[:, None]), 1)[:, :, 0]
    got = G.cpu().numpy().reshape((prob.n_basis, prob.n_basis), order="F")
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_far_grad_matches_fd_and_closed_form(cuda):
    fp = synth.make_far_problem(12, 10.0, 4000, 6, config_k=8, all_far=True)
    plan = api.FarPlan.from_problem(fp)
    inp = api.FarInputs(fp)
    rng = np.random.default_rng(6)
    M = inp.moments.clone().requires_grad_(True)
    gV = api.to_device(rng.standard_normal(fp.n_full_points))
    L = torch.dot(gV, autograd.far_field(M, plan, inp))
    L.backward()
    G = M.grad.detach()
    dM = api.to_device(rng.standard_normal(M.numel()))
    h = 1e-6
    with torch.no_grad():
        Lp = torch.dot(gV, autograd.far_field(M.detach() + h * dM, plan, inp))
        Lm = torch.dot(gV, autograd.far_field(M.detach() - h * dM, plan, inp))
    fd = float((Lp - Lm) / (2 * h))
    an = float(torch.dot(G, dM))
    print(f"[far autograd] fd {fd:.10e} analytic {an:.10e}", flush=True)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    # closed form: dL/dM_lm(a) = sum_p gV(p) Y_lm(r^) / r^(l+1)
    mp = fp.i_valid_point_2_i_full_points_map
    pts = fp.coord_points[:, mp].T
    g = gV.cpu().numpy()[mp]
    Gn = G.cpu().numpy().reshape(fp.multipole_moments.shape, order="F")
    for a in (0, 5, 11):
        d = pts - fp.coords_center[:, a]
        r = np.linalg.norm(d, axis=1)
        Y = mt.real_sph_harm(6, d / r[:, None])
        for l in range(7):
            ref = (g[:, None] * Y[:, l * l:(l + 1) ** 2] / r[:, None] ** (l + 1)).sum(axis=0)
            np.testing.assert_allclose(Gn[l * l:(l + 1) ** 2, a, 0], ref, rtol=1e-10, atol=1e-14)
