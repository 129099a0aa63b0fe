"""Full first-order H (LDA + GGA terms of evaluate_first_order_H) on the GPU against the CPU oracle.

Bar: within 1e-12 of max|H1| (FP64 both sides).  Cases: random XC derivatives, gradient and hessian
waves, T = 2 with both a full and a short tile, every j_coord, accumulation into a nonzero H, batches
over three 64-row tiles and many 16-point chunks, batches of up to 600 points (the prefactor pass
works in 256-point passes); with the XC derivatives zero it must equal the
LDA adjoint frpa_rho1_adjoint (g = partition_tab * v_hartree_gradient + H_prefactor1); and the
reference-named symbol evaluate_first_order_h_reduce_memory_c_v2_ through ctypes.
"""
import ctypes

import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import _lib, api, synth

pytestmark = pytest.mark.gpu


def _problem(np_range=(20, 200)):
    return synth.make_rho1_problem(12, 11.0, 40, np_range, (30, 150), config_k=12, zero_nc=(5,))


def _xc(prob, T, seed, gga=True):
    rng = np.random.default_rng(seed)
    nmb, nb, nf = prob.n_max_batch_size, prob.n_batches, prob.n_full_points
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    z = lambda *s: np.zeros(s, order="F")
    a = dict(vrho=z(nmb * nb), vsigma=F(3 * nmb * nb) if gga else z(3 * nmb * nb),
             gradient_rho=F(3, nf) if gga else z(3, nf),
             gwave=synth.random_like_wave(prob, 3, seed + 1), hwave=synth.random_like_wave(prob, 6, seed + 2),
             H_prefactor1=F(nmb, nb, T), grad_rho1=F(3, nmb, nb, T) if gga else z(3, nmb, nb, T),
             v_hartree_gradient=F(nf, T), first_order_rho=F(nf, T) if gga else z(nf, T),
             v2rho2=F(3 * nmb * nb) if gga else z(3 * nmb * nb), v2rhosigma=F(6 * nmb * nb) if gga else z(6 * nmb * nb),
             v2sigma2=F(6 * nmb * nb) if gga else z(6 * nmb * nb), partition_tab=rng.uniform(0.1, 1.0, nf))
    return a


def _dev_xc(a):
    return {k: api.to_device(v) for k, v in a.items() if k not in ("vrho", "gwave", "hwave")}


@pytest.mark.parametrize("T,n_tile,jb,jc,np_range", [(2, 2, 2, 3, (20, 200)), (2, 1, 7, 1, (20, 200)),
                                                     (1, 1, 4, 2, (20, 200)), (2, 2, 5, 1, (150, 600))])
def test_h1_gga_against_oracle(cuda, oracle, T, n_tile, jb, jc, np_range):
    prob = _problem(np_range)
    a = _xc(prob, T, 3 * jb + jc)
    H0 = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, size=(prob.n_basis, prob.n_basis, T)))
    ref = oracle.h1(prob, T, jb, jb + n_tile - 1, jc, H0.copy(order="F"), a)
    plan = api.Rho1Plan.from_problem(prob)
    H = api.to_device(H0)
    plan.h1(api.to_device(prob.wave), api.to_device(a["gwave"]), api.to_device(a["hwave"]), _dev_xc(a),
            api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, jb, n_tile=n_tile, j_coord=jc, T=T,
            H=H)
    torch.cuda.synchronize()
    got = H.cpu().numpy().reshape(H0.shape, order="F")
    err = np.abs(got - ref).max() / np.abs(ref[:, :, :n_tile] - H0[:, :, :n_tile]).max()
    print(f"[h1 gga T={T} n_tile={n_tile} jb={jb} jc={jc} np_max={prob.n_point_batches.max()}] rel {err:.2e}",
          flush=True)
    assert err <= 1e-12
    assert np.array_equal(got[:, :, n_tile:], H0[:, :, n_tile:])


def test_h1_without_xc_derivatives_is_the_lda_adjoint(cuda):
    prob = _problem()
    T = 1
    a = _xc(prob, T, 9, gga=False)
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    H = plan.h1(wave, api.to_device(a["gwave"]), api.to_device(a["hwave"]), _dev_xc(a),
                api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, 3, j_coord=2)
    # g(q) = w(q) v_hartree_gradient(q) + H_prefactor1(p, b) at q = map(prefix(b) + p)
    g = a["partition_tab"] * a["v_hartree_gradient"][:, 0]
    for b in range(prob.n_batches):
        s0, s1 = prob.n_point_batches_prefix_sum[b], prob.n_point_batches_prefix_sum[b + 1]
        g[prob.i_valid_point_2_i_full_points_map[s0:s1]] += a["H_prefactor1"][:s1 - s0, b, 0]
    Hl = plan.adjoint(wave, api.to_device(np.asfortranarray(g)), precision=api.FP64)
    torch.cuda.synchronize()
    err = float((H - Hl).abs().max() / Hl.abs().max())
    print(f"[h1 gga with zero XC vs LDA adjoint] rel {err:.2e}", flush=True)
    assert err <= 1e-12


def test_h1_reference_entry_point(cuda, oracle, monkeypatch):
    monkeypatch.setenv("FRPA_ATOM_TILE_SIZE", "2")
    T = 2
    prob = _problem()
    a = _xc(prob, T, 21)
    L = _lib.lib()
    P = ctypes.c_void_p
    keep = []

    def ip(v):
        c = ctypes.c_int(int(v))
        keep.append(c)
        return ctypes.byref(c)

    def ptr(x, dt):
        x = np.asfortranarray(np.asarray(x, dtype=dt))
        keep.append(x)
        return x.ctypes.data_as(P)

    I = lambda x: ptr(x, np.int32)
    D = lambda x: ptr(x, np.float64)
    nb, nmb, nf = prob.n_batches, prob.n_max_batch_size, prob.n_full_points
    zi = np.zeros(64, np.int32)
    zd = np.zeros(nf + 64)
    ibd, avn = prob.dense_i_basis_batches(), prob.atom_valid_n_compute_c_batches()
    L.init_dfpt_device_info()
    L.init_dfpt_device_data(
        ip(nb), ip(nf), ip(0), ip(prob.n_basis), ip(prob.n_atoms), ip(prob.n_max_compute_ham),
        ip(prob.n_centers_basis_I), ip(nmb), ip(0), ip(1), ip(2), ip(0), ip(0), ip(1), ip(1), ip(1), ip(1),
        I(prob.basis.atom + 1), I(prob.batch_sizes), I(prob.n_point_batches), I(prob.n_point_batches_prefix_sum),
        I(prob.i_valid_point_2_i_full_points_map), I(prob.n_compute_c_batches), I(ibd), I(avn),
        I(prob.wave_offset_i32()), D(prob.wave), D(a["gwave"]), D(np.zeros((prob.n_basis, prob.n_basis))),
        I(zi), I(zi), I(zi), D(prob.coord_points), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), D(zd), D(zd),
        D(zd), D(zd), D(zd), D(zd),
        D(np.zeros((prob.n_basis, prob.n_basis, T))), D(np.zeros((nf, T))), D(np.zeros((nmb, nb, T))),
        D(np.zeros((3, nmb, nb, T))), D(np.zeros((3, nmb, nb, T))), I(zi), D(zd), D(zd), D(zd), D(zd), D(zd))
    H0 = np.asfortranarray(np.random.default_rng(4).uniform(-1, 1, size=(prob.n_basis, prob.n_basis, T)))
    H = H0.copy(order="F")
    L.evaluate_first_order_h_reduce_memory_c_v2_(
        ip(3), ip(4), ip(nf), ip(2), ip(prob.n_basis), ip(prob.n_atoms), ip(prob.n_max_compute_ham), ip(nmb), ip(nb),
        ip(prob.n_centers_basis_I), I(prob.n_point_batches), I(prob.n_compute_c_batches), I(ibd), I(avn),
        I(prob.n_point_batches_prefix_sum), I(prob.i_valid_point_2_i_full_points_map), D(a["vrho"]),
        D(a["vsigma"]), D(a["gradient_rho"]), D(zd), I(prob.wave_offset_i32()), D(prob.wave), D(a["gwave"]),
        D(a["hwave"]), H.ctypes.data_as(P), D(a["H_prefactor1"]), D(a["grad_rho1"]), D(a["v_hartree_gradient"]),
        D(a["first_order_rho"]), D(a["v2rho2"]), D(a["v2rhosigma"]), D(a["v2sigma2"]), D(a["partition_tab"]), I(zi))
    L.free_dfpt_device_data()
    L.free_dfpt_device_info()
    ref = oracle.h1(prob, T, 3, 4, 2, H0.copy(order="F"), a)
    err = np.abs(H - ref).max() / np.abs(ref - H0).max()
    print(f"[h1 reference entry point] rel {err:.2e}", flush=True)
    assert err <= 1e-12
