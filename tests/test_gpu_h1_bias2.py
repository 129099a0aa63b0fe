"""Bias part 2 of the first-order H (evaluate_first_order_h_bias_part2) on the GPU against the CPU oracle.

Bar: within 1e-12 of max|added H| (FP64 both sides, FP64 atomics on the GPU).  Cases: random
H_times_psi, vrho, vsigma, gradient_rho, partition_tab and gradient wave; a short tile (n_tile < T);
a tile of 8 atoms (112 stacked rows, two sweeps of 64 over the points); a batch with n_compute_c = 0;
accumulation into a nonzero H; the reference-named drop-in symbol after init_dfpt_device_data.
"""
import ctypes

import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import _lib, api, synth

pytestmark = pytest.mark.gpu


def _inputs(prob, seed):
    rng = np.random.default_rng(seed)
    nmb, nb, nf = prob.n_max_batch_size, prob.n_batches, prob.n_full_points
    return dict(gw=synth.random_like_wave(prob, 3, seed + 1), hpsi=synth.random_like_wave(prob, 1, seed + 2),
                vrho=np.asfortranarray(rng.uniform(-1, 1, (nmb, nb))),
                vsig=np.asfortranarray(rng.uniform(-1, 1, (3, nmb, nb))),
                grho=np.asfortranarray(rng.uniform(-1, 1, (3, nf))), w=rng.uniform(0.1, 1.0, nf))


@pytest.mark.parametrize("T,n_tile,jb,jc", [(3, 2, 4, 2), (8, 8, 1, 3), (1, 1, 7, 1)])
def test_h_bias2_against_oracle(cuda, oracle, T, n_tile, jb, jc):
    prob = synth.make_rho1_problem(16, 12.0, 60, (20, 200), (20, 150), config_k=8, zero_nc=(3,))
    a = _inputs(prob, 10 + T)
    H0 = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, (prob.n_basis, prob.n_basis, T)))
    ref = oracle.h1_bias2(prob, T, jb, jb + n_tile - 1, jc, H0.copy(order="F"), a["vrho"], a["vsig"], a["grho"],
                          a["hpsi"], a["gw"], a["w"])
    plan = api.Rho1Plan.from_problem(prob)
    H = api.to_device(H0)
    plan.h1_bias_part2(api.to_device(prob.wave), api.to_device(a["gw"]), api.to_device(a["hpsi"]),
                       api.to_device(a["vrho"]), api.to_device(a["vsig"]), api.to_device(a["grho"]),
                       api.to_device(a["w"]), api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, jb,
                       n_tile=n_tile, j_coord=jc, T=T, H=H)
    torch.cuda.synchronize()
    got = H.cpu().numpy().reshape(H0.shape, order="F")
    avn = prob.atom_valid_n_compute_c_batches()
    rows = int(max((avn[jb - 1 + n_tile, b] - avn[jb - 1, b]) for b in range(prob.n_batches)))
    err = np.abs(got - ref).max() / np.abs(ref - H0).max()
    print(f"[H bias part 2 T={T} n_tile={n_tile} jb={jb} jc={jc} max stacked rows {rows}] rel {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(got[:, :, n_tile:], H0[:, :, n_tile:])


def test_h_bias2_reference_entry_point(cuda, oracle, monkeypatch):
    """evaluate_first_order_h_bias_part2_ through ctypes after init_dfpt_device_data, ATOM_TILE_SIZE 2."""
    monkeypatch.setenv("FRPA_ATOM_TILE_SIZE", "2")
    T = 2
    prob = synth.make_rho1_problem(16, 12.0, 60, (20, 200), (20, 150), config_k=8, zero_nc=(3,))
    a = _inputs(prob, 41)
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
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    zi = np.zeros(max(prob.n_atoms, 8) + 8, np.int32)
    zd = np.zeros(prob.n_full_points + 8)
    ibd, avn = prob.dense_i_basis_batches(), prob.atom_valid_n_compute_c_batches()
    L.init_dfpt_device_info()
    try:
        L.init_dfpt_device_data(
            ip(nb), ip(prob.n_full_points), ip(0), ip(prob.n_basis), ip(prob.n_atoms), ip(prob.n_max_compute_ham),
            ip(prob.n_centers_basis_I), ip(nmb), ip(0), ip(1), ip(2), ip(0), ip(0), ip(1), ip(1), ip(1), ip(1),
            I(prob.basis.atom + 1), I(prob.batch_sizes), I(prob.n_point_batches), I(prob.n_point_batches_prefix_sum),
            I(prob.i_valid_point_2_i_full_points_map), I(prob.n_compute_c_batches), I(ibd), I(avn),
            I(prob.wave_offset_i32()), D(prob.wave), D(a["gw"]), D(np.zeros((prob.n_basis, prob.n_basis))),
            I(zi), I(zi), I(zi), D(prob.coord_points), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), D(zd), D(zd),
            D(zd), D(zd), D(zd), D(zd),
            D(np.zeros((prob.n_basis, prob.n_basis, T))), D(np.zeros((prob.n_full_points, T))),
            D(np.zeros((nmb, nb, T))), D(np.zeros((3, nmb, nb, T))), D(np.zeros((3, nmb, nb, T))),
            I(zi), D(zd), D(zd), D(zd), D(zd), D(zd))
        H0 = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, (prob.n_basis, prob.n_basis, T)))
        H = H0.copy(order="F")
        L.evaluate_first_order_h_bias_part2_(
            ip(5), ip(6), ip(prob.n_full_points), ip(3), ip(prob.n_basis), ip(prob.n_atoms),
            ip(prob.n_max_compute_ham), ip(nmb), ip(nb), ip(prob.n_centers_basis_I), I(prob.n_point_batches),
            I(prob.n_compute_c_batches), I(ibd), I(avn), I(prob.n_point_batches_prefix_sum),
            I(prob.i_valid_point_2_i_full_points_map), D(a["vrho"]), D(a["vsig"]), D(a["grho"]), D(a["hpsi"]),
            I(prob.wave_offset_i32()), D(prob.wave), D(a["gw"]), H.ctypes.data_as(P), D(zd), D(zd), D(zd), D(zd),
            D(zd), D(zd), D(zd), D(a["w"]), I(zi))
    finally:
        L.free_dfpt_device_data()
        L.free_dfpt_device_info()
    ref = oracle.h1_bias2(prob, T, 5, 6, 3, H0.copy(order="F"), a["vrho"], a["vsig"], a["grho"], a["hpsi"], a["gw"],
                          a["w"])
    err = np.abs(H - ref).max() / np.abs(ref - H0).max()
    print(f"[H bias part 2 reference entry point] rel {err:.2e}", flush=True)
    assert err <= 1e-12
