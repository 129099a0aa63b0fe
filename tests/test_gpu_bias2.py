"""Bias input of rho1 (generate_first_order_rho_bias_part2) on the GPU against the CPU oracle.

Bar: within 1e-12 of max|bias| (FP64 both sides).  Cases: a tile of 8 atoms (stacked rows over
several 32-row sweeps), batches of up to 600 points (several 256-point passes), a batch with
n_compute_c = 0 and atoms without functions in a batch, a short tile with a nonsymmetric DM; entries
the generator does not own keep their values; the DM + DM^T table, direct DM gathers and unaligned
wave blocks agree bit for bit; the reference-named drop-in symbol with the reference's tile extent 4.
"""
import ctypes
import dataclasses

import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import _lib, api, synth

pytestmark = pytest.mark.gpu


def _setup(seed, np_range=(20, 200)):
    prob = synth.make_rho1_problem(16, 12.0, 60, np_range, (20, 150), config_k=8, zero_nc=(3,))
    gw = synth.random_like_wave(prob, 3, seed)
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((prob.n_basis, prob.n_basis))
    return prob, gw, A


@pytest.mark.parametrize("T,n_tile,jb,jc,sym,np_range", [(8, 8, 1, 2, True, (20, 200)),
                                                          (3, 2, 5, 3, False, (20, 200)),
                                                          (4, 4, 9, 1, False, (150, 600))])
def test_bias2_against_oracle(cuda, oracle, T, n_tile, jb, jc, sym, np_range):
    prob, gw, A = _setup(T + jb, np_range)
    DM = (A + A.T) / 2 if sym else A
    init = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, size=(prob.n_max_batch_size, prob.n_batches, T)))
    ref = oracle.bias2(prob, DM, gw, T, jb, j_atom_end=jb + n_tile - 1, j_coord=jc, bias=init)
    plan = api.Rho1Plan.from_problem(prob)
    out = api.to_device(init)
    plan.bias_part2(api.to_device(prob.wave), api.to_device(gw), api.to_device(np.asfortranarray(DM)),
                    api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, jb, n_tile=n_tile,
                    j_coord=jc, T=T, out=out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(init.shape, order="F")
    err = np.abs(got - ref).max() / np.abs(ref[:, :, :n_tile]).max()
    print(f"[bias2 T={T} n_tile={n_tile} jb={jb} jc={jc} sym={sym} np_max={prob.n_point_batches.max()}] "
          f"rel {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(got[:, :, n_tile:], init[:, :, n_tile:])
    assert np.array_equal(got[:, 3, :], init[:, 3, :])  # n_compute_c = 0


def _gapped(prob, gw):
    """The same problem with one unused double after every wave block (block offsets no longer on
    4-double boundaries); the gradient wave re-laid out to match."""
    off = prob.wave_offset
    noff = off + np.arange(len(off), dtype=off.dtype)
    wave = np.full(int(noff[-1]), np.nan)
    gwave = np.full(3 * int(noff[-1]), np.nan)
    for b in range(prob.n_batches):
        n = int(off[b + 1] - off[b])
        wave[noff[b]:noff[b] + n] = prob.wave[off[b]:off[b + 1]]
        gwave[3 * noff[b]:3 * noff[b] + 3 * n] = gw[3 * off[b]:3 * off[b + 1]]
    return dataclasses.replace(prob, wave_offset=noff, wave=wave), gwave


def _run(prob, gw, DM, T, n_tile, jb, jc, init):
    plan = api.Rho1Plan.from_problem(prob)
    out = api.to_device(init)
    plan.bias_part2(api.to_device(prob.wave), api.to_device(gw), api.to_device(np.asfortranarray(DM)),
                    api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, jb, n_tile=n_tile,
                    j_coord=jc, T=T, out=out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(init.shape, order="F")


@pytest.mark.parametrize("rows", ["0", "64"])
def test_bias2_paths_agree_bit_for_bit(cuda, monkeypatch, rows):
    """The DM + DM^T table, gathers straight from DM (table off, or a tile with more rows than the
    table holds: 8 atoms, > 64 rows) and unaligned wave blocks all give the same bits."""
    T = 8
    prob, gw, A = _setup(5)
    init = np.zeros((prob.n_max_batch_size, prob.n_batches, T), order="F")
    ref = _run(prob, gw, A, T, T, 1, 2, init)
    monkeypatch.setenv("FRPA_BIAS2_TABLE_ROWS", rows)
    got = _run(prob, gw, A, T, T, 1, 2, init)
    gp, ggw = _gapped(prob, gw)
    assert gp.wave_offset[1] % 4 != 0
    got_gapped = _run(gp, ggw, A, T, T, 1, 2, init)
    monkeypatch.delenv("FRPA_BIAS2_TABLE_ROWS")
    got_gapped_tab = _run(gp, ggw, A, T, T, 1, 2, init)
    nR = prob.atom_valid_n_compute_c_batches()
    print(f"[bias2 paths rows={rows}] max stacked rows {int((nR[T] - nR[0]).max())}, "
          f"nonzero {int(np.count_nonzero(ref))}", flush=True)
    assert np.count_nonzero(ref) > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(got_gapped, ref)
    assert np.array_equal(got_gapped_tab, ref)


def test_bias2_reference_entry_point(cuda, oracle, monkeypatch):
    """generate_first_order_rho_bias_part2_ through ctypes after init_dfpt_device_data, ATOM_TILE_SIZE 4."""
    monkeypatch.setenv("FRPA_ATOM_TILE_SIZE", "4")
    T = 4
    prob, gw, A = _setup(11)
    DM = np.asfortranarray((A + A.T) / 2)
    L = _lib.lib()
    P = ctypes.c_void_p
    keep = []

    def ip(v):
        c = ctypes.c_int(int(v))
        keep.append(c)
        return ctypes.byref(c)

    def ptr(a, dt):
        a = np.asfortranarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data_as(P)

    I = lambda a: ptr(a, np.int32)
    D = lambda a: ptr(a, np.float64)
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    zi = np.zeros(max(prob.n_atoms, 8) + 8, np.int32)
    zd = np.zeros(prob.n_full_points + 8)
    ibd, avn = prob.dense_i_basis_batches(), prob.atom_valid_n_compute_c_batches()
    L.init_dfpt_device_info()
    L.init_dfpt_device_data(
        ip(nb), ip(prob.n_full_points), ip(0), ip(prob.n_basis), ip(prob.n_atoms), ip(prob.n_max_compute_ham),
        ip(prob.n_centers_basis_I), ip(nmb), ip(0), ip(1), ip(2), ip(0), ip(0), ip(1), ip(1), ip(1), ip(1),
        I(prob.basis.atom + 1), I(prob.batch_sizes), I(prob.n_point_batches), I(prob.n_point_batches_prefix_sum),
        I(prob.i_valid_point_2_i_full_points_map), I(prob.n_compute_c_batches), I(ibd), I(avn),
        I(prob.wave_offset_i32()), D(prob.wave), D(gw), D(DM),
        I(zi), I(zi), I(zi), D(prob.coord_points), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), D(zd), D(zd),
        D(zd), D(zd), D(zd), D(zd),
        D(np.zeros((prob.n_basis, prob.n_basis, T))), D(np.zeros((prob.n_full_points, T))),
        D(np.zeros((nmb, nb, T))), D(np.zeros((3, nmb, nb, T))), D(np.zeros((3, nmb, nb, T))),
        I(zi), D(zd), D(zd), D(zd), D(zd), D(zd))
    init = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, size=(nmb, nb, T)))
    bias = init.copy(order="F")
    L.generate_first_order_rho_bias_part2_(
        ip(nb), ip(2), ip(5), ip(1), ip(prob.n_full_points), ip(prob.n_basis), ip(prob.n_atoms),
        ip(prob.n_max_compute_ham), ip(prob.n_centers_basis_I), ip(nmb), I(prob.batch_sizes),
        I(prob.n_point_batches), I(prob.n_compute_c_batches), I(ibd), I(avn), I(prob.wave_offset_i32()),
        D(prob.wave), D(gw), bias.ctypes.data_as(P), D(zd), D(DM))
    L.free_dfpt_device_data()
    L.free_dfpt_device_info()
    ref = oracle.bias2(prob, DM, gw, T, 2, j_atom_end=5, j_coord=1, bias=init)
    err = np.abs(bias - ref).max() / np.abs(ref).max()
    print(f"[bias2 reference entry point] rel {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("T,n_tile,jb,jc,np_range,rows", [(8, 8, 1, 2, (20, 200), None),
                                                          (3, 2, 5, 3, (150, 600), None),
                                                          (4, 3, 9, 1, (20, 200), "0")])
def test_bias_inline_against_oracle(cuda, oracle, monkeypatch, T, n_tile, jb, jc, np_range, rows):
    """The v6 inline bias (one side of a nonsymmetric DM) ADDED to a random bias, within 1e-12 of
    max|inline|; the table and the direct DM gathers (rows = 0) give the same bits; entries of other
    perturbations, of the n_compute_c = 0 batch and past n_points are unchanged."""
    prob, gw, A = _setup(T + jb + 100, np_range)
    init = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, size=(prob.n_max_batch_size, prob.n_batches, T)))
    ref = oracle.bias_inline(prob, A, gw, T, jb, j_atom_end=jb + n_tile - 1, j_coord=jc, bias=init)
    plan = api.Rho1Plan.from_problem(prob)
    args = (api.to_device(prob.wave), api.to_device(gw), api.to_device(np.asfortranarray(A)),
            api.to_device(prob.atom_valid_n_compute_c_batches()), prob.n_atoms, jb)
    got = plan.bias_inline(*args, api.to_device(init), n_tile=n_tile, j_coord=jc, T=T)
    if rows is not None:
        monkeypatch.setenv("FRPA_BIAS2_TABLE_ROWS", rows)
        direct = plan.bias_inline(*args, api.to_device(init), n_tile=n_tile, j_coord=jc, T=T)
        assert torch.equal(direct, got)
    torch.cuda.synchronize()
    got = got.cpu().numpy().reshape(init.shape, order="F")
    err = np.abs(got - ref).max() / np.abs(ref - init).max()
    print(f"[inline bias T={T} n_tile={n_tile} jb={jb} jc={jc} np_max={prob.n_point_batches.max()} rows={rows}] "
          f"rel {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(got[:, :, n_tile:], init[:, :, n_tile:])
    assert np.array_equal(got[:, 3, :], init[:, 3, :])  # n_compute_c = 0
    for b in range(prob.n_batches):
        assert np.array_equal(got[prob.n_point_batches[b]:, b, :], init[prob.n_point_batches[b]:, b, :])


def test_rho1_inline_bias_reference_entry_points(cuda, oracle, monkeypatch):
    """FRPA_RHO1_INLINE_BIAS=1: the rho1 GPU entry point (density_matrix of init_dfpt_device_data) and
    the CPU function's symbol (density_matrix of the call) compute the v6 variant, rho1 = P1 term +
    inline bias + bias input, within 1e-10 at the default precision (gradient 1e-10); the host bias
    input is not modified; without the switch the inline term is absent."""
    monkeypatch.setenv("FRPA_ATOM_TILE_SIZE", "4")
    for k in ("FRPA_RHO1_PRECISION", "FRPA_RHO1_GRADIENT", "FRPA_ALL_KERNELS_ON_GPU"):
        monkeypatch.delenv(k, raising=False)
    T, jb, je, jc = 4, 2, 4, 2
    prob, gw, A = _setup(23)
    DM = np.asfortranarray(A)
    L = _lib.lib()
    P = ctypes.c_void_p
    keep = []

    def ip(v):
        c = ctypes.c_int(int(v))
        keep.append(c)
        return ctypes.byref(c)

    def ptr(a, dt):
        a = np.asfortranarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data_as(P)

    I = lambda a: ptr(a, np.int32)
    D = lambda a: ptr(a, np.float64)
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    rng = np.random.default_rng(29)
    P1 = np.asfortranarray(rng.uniform(-1, 1, size=(prob.n_basis, prob.n_basis, T)))
    bias_in = np.asfortranarray(rng.uniform(-1, 1, size=(nmb, nb, T)))
    gbias = np.asfortranarray(rng.uniform(-1, 1, size=(3, nmb, nb, T)))
    bias = bias_in.copy(order="F")
    rho = np.zeros((prob.n_full_points, T), order="F")
    grad = np.zeros((3, nmb, nb, T), order="F")
    zi = np.zeros(max(prob.n_atoms, 8) + 8, np.int32)
    zd = np.zeros(prob.n_full_points + 8)
    ibd, avn = prob.dense_i_basis_batches(), prob.atom_valid_n_compute_c_batches()
    L.init_dfpt_device_info()
    L.init_dfpt_device_data(
        ip(nb), ip(prob.n_full_points), ip(0), ip(prob.n_basis), ip(prob.n_atoms), ip(prob.n_max_compute_ham),
        ip(prob.n_centers_basis_I), ip(nmb), ip(0), ip(1), ip(2), ip(0), ip(0), ip(1), ip(1), ip(1), ip(1),
        I(prob.basis.atom + 1), I(prob.batch_sizes), I(prob.n_point_batches), I(prob.n_point_batches_prefix_sum),
        I(prob.i_valid_point_2_i_full_points_map), I(prob.n_compute_c_batches), I(ibd), I(avn),
        I(prob.wave_offset_i32()), D(prob.wave), D(gw), D(DM),
        I(zi), I(zi), I(zi), D(prob.coord_points), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), I(zi), D(zd), D(zd),
        D(zd), D(zd), D(zd), D(zd),
        D(np.zeros((prob.n_basis, prob.n_basis, T))), D(np.zeros((prob.n_full_points, T))),
        D(np.zeros((nmb, nb, T))), D(np.zeros((3, nmb, nb, T))), D(np.zeros((3, nmb, nb, T))),
        I(zi), D(zd), D(zd), D(zd), D(zd), D(zd))

    def call(fn):
        fn(ip(nb), ip(jb), ip(je), ip(jc), ip(prob.n_full_points), ip(prob.n_basis), ip(prob.n_atoms),
           ip(prob.n_max_compute_ham), ip(prob.n_centers_basis_I), ip(nmb), I(prob.batch_sizes),
           I(prob.n_point_batches), I(prob.n_point_batches_prefix_sum), I(prob.i_valid_point_2_i_full_points_map),
           I(prob.n_compute_c_batches), I(ibd), I(avn), I(prob.wave_offset_i32()), D(prob.wave), D(gw),
           P1.ctypes.data_as(P), rho.ctypes.data_as(P), bias.ctypes.data_as(P), grad.ctypes.data_as(P),
           gbias.ctypes.data_as(P), I(prob.basis.atom + 1), D(zd), D(DM))

    bias_v6 = oracle.bias_inline(prob, DM, gw, T, jb, j_atom_end=je, j_coord=jc, bias=bias_in)
    ref, gref = oracle.rho1(prob, P1, T=T, bias=bias_v6, gwave=gw, gbias=gbias, j_atom_begin=jb, j_atom_end=je)
    ref_plain, _ = oracle.rho1(prob, P1, T=T, bias=bias_in, gwave=gw, gbias=gbias, j_atom_begin=jb, j_atom_end=je)
    assert np.abs(ref - ref_plain).max() > 1e-2 * np.abs(ref).max()
    try:
        for name in ("evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_",
                     "evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_"):
            for inline, r in (("1", ref), (None, ref_plain)):
                if inline:
                    monkeypatch.setenv("FRPA_RHO1_INLINE_BIAS", inline)
                else:
                    monkeypatch.delenv("FRPA_RHO1_INLINE_BIAS", raising=False)
                rho[:] = -7.0
                call(getattr(L, name))
                rel = np.abs(rho - r).max() / np.abs(r).max()
                grel = np.abs(grad - gref).max() / np.abs(gref).max()
                print(f"[{name} inline bias {inline}] rel {rel:.3e}, gradient rel {grel:.3e}", flush=True)
                assert rel <= 1e-10 and grel <= 1e-10
                assert np.array_equal(bias, bias_in)
                assert np.all(rho[:, je - jb + 1:] == 0.0)
    finally:
        L.free_dfpt_device_data()
        L.free_dfpt_device_info()
