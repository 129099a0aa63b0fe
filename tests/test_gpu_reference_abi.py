"""The reference's own entry points, called through ctypes with host buffers exactly as its C++
runners call them (fast_rap_src_dfpt_gpu_eval/direct_test/evaluate_first_order_rho_run.cpp,
sum_up_run.cpp): init_dfpt_device_info, init_dfpt_device_data (57 arguments),
evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_ (28 pointer arguments),
sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_ (scalars by value), free_*.
Outputs against the CPU oracle of the same reference functions, ATOM_TILE_SIZE = 2 with a short
tile, GGA gradient and bias included; the whole Hartree sum (near-field splines of two species plus
far-field multipoles) accumulates into delta_v_hartree."""
import ctypes

import numpy as np
import pytest

from fast_rpa_eval_amd import _lib, synth

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _ip(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _arr(a, dtype):
    a = np.asfortranarray(np.asarray(a, dtype=dtype))
    return a, a.ctypes.data_as(P)


def _inputs(prob, fp, T, seed):
    rng = np.random.default_rng(seed)
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    keep = {}

    def ptr(name, a, dtype):
        arr, p = _arr(a, dtype)
        keep[name] = arr
        return p

    d = dict(
        ibasis=ptr("ibasis", prob.dense_i_basis_batches(), np.int32),
        avn=ptr("avn", prob.atom_valid_n_compute_c_batches(), np.int32),
        woff=ptr("woff", prob.wave_offset_i32(), np.int32),
        wave=ptr("wave", prob.wave, np.float64),
        gwave=ptr("gwave", synth.random_like_wave(prob, 3, seed + 1), np.float64),
        P1=ptr("P1", rng.uniform(-1, 1, size=(prob.n_basis, prob.n_basis, T)), np.float64),
        bias=ptr("bias", rng.uniform(-1, 1, size=(nmb, nb, T)), np.float64),
        gbias=ptr("gbias", rng.uniform(-1, 1, size=(3, nmb, nb, T)), np.float64),
        rho=ptr("rho", np.zeros((prob.n_full_points, T)), np.float64),
        grad=ptr("grad", np.zeros((3, nmb, nb, T)), np.float64),
        dv=ptr("dv", rng.uniform(-1, 1, size=(fp.n_full_points, T)), np.float64),
    )
    keep["dv0"] = keep["dv"].copy(order="F")
    return d, keep


def test_reference_entry_points_against_oracle(cuda, oracle, monkeypatch):
    monkeypatch.setenv("FRPA_ATOM_TILE_SIZE", "2")
    monkeypatch.delenv("FRPA_RHO1_PRECISION", raising=False)
    monkeypatch.delenv("FRPA_RHO1_GRADIENT", raising=False)
    T = 2
    prob = synth.make_rho1_problem(10, 11.0, 40, (20, 120), (10, 90), config_k=4)
    fp = synth.make_far_problem(10, 11.0, 4000, 6, config_k=4, xyz=prob.atom_xyz, n_tile=T,
                                outer_radius="decreasing", all_far=False)
    synth.add_near_field(fp, 60, "smooth", n_species=2, seed=9)
    nf = fp.near
    nhg = nf["n_hartree_grid"]
    tb = fp.tables
    L = _lib.lib()
    d, keep = _inputs(prob, fp, T, 21)
    z_i = np.zeros(max(prob.n_atoms, 1) + 8, np.int32)
    z_d = np.zeros(max(prob.n_full_points, fp.n_full_points) + 8)
    zi, zd = z_i.ctypes.data_as(P), z_d.ctypes.data_as(P)
    ints = {k: _arr(v, np.int32) for k, v in dict(
        basis_atom=prob.basis.atom + 1, batch_sizes=prob.batch_sizes, npb=prob.n_point_batches,
        pref=prob.n_point_batches_prefix_sum, map=prob.i_valid_point_2_i_full_points_map,
        ncc=prob.n_compute_c_batches, species=fp.species, lh=fp.l_hartree, nccl=tb["n_cc_lm_ijk"],
        chp=fp.centers_hartree_potential, c2a=fp.center_to_atom, icc=tb["index_cc"],
        ijk=tb["index_ijk_max_cc"], lmfd=fp.l_hartree_max_far_distance,
        fmap=fp.i_valid_point_2_i_full_points_map, spc=nf["species_center"], ngrid=nf["n_grid"]).items()}
    dbl = {k: _arr(v, np.float64) for k, v in dict(
        cp=prob.coord_points, fcp=fp.coord_points, cc=tb["cc"], cen=fp.coords_center,
        mrsq=fp.multipole_radius_sq, opr=fp.outer_potential_radius, M=fp.multipole_moments,
        rmin=nf["r_grid_min"], linc=nf["log_r_grid_inc"], spl=nf["spl"]).items()}
    I = lambda k: ints[k][1]
    D = lambda k: dbl[k][1]

    L.init_dfpt_device_info()
    args = [_ip(prob.n_batches), _ip(prob.n_full_points), _ip(fp.n_valid_points), _ip(prob.n_basis),
            _ip(prob.n_atoms), _ip(prob.n_max_compute_ham), _ip(prob.n_centers_basis_I), _ip(prob.n_max_batch_size),
            _ip(fp.l_pot_max), _ip(1), _ip(nhg), _ip(fp.n_centers), _ip(len(fp.centers_hartree_potential)),
            _ip(fp.hartree_force_l_add), _ip(len(fp.l_hartree)), _ip(fp.l_max_analytic_multipole),
            _ip(tb["index_cc_dim_0"]),
            I("basis_atom"), I("batch_sizes"), I("npb"), I("pref"), I("map"), I("ncc"), d["ibasis"], d["avn"],
            d["woff"], d["wave"], d["gwave"], zd,
            zi, I("species"), I("spc"), D("cp"), I("lh"), I("ngrid"), I("nccl"), I("chp"), I("c2a"), I("icc"),
            I("ijk"), D("cc"), D("cen"), D("rmin"), D("linc"), zd, zd,
            d["P1"], d["rho"], d["bias"], d["grad"], d["gbias"],
            I("lmfd"), D("mrsq"), D("opr"), D("M"), D("spl"), d["dv"]]
    assert len(args) == 57
    L.init_dfpt_device_data(*args)

    def rho1_call(jb, je):
        L.evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_(
            _ip(prob.n_batches), _ip(jb), _ip(je), _ip(1), _ip(prob.n_full_points), _ip(prob.n_basis),
            _ip(prob.n_atoms), _ip(prob.n_max_compute_ham), _ip(prob.n_centers_basis_I), _ip(prob.n_max_batch_size),
            I("batch_sizes"), I("npb"), I("pref"), I("map"), I("ncc"), d["ibasis"], d["avn"], d["woff"], d["wave"],
            d["gwave"], d["P1"], d["rho"], d["bias"], d["grad"], d["gbias"], I("basis_atom"), zd, zd)

    # a short tile: j_atom_begin = 4, j_atom_end = 4 -> one perturbation of the two; the default
    # precision (INT8 split, 5 digits, GGA gradient on every tile pair) within 1e-10, FP64 within 1e-12
    ref, gref = oracle.rho1(prob, keep["P1"], T=T, bias=keep["bias"], gwave=keep["gwave"], gbias=keep["gbias"],
                            j_atom_begin=4, j_atom_end=4)
    for prec, tol in ((None, 1e-10), ("fp64", 1e-12), ("fp64sym", 1e-12)):  # fp64sym: full product (gradient on)
        if prec:
            monkeypatch.setenv("FRPA_RHO1_PRECISION", prec)
        keep["rho"][:] = -7.0
        rho1_call(4, 4)
        rel = np.abs(keep["rho"] - ref).max() / np.abs(ref).max()
        grel = np.abs(keep["grad"] - gref).max() / np.abs(gref).max()
        print(f"[reference rho1 entry, precision {prec or 'default'}] rel {rel:.3e}, gradient rel {grel:.3e}",
              flush=True)
        assert rel <= tol and grel <= tol
        assert np.all(keep["rho"][:, 1] == 0.0)

    # LDA caller without the gradient, INT8 split: rho1 within 1e-10, gradient zeroed
    monkeypatch.setenv("FRPA_RHO1_GRADIENT", "0")
    monkeypatch.setenv("FRPA_RHO1_PRECISION", "ozaki")
    rho1_call(3, 4)
    ref2, _ = oracle.rho1(prob, keep["P1"], T=T, bias=keep["bias"], j_atom_begin=3, j_atom_end=4)
    rel2 = np.abs(keep["rho"] - ref2).max() / np.abs(ref2).max()
    print(f"[reference rho1 entry, INT8 split, no gradient] rel {rel2:.3e}", flush=True)
    assert rel2 <= 1e-10 and np.all(keep["grad"] == 0.0)
    # ... and FP64 on the symmetric half of P1
    monkeypatch.setenv("FRPA_RHO1_PRECISION", "fp64sym")
    keep["rho"][:] = -7.0
    rho1_call(3, 4)
    rel3 = np.abs(keep["rho"] - ref2).max() / np.abs(ref2).max()
    print(f"[reference rho1 entry, FP64 symmetric half, no gradient] rel {rel3:.3e}", flush=True)
    assert rel3 <= 1e-12 and np.all(keep["grad"] == 0.0)

    # whole potential (near-field splines + far-field multipoles): accumulates into delta_v_hartree,
    # both perturbations, two species with their own log grids
    L.sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_.argtypes = [ctypes.c_int] * 14 + [P] * 24
    L.sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_(
        1, T, fp.n_full_points, fp.l_max_analytic_multipole, tb["index_cc_dim_0"], fp.n_valid_points, fp.l_pot_max,
        1, nhg, fp.n_centers, len(fp.centers_hartree_potential), fp.n_atoms, fp.hartree_force_l_add,
        len(fp.l_hartree),
        zi, I("species"), I("spc"), I("lmfd"), I("fmap"), D("fcp"), D("mrsq"), D("opr"), D("M"), D("spl"), I("lh"),
        I("ngrid"), I("nccl"), I("chp"), I("c2a"), I("icc"), I("ijk"), D("cc"), D("cen"), D("rmin"), D("linc"), zd,
        zd, d["dv"])
    vfar = oracle.far(fp, delta_v=keep["dv0"].copy(order="F"), j_atom_begin=1, j_atom_end=T)
    vref = oracle.near(fp, delta_v=vfar.copy(order="F"), j_atom_begin=1, j_atom_end=T)
    assert np.abs(vref - vfar).max() > 0.1 * np.abs(vfar - keep["dv0"]).max()  # the near part matters
    vrel = np.abs(keep["dv"] - vref).max() / np.abs(vref - keep["dv0"]).max()
    for t in range(T):
        dt = np.abs(keep["dv"][:, t] - vref[:, t]).max() / np.abs(vref[:, t] - keep["dv0"][:, t]).max()
        print(f"[reference sum_up_whole_potential entry] perturbation {t}: rel {dt:.3e}", flush=True)
    assert vrel <= 1e-12
    L.free_dfpt_device_data()
    L.free_dfpt_device_info()
