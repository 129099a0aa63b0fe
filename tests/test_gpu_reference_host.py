"""A C++ host program compiled against the reference's own declarations (tests/cxx/reference_host.cpp:
C++ linkage for init_dfpt_device_info / init_dfpt_device_data / free_* and the Hartree _cu_host_,
fast_rap_src_dfpt_gpu_eval/device.hpp:191-192, device_data.hpp:159-227, sum_up_whole_potential.h:50-95;
C linkage for the rho1 functions and the CPU Hartree sum, evaluate_first_order_rho.h:3-69,
sum_up_whole_potential.h:3-48) links against libfrpa.so and runs the direct-test runners' calls:
the rho1 GPU entry point, the rho1 CPU function's symbol, the device-resident (ALL_KERNELS_ON_GPU)
mode with the caller filling and reading the device arrays, the Hartree GPU entry point and the
Hartree CPU function's symbol.  Every output against the CPU oracle (ATOM_TILE_SIZE = 2 with a short
tile, GGA gradient and bias, near + far field accumulated into delta_v_hartree): rho1 and its gradient
at the entry points' default precision (the INT8 split, 5 digits) within 1e-10, the Hartree sum
(FP64) within 1e-12."""
import os
import subprocess

import numpy as np
import pytest

from fast_rpa_eval_amd import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cxx", "reference_host")


def _write(d, arrays, scalars):
    for k, (a, dt) in arrays.items():
        np.asfortranarray(np.asarray(a, dtype=dt)).ravel(order="F").tofile(os.path.join(d, k + ".bin"))
    with open(os.path.join(d, "scalars.txt"), "w") as f:
        for k, v in scalars.items():
            f.write(f"{k} {int(v)}\n")


def _read(d, name, shape):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=np.float64).reshape(shape, order="F")


def test_cxx_host_against_oracle(cuda, oracle, tmp_path):
    assert os.path.exists(BIN), "tests/cxx/reference_host not built (make -C tests/cxx)"
    T, jb, je = 2, 4, 4  # a short tile: one perturbation of the two
    prob = synth.make_rho1_problem(10, 11.0, 40, (20, 120), (10, 90), config_k=4)
    fp = synth.make_far_problem(10, 11.0, 4000, 6, config_k=4, xyz=prob.atom_xyz, n_tile=T,
                                outer_radius="decreasing", all_far=False)
    synth.add_near_field(fp, 60, "smooth", n_species=2, seed=9)
    nf, tb = fp.near, fp.tables
    rng = np.random.default_rng(31)
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    gwave = synth.random_like_wave(prob, 3, 32)
    P1 = rng.uniform(-1, 1, size=(prob.n_basis, prob.n_basis, T))
    bias = rng.uniform(-1, 1, size=(nmb, nb, T))
    gbias = rng.uniform(-1, 1, size=(3, nmb, nb, T))
    dv0 = rng.uniform(-1, 1, size=(fp.n_full_points, T))
    i32, f64 = np.int32, np.float64
    arrays = dict(
        basis_atom=(prob.basis.atom + 1, i32), batch_sizes=(prob.batch_sizes, i32),
        n_point_batches=(prob.n_point_batches, i32), n_point_batches_prefix_sum=(prob.n_point_batches_prefix_sum, i32),
        i_valid_point_2_i_full_points_map=(prob.i_valid_point_2_i_full_points_map, i32),
        n_compute_c_batches=(prob.n_compute_c_batches, i32), i_basis_batches=(prob.dense_i_basis_batches(), i32),
        atom_valid_n_compute_c_batches=(prob.atom_valid_n_compute_c_batches(), i32),
        i_batch_2_wave_offset=(prob.wave_offset_i32(), i32), wave_batches_compress=(prob.wave, f64),
        gradient_wave_batches_compress=(gwave, f64), density_matrix=(np.zeros((prob.n_basis, prob.n_basis)), f64),
        n_radial=(np.ones(len(fp.l_hartree)), i32), species=(fp.species, i32),
        species_center=(nf["species_center"], i32), coord_points=(prob.coord_points, f64),
        l_hartree=(fp.l_hartree, i32), n_grid=(nf["n_grid"], i32), n_cc_lm_ijk=(tb["n_cc_lm_ijk"], i32),
        centers_hartree_potential=(fp.centers_hartree_potential, i32), center_to_atom=(fp.center_to_atom, i32),
        index_cc=(tb["index_cc"], i32), index_ijk_max_cc=(tb["index_ijk_max_cc"], i32), cc=(tb["cc"], f64),
        coords_center=(fp.coords_center, f64), r_grid_min=(nf["r_grid_min"], f64),
        log_r_grid_inc=(nf["log_r_grid_inc"], f64), scale_radial=(np.ones(len(fp.l_hartree)), f64),
        partition_tab=(np.ones(max(prob.n_full_points, fp.n_full_points)), f64),
        first_order_density_matrix=(P1, f64), first_order_rho_bias_part2=(bias, f64),
        first_order_gradient_rho_bias_batches_atoms=(gbias, f64),
        l_hartree_max_far_distance=(fp.l_hartree_max_far_distance, i32),
        multipole_radius_sq=(fp.multipole_radius_sq, f64), outer_potential_radius=(fp.outer_potential_radius, f64),
        multipole_moments=(fp.multipole_moments, f64), current_delta_v_hart_part_spl_tile=(nf["spl"], f64),
        delta_v_hartree=(dv0, f64), far_coord_points=(fp.coord_points, f64),
        far_i_valid_point_2_i_full_points_map=(fp.i_valid_point_2_i_full_points_map, i32))
    scalars = dict(
        n_my_batches_work=nb, n_full_points=prob.n_full_points, n_valid_points=fp.n_valid_points,
        n_basis=prob.n_basis, n_atoms=prob.n_atoms, n_max_compute_ham=prob.n_max_compute_ham,
        n_centers_basis_I=prob.n_centers_basis_I, n_max_batch_size=nmb, l_pot_max=fp.l_pot_max, n_max_radial=1,
        n_hartree_grid=nf["n_hartree_grid"], n_centers=fp.n_centers,
        n_centers_hartree_potential=len(fp.centers_hartree_potential), hartree_force_l_add=fp.hartree_force_l_add,
        n_species=len(fp.l_hartree), l_max_analytic_multipole=fp.l_max_analytic_multipole,
        index_cc_dim_0=tb["index_cc_dim_0"], atom_tile_size=T, j_atom_begin=jb, j_atom_end=je,
        far_n_full_points=fp.n_full_points, far_n_valid_points=fp.n_valid_points)
    d = str(tmp_path)
    _write(d, arrays, scalars)
    env = dict(os.environ, FRPA_ATOM_TILE_SIZE=str(T))
    for k in ("FRPA_RHO1_PRECISION", "FRPA_RHO1_GRADIENT", "FRPA_ALL_KERNELS_ON_GPU", "FRPA_FAR_VARIANT",
              "FRPA_RHO1_INLINE_BIAS"):
        env.pop(k, None)
    print("[reference_host] running the C++ host program", flush=True)
    r = subprocess.run([BIN, d], env=env, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.returncode

    ref, gref = oracle.rho1(prob, P1, T=T, bias=bias, gwave=gwave, gbias=gbias, j_atom_begin=jb, j_atom_end=je)
    for mode in ("cu_host", "cpu_symbol", "resident"):
        rho = _read(d, "out_rho_" + mode, (prob.n_full_points, T))
        grad = _read(d, "out_grad_" + mode, (3, nmb, nb, T))
        rel = np.abs(rho - ref).max() / np.abs(ref).max()
        grel = np.abs(grad - gref).max() / np.abs(gref).max()
        print(f"[reference_host rho1 {mode}] rel {rel:.3e}, gradient rel {grel:.3e}", flush=True)
        assert rel <= 1e-10 and grel <= 1e-10
        assert np.all(rho[:, 1] == 0.0)  # the tile's second perturbation is outside [jb, je]
    vfar = oracle.far(fp, delta_v=np.asfortranarray(dv0.copy()), j_atom_begin=jb, j_atom_end=je)
    vref = oracle.near(fp, delta_v=vfar.copy(order="F"), j_atom_begin=jb, j_atom_end=je)
    for mode in ("cu_host", "cpu_symbol"):
        dv = _read(d, "out_dv_" + mode, (fp.n_full_points, T))
        rel = np.abs(dv - vref).max() / np.abs(vref - dv0).max()
        print(f"[reference_host sum_up {mode}] rel {rel:.3e}", flush=True)
        assert rel <= 1e-12
