"""ctypes front of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker or the timed CPU baseline.  The product path never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
F64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i, p = ctypes.c_int, ctypes.c_void_p
        L.oracle_evaluate_first_order_rho.argtypes = [i] * 11 + [p] * 15
        L.oracle_sum_up_far.argtypes = [i] * 12 + [p] * 16
        L.oracle_evaluate_first_order_h.argtypes = [i] * 11 + [p] * 22
        L.oracle_sum_up_near.argtypes = [i] * 9 + [p] * 15
        L.oracle_generate_first_order_rho_bias_part2.argtypes = [i] * 9 + [p] * 9
        L.oracle_rho_bias_inline.argtypes = [i] * 9 + [p] * 9
        L.oracle_evaluate_first_order_h_bias_part2.argtypes = [i] * 10 + [p] * 15
        d = ctypes.c_double
        L.oracle_ylm.argtypes = [i, d, d, d, d, p]
        L.oracle_set_threads.argtypes = [i]
        L.oracle_set_threads.restype = i
        _LIB = L
    return _LIB


def _p(a, dtype):
    """Pointer to the memory of a copy-free view; arrays of two or more dimensions must be in
    Fortran order (the C restatement indexes them column-major; a C-order array would be read
    transposed and outputs written to the wrong elements)."""
    a = np.asarray(a)
    assert a.dtype == dtype, (a.dtype, dtype)
    if not a.flags.f_contiguous:
        raise ValueError(f"array of shape {a.shape} must be Fortran-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    return lib().oracle_set_threads(int(n))


def rho1(prob, P1, T=1, bias=None, gwave=None, gbias=None, j_atom_begin=1, j_atom_end=None, i_basis_batches=None):
    """evaluate_first_order_rho on the CPU; returns (first_order_rho (n_full,T) F, gradient (3,nmb,nb,T) F)."""
    nb = prob.n_batches
    if j_atom_end is None:
        j_atom_end = j_atom_begin + T - 1
    if i_basis_batches is None:
        i_basis_batches = prob.dense_i_basis_batches()
    nmb = prob.n_max_batch_size
    if bias is None:
        bias = np.zeros((nmb, nb, T), order="F")
    if gwave is None:
        gwave = np.zeros(3 * int(prob.wave_offset[-1]))
    if gbias is None:
        gbias = np.zeros((3, nmb, nb, T), order="F")
    rho = np.zeros((prob.n_full_points, T), order="F")
    grad = np.zeros((3, nmb, nb, T), order="F")
    P1 = np.asfortranarray(P1, dtype=np.float64)
    lib().oracle_evaluate_first_order_rho(
        T, nb, j_atom_begin, j_atom_end, 1, prob.n_full_points, prob.n_basis, prob.n_atoms,
        prob.n_max_compute_ham, prob.n_centers_basis_I, nmb,
        _p(prob.batch_sizes, np.int32), _p(prob.n_point_batches, np.int32),
        _p(prob.n_point_batches_prefix_sum, np.int32), _p(prob.i_valid_point_2_i_full_points_map, np.int32),
        _p(prob.n_compute_c_batches, np.int32), _p(i_basis_batches, np.int32),
        _p(prob.atom_valid_n_compute_c_batches(), np.int32), _p(prob.wave_offset_i32(), np.int32),
        _p(prob.wave, np.float64), _p(gwave, np.float64), _p(P1, np.float64), _p(rho, np.float64),
        _p(np.asfortranarray(bias), np.float64), _p(grad, np.float64), _p(np.asfortranarray(gbias), np.float64))
    return rho, grad


def far(fp, delta_v=None, j_atom_begin=1, j_atom_end=None):
    """Far-field branch of sum_up_whole_potential on the CPU; delta_v (n_full, T) F accumulated into."""
    T = fp.n_tile
    if j_atom_end is None:
        j_atom_end = j_atom_begin + T - 1
    if delta_v is None:
        delta_v = np.zeros((fp.n_full_points, T), order="F")
    tb = fp.tables
    lib().oracle_sum_up_far(
        T, j_atom_begin, j_atom_end, fp.n_full_points, fp.l_max_analytic_multipole, tb["index_cc_dim_0"],
        fp.n_valid_points, fp.l_pot_max, fp.n_centers, len(fp.centers_hartree_potential), fp.n_atoms,
        fp.hartree_force_l_add,
        _p(fp.species, np.int32), _p(fp.l_hartree_max_far_distance, np.int32),
        _p(fp.i_valid_point_2_i_full_points_map, np.int32), _p(fp.coord_points, np.float64),
        _p(fp.multipole_radius_sq, np.float64), _p(fp.outer_potential_radius, np.float64),
        _p(fp.multipole_moments, np.float64), _p(fp.l_hartree, np.int32), _p(tb["n_cc_lm_ijk"], np.int32),
        _p(fp.centers_hartree_potential, np.int32), _p(fp.center_to_atom, np.int32),
        _p(tb["index_cc"], np.int32), _p(tb["index_ijk_max_cc"], np.int32), _p(tb["cc"], np.float64),
        _p(fp.coords_center, np.float64), _p(delta_v, np.float64))
    return delta_v


def near(fp, delta_v=None, j_atom_begin=1, j_atom_end=None):
    """Near-field branch of sum_up_whole_potential on the CPU (fp carries the near inputs of
    synth.add_near_field); delta_v (n_full, T) F accumulated into."""
    T = fp.n_tile
    if j_atom_end is None:
        j_atom_end = j_atom_begin + T - 1
    if delta_v is None:
        delta_v = np.zeros((fp.n_full_points, T), order="F")
    nf = fp.near
    lib().oracle_sum_up_near(
        T, j_atom_begin, j_atom_end, fp.n_full_points, fp.n_valid_points, fp.l_pot_max, nf["n_hartree_grid"],
        len(fp.centers_hartree_potential), fp.n_atoms,
        _p(fp.species, np.int32), _p(nf["species_center"], np.int32),
        _p(fp.i_valid_point_2_i_full_points_map, np.int32), _p(fp.coord_points, np.float64),
        _p(fp.multipole_radius_sq, np.float64), _p(fp.outer_potential_radius, np.float64),
        _p(nf["spl"], np.float64), _p(fp.l_hartree, np.int32), _p(nf["n_grid"], np.int32),
        _p(fp.centers_hartree_potential, np.int32), _p(fp.center_to_atom, np.int32),
        _p(fp.coords_center, np.float64), _p(nf["r_grid_min"], np.float64), _p(nf["log_r_grid_inc"], np.float64),
        _p(delta_v, np.float64))
    return delta_v


def ylm(lmax, xyz):
    """SHEval-convention Y_lm at unit-vector rows xyz, through the reference's trigonometric route
    (sin/cos of theta and phi as tab_single_trigonom); returns (n, (lmax+1)^2)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    out = np.zeros((len(xyz), (lmax + 1) ** 2))
    buf = np.zeros((lmax + 1) ** 2)
    for n, (x, y, z) in enumerate(xyz):
        ab = np.hypot(x, y)
        if max(abs(x), abs(y)) > 1e-36:
            cphi, sphi = x / ab, y / ab
        else:
            cphi, sphi, ab = 1.0, 0.0, 0.0
        abc = np.hypot(ab, z)
        sth, cth = (ab / abc, z / abc) if max(abs(x), abs(y), abs(z)) > 1e-36 else (1.0, 0.0)
        lib().oracle_ylm(lmax, sth, cth, sphi, cphi, buf.ctypes.data_as(ctypes.c_void_p))
        out[n] = buf
    return out


def bias2(prob, density_matrix, gwave, T, j_atom_begin, j_atom_end=None, j_coord=1, bias=None):
    """generate_first_order_rho_bias_part2 on the CPU; returns bias (n_max_batch_size, n_batches, T) F
    (entries of batches with n_compute_c = 0 keep the given `bias`, zeros by default)."""
    if j_atom_end is None:
        j_atom_end = j_atom_begin + T - 1
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    out = np.zeros((nmb, nb, T), order="F") if bias is None else np.asfortranarray(bias).copy(order="F")
    lib().oracle_generate_first_order_rho_bias_part2(
        T, nb, j_atom_begin, j_atom_end, j_coord, prob.n_basis, prob.n_atoms, prob.n_centers_basis_I, nmb,
        _p(prob.n_point_batches, np.int32), _p(prob.n_compute_c_batches, np.int32),
        _p(prob.dense_i_basis_batches(), np.int32), _p(prob.atom_valid_n_compute_c_batches(), np.int32),
        _p(prob.wave_offset_i32(), np.int32), _p(prob.wave, np.float64), _p(gwave, np.float64), _p(out, np.float64),
        _p(np.asfortranarray(density_matrix, dtype=np.float64), np.float64))
    return out


def bias_inline(prob, density_matrix, gwave, T, j_atom_begin, j_atom_end=None, j_coord=1, bias=None):
    """The v6 variant's inline bias on the CPU, ADDED to `bias` (zeros by default) for the atoms
    j_atom_begin .. j_atom_end; returns (n_max_batch_size, n_batches, T) F."""
    if j_atom_end is None:
        j_atom_end = j_atom_begin + T - 1
    nb, nmb = prob.n_batches, prob.n_max_batch_size
    out = np.zeros((nmb, nb, T), order="F") if bias is None else np.asfortranarray(bias).copy(order="F")
    lib().oracle_rho_bias_inline(
        T, nb, j_atom_begin, j_atom_end, j_coord, prob.n_basis, prob.n_atoms, prob.n_centers_basis_I, nmb,
        _p(prob.n_point_batches, np.int32), _p(prob.n_compute_c_batches, np.int32),
        _p(prob.dense_i_basis_batches(), np.int32), _p(prob.atom_valid_n_compute_c_batches(), np.int32),
        _p(prob.wave_offset_i32(), np.int32), _p(prob.wave, np.float64), _p(gwave, np.float64), _p(out, np.float64),
        _p(np.asfortranarray(density_matrix, dtype=np.float64), np.float64))
    return out


def h1(prob, T, j_atom_begin, j_atom_end, j_coord, H, arrays):
    """evaluate_first_order_h on the CPU, accumulating into H (n_basis, n_basis, T) F.

    arrays: dict with vrho, vsigma, gradient_rho, gwave, hwave, H_prefactor1, grad_rho1,
    v_hartree_gradient, first_order_rho, v2rho2, v2rhosigma, v2sigma2, partition_tab.
    """
    a = {k: np.asfortranarray(v) for k, v in arrays.items()}
    lib().oracle_evaluate_first_order_h(
        T, j_atom_begin, j_atom_end, prob.n_full_points, j_coord, prob.n_basis, prob.n_atoms,
        prob.n_max_compute_ham, prob.n_max_batch_size, prob.n_batches, prob.n_centers_basis_I,
        _p(prob.n_point_batches, np.int32), _p(prob.n_compute_c_batches, np.int32),
        _p(prob.dense_i_basis_batches(), np.int32), _p(prob.atom_valid_n_compute_c_batches(), np.int32),
        _p(prob.n_point_batches_prefix_sum, np.int32), _p(prob.i_valid_point_2_i_full_points_map, np.int32),
        _p(a["vrho"], np.float64), _p(a["vsigma"], np.float64), _p(a["gradient_rho"], np.float64),
        _p(prob.wave_offset_i32(), np.int32), _p(prob.wave, np.float64), _p(a["gwave"], np.float64),
        _p(a["hwave"], np.float64), _p(H, np.float64), _p(a["H_prefactor1"], np.float64),
        _p(a["grad_rho1"], np.float64), _p(a["v_hartree_gradient"], np.float64),
        _p(a["first_order_rho"], np.float64), _p(a["v2rho2"], np.float64), _p(a["v2rhosigma"], np.float64),
        _p(a["v2sigma2"], np.float64), _p(a["partition_tab"], np.float64))
    return H


def h1_bias2(prob, T, j_atom_begin, j_atom_end, j_coord, H, vrho, vsigma, gradient_rho, hpsi, gwave, partition_tab):
    """evaluate_first_order_h_bias_part2 on the CPU, ADDED to H (n_basis, n_basis, T) F; vrho (nmb, nb),
    vsigma (3, nmb, nb), gradient_rho (3, n_full), hpsi like the wave, gwave 3x the wave."""
    lib().oracle_evaluate_first_order_h_bias_part2(
        T, j_atom_begin, j_atom_end, prob.n_full_points, j_coord, prob.n_basis, prob.n_atoms,
        prob.n_max_batch_size, prob.n_batches, prob.n_centers_basis_I,
        _p(prob.n_point_batches, np.int32), _p(prob.n_compute_c_batches, np.int32),
        _p(prob.dense_i_basis_batches(), np.int32), _p(prob.atom_valid_n_compute_c_batches(), np.int32),
        _p(prob.n_point_batches_prefix_sum, np.int32), _p(prob.i_valid_point_2_i_full_points_map, np.int32),
        _p(np.asfortranarray(vrho), np.float64), _p(np.asfortranarray(vsigma), np.float64),
        _p(np.asfortranarray(gradient_rho), np.float64), _p(hpsi, np.float64), _p(prob.wave_offset_i32(), np.int32),
        _p(prob.wave, np.float64), _p(gwave, np.float64), _p(H, np.float64),
        _p(np.asfortranarray(partition_tab), np.float64))
    return H
