/*
 * CPU FP64 oracle: a plain-C restatement of the reference's CPU paths for the two DFPT grid hot
 * paths and the first-order H integration (their adjoint).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker or the timed CPU baseline.  The product library
 * (fast_rpa_eval_amd/libfrpa.so) never links or calls it.
 *
 * Every function keeps the reference's argument list and column-major layouts; the reference's
 * compile-time ATOM_TILE_SIZE (fast_rap_src_dfpt_gpu_eval/setting.h:19-21) becomes the run-time
 * argument `atom_tile_size`.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   oracle_evaluate_first_order_rho  brute-force sum_ij phi_i P_ij phi_j, rank-1 and unit-matrix
 *                                    closed forms, linearity, nc = 0 batches, NaN padding rows.
 *   oracle_sum_up_far                closed form sum_lm M_lm Y_lm / r^(l+1) (scipy + SHEval
 *                                    coefficients), reference AOT term table fixture, branch and
 *                                    l-reduction special cases.
 *   oracle_ylm                       scipy's spherical harmonics to l = 10 and the SHEval
 *                                    polynomial fixture (l <= 3).
 *   oracle_sum_up_near               closed form sum_lm f_lm(i_r_log) Y_lm for tiles cubic in the
 *                                    grid index (exact for a cubic Hermite spline, clamped ends
 *                                    included), node values, near/far split and l reduction.
 *   oracle_generate_first_order_rho_bias_part2
 *                                    central differences of rho = phi^T DM phi under a displacement
 *                                    of atom j (its basis functions move with it), empty-V zeroing,
 *                                    n_compute_c = 0 batches untouched.
 *   oracle_rho_bias_inline           central differences of sum_ik phi_i DM_ik phi_k with only the
 *                                    first factor moving with atom j; inline(DM) + inline(DM^T)
 *                                    equals bias part 2 of DM.
 *   oracle_evaluate_first_order_h_bias_part2
 *                                    central differences of M(i, k) = sum_p C(i, p) phi_k(p) under a
 *                                    displacement of atom j (only its functions phi_k move); the
 *                                    vrho / vsigma prefactors equal the derivative of a model
 *                                    functional f(rho, sigma) along phi_i (central differences in eps).
 *   oracle_evaluate_first_order_h    adjoint identity <v, rho1(P)> = <H1(v), P>, exact
 *                                    d rho1 / d P_ij = phi_i phi_j, GGA terms against an independent
 *                                    numpy restatement on tiny inputs.
 */
#ifndef FRPA_ORACLE_H
#define FRPA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.cpp:12-210 */
void oracle_evaluate_first_order_rho(
    int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end, int j_coord, int n_full_points,
    int n_basis, int n_atoms, int n_max_compute_ham, int n_centers_basis_I, int n_max_batch_size,
    const int *batch_sizes, const int *n_point_batches, const int *n_point_batches_prefix_sum,
    const int *i_valid_point_2_i_full_points_map, const int *n_compute_c_batches, const int *i_basis_batches,
    const int *atom_valid_n_compute_c_batches, const int *i_batch_2_wave_offset,
    const double *wave_batches_compress, const double *gradient_wave_batches_compress,
    const double *first_order_density_matrix, double *first_order_rho, const double *first_order_rho_bias_part2,
    double *first_order_gradient_rho, const double *first_order_gradient_rho_bias_batches_atoms);

/* Far branch of fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cpp:195-529 (with :91-187);
 * pairs with dist^2 < multipole_radius_sq are the near branch (oracle_sum_up_near) and add nothing
 * here, as in the far-only reference of task3_jit/version5_sumup_far_part_with_jit_or_aot. */
void oracle_sum_up_far(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int l_max_analytic_multipole,
    int index_cc_dim_0, int n_valid_points, int l_pot_max, int n_centers, int n_centers_hartree_potential,
    int n_atoms, int hartree_force_l_add,
    const int *species, const int *l_hartree_max_far_distance, const int *i_valid_point_2_i_full_points_map,
    const double *coord_points, const double *multipole_radius_sq, const double *outer_potential_radius,
    const double *multipole_moments, const int *l_hartree, const int *n_cc_lm_ijk,
    const int *centers_hartree_potential, const int *center_to_atom, const int *index_cc,
    const int *index_ijk_max_cc, const double *cc, const double *coords_center, double *delta_v_hartree);

/* Near branch of fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cpp:382-477 (with :12-49 and
 * sheval.hpp): pairs with dist^2 < multipole_radius_sq, cubic-Hermite spline of the per-atom
 * (lm, 2, n_hartree_grid) tile on the log grid, dotted with Y_lm; accumulated into delta_v_hartree. */
void oracle_sum_up_near(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int n_valid_points, int l_pot_max,
    int n_hartree_grid, int n_centers_hartree_potential, int n_atoms,
    const int *species, const int *species_center, const int *i_valid_point_2_i_full_points_map,
    const double *coord_points, const double *multipole_radius_sq, const double *outer_potential_radius,
    const double *current_delta_v_hart_part_spl_tile, const int *l_hartree, const int *n_grid,
    const int *centers_hartree_potential, const int *center_to_atom, const double *coords_center,
    const double *r_grid_min, const double *log_r_grid_inc, double *delta_v_hartree);

/* Real spherical harmonics of the SHEval convention (sheval.hpp:6-343), ylm[(lmax+1)^2], from
 * (sin theta, cos theta, sin phi, cos phi); lmax <= 31. */
void oracle_ylm(int lmax, double sintheta, double costheta, double sinphi, double cosphi, double *ylm);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H.cpp:13-222 (all LDA and GGA terms). */
void oracle_evaluate_first_order_h(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int j_coord, int n_basis,
    int n_atoms, int n_max_compute_ham, int n_max_batch_size, int n_my_batches_work, int n_centers_basis_I,
    const int *n_point_batches, const int *n_compute_c_batches, const int *i_basis_batches,
    const int *atom_valid_n_compute_c_batches, const int *n_point_batches_prefix_sum,
    const int *i_valid_point_2_i_full_points_map, const double *vrho_batches, const double *vsigma_batches,
    const double *gradient_rho, const int *i_batch_2_wave_offset, const double *wave_batches_compress,
    const double *gradient_wave_batches_compress, const double *hessian_wave_batches_compress,
    double *first_order_H_tile, const double *H_prefactor1_batches_tile,
    const double *first_order_gradient_rho_batches_tile, const double *v_hartree_gradient,
    const double *first_order_rho, const double *v2rho2_batches, const double *v2rhosigma_batches,
    const double *v2sigma2_batches, const double *partition_tab);

/* fast_rap_src_dfpt_gpu_eval/generate_first_order_rho_bias_part2.cpp:9-167: the bias input of rho1
 * (d rho / d R_j of the unperturbed density for the tile's atoms); tile extent atom_tile_size. */
void oracle_generate_first_order_rho_bias_part2(
    int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end, int j_coord, int n_basis,
    int n_atoms, int n_centers_basis_I, int n_max_batch_size, const int *n_point_batches,
    const int *n_compute_c_batches, const int *i_basis_batches, const int *atom_valid_n_compute_c_batches,
    const int *i_batch_2_wave_offset, const double *wave_batches_compress,
    const double *gradient_wave_batches_compress, double *first_order_rho_bias_part2,
    const double *density_matrix);

/* task4_common_calculation_eliminated/version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:179-231
 * (the inline bias of the v6 variant); ADDS into first_order_rho_bias */
void oracle_rho_bias_inline(int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end,
                            int j_coord, int n_basis, int n_atoms, int n_centers_basis_I, int n_max_batch_size,
                            const int *n_point_batches, const int *n_compute_c_batches, const int *i_basis_batches,
                            const int *atom_valid_n_compute_c_batches, const int *i_batch_2_wave_offset,
                            const double *wave_batches_compress, const double *gradient_wave_batches_compress,
                            double *first_order_rho_bias, const double *density_matrix);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H_bias_part2.cpp:13-218 (ADDS into
 * first_order_H_tile; the arguments the reference does not read are omitted) */
void oracle_evaluate_first_order_h_bias_part2(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int j_coord, int n_basis, int n_atoms,
    int n_max_batch_size, int n_my_batches_work, int n_centers_basis_I, const int *n_point_batches,
    const int *n_compute_c_batches, const int *i_basis_batches, const int *atom_valid_n_compute_c_batches,
    const int *n_point_batches_prefix_sum, const int *i_valid_point_2_i_full_points_map, const double *vrho_batches,
    const double *vsigma_batches, const double *gradient_rho, const double *H_times_psi_batches_compress,
    const int *i_batch_2_wave_offset, const double *wave_batches_compress,
    const double *gradient_wave_batches_compress, double *first_order_H_tile, const double *partition_tab);

/* Number of OpenMP threads used by the two grid loops (rho1 over batches, far field over points);
 * 1 reproduces the reference's single-threaded build.  Returns the previous value. */
int oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
