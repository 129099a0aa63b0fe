// The reference entry points that the reference declares with C++ linkage, with that linkage and
// its parameter lists (fast_rap_src_dfpt_gpu_eval/device.hpp:191-192, device_data.hpp:159-227,
// sum_up_whole_potential.h:50-95), so that a host compiled against those headers links against this
// library unchanged.  Each forwards to the implementation shared with the C-linkage form.
#include "reference_impl.hpp"

void init_dfpt_device_info() { frpa_ref::init_device_info(); }
void free_dfpt_device_info() { frpa_ref::free_device_info(); }

void init_dfpt_device_data(
    int *n_my_batches_work_, int *n_full_points_, int *n_valid_points_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_, int *l_pot_max_, int *n_max_radial_,
    int *n_hartree_grid_, int *n_centers_, int *n_centers_hartree_potential_, int *hartree_force_l_add_,
    int *n_species_, int *l_max_analytic_multipole_, int *index_cc_dim_0_,
    int *basis_atom_ptr, int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *density_matrix_ptr,
    int *n_radial_ptr, int *species_ptr, int *species_center_ptr, double *coord_points_ptr, int *l_hartree_ptr,
    int *n_grid_ptr, int *n_cc_lm_ijk_ptr, int *centers_hartree_potential_ptr, int *center_to_atom_ptr,
    int *index_cc_ptr, int *index_ijk_max_cc_ptr, double *cc_ptr, double *coords_center_ptr,
    double *r_grid_min_ptr, double *log_r_grid_inc_ptr, double *scale_radial_ptr, double *partition_tab_ptr,
    double *first_order_density_matrix_ptr, double *first_order_rho_ptr, double *first_order_rho_bias_part2_ptr,
    double *first_order_gradient_rho_ptr, double *first_order_gradient_rho_bias_batches_atoms_ptr,
    int *l_hartree_max_far_distance_ptr, double *multipole_radius_sq_ptr, double *outer_potential_radius_ptr,
    double *multipole_moments_ptr, double *current_delta_v_hart_part_spl_tile_ptr, double *delta_v_hartree_ptr) {
  frpa_ref::init_device_data(
      n_my_batches_work_, n_full_points_, n_valid_points_, n_basis_, n_atoms_, n_max_compute_ham_,
      n_centers_basis_I_, n_max_batch_size_, l_pot_max_, n_max_radial_, n_hartree_grid_, n_centers_,
      n_centers_hartree_potential_, hartree_force_l_add_, n_species_, l_max_analytic_multipole_, index_cc_dim_0_,
      basis_atom_ptr, batch_sizes_ptr, n_point_batches_ptr, n_point_batches_prefix_sum_ptr,
      i_valid_point_2_i_full_points_map_ptr, n_compute_c_batches_ptr, i_basis_batches_ptr,
      atom_valid_n_compute_c_batches_ptr, i_batch_2_wave_offset_ptr, wave_batches_compress_ptr,
      gradient_wave_batches_compress_ptr, density_matrix_ptr, n_radial_ptr, species_ptr, species_center_ptr,
      coord_points_ptr, l_hartree_ptr, n_grid_ptr, n_cc_lm_ijk_ptr, centers_hartree_potential_ptr,
      center_to_atom_ptr, index_cc_ptr, index_ijk_max_cc_ptr, cc_ptr, coords_center_ptr, r_grid_min_ptr,
      log_r_grid_inc_ptr, scale_radial_ptr, partition_tab_ptr, first_order_density_matrix_ptr, first_order_rho_ptr,
      first_order_rho_bias_part2_ptr, first_order_gradient_rho_ptr, first_order_gradient_rho_bias_batches_atoms_ptr,
      l_hartree_max_far_distance_ptr, multipole_radius_sq_ptr, outer_potential_radius_ptr, multipole_moments_ptr,
      current_delta_v_hart_part_spl_tile_ptr, delta_v_hartree_ptr);
}

void free_dfpt_device_data() { frpa_ref::free_device_data(); }

void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_(
    const int j_atom_begin, const int j_atom_end, const int n_full_points, const int l_max_analytic_multipole,
    const int index_cc_dim_0, const int n_valid_points, const int l_pot_max, const int n_max_radial,
    const int n_hartree_grid, const int n_centers, const int n_centers_hartree_potential, const int n_atoms,
    const int hartree_force_l_add, const int n_species, const int *n_radial_ptr, const int *species_ptr,
    const int *species_center_ptr, const int *l_hartree_max_far_distance_ptr,
    const int *i_valid_point_2_i_full_points_map_ptr, const double *coord_points_ptr,
    const double *multipole_radius_sq_ptr, const double *outer_potential_radius_ptr,
    const double *multipole_moments_ptr, const double *current_delta_v_hart_part_spl_tile_ptr,
    const int *l_hartree_ptr, const int *n_grid_ptr, const int *n_cc_lm_ijk_ptr,
    const int *centers_hartree_potential_ptr, const int *center_to_atom_ptr, const int *index_cc_ptr,
    const int *index_ijk_max_cc_ptr, const double *cc_ptr, const double *coords_center_ptr,
    const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, const double *scale_radial_ptr,
    const double *partition_tab_ptr, double *delta_v_hartree_ptr) {
  frpa_ref::sum_up_cu_host(
      j_atom_begin, j_atom_end, n_full_points, l_max_analytic_multipole, index_cc_dim_0, n_valid_points, l_pot_max,
      n_max_radial, n_hartree_grid, n_centers, n_centers_hartree_potential, n_atoms, hartree_force_l_add, n_species,
      n_radial_ptr, species_ptr, species_center_ptr, l_hartree_max_far_distance_ptr,
      i_valid_point_2_i_full_points_map_ptr, coord_points_ptr, multipole_radius_sq_ptr, outer_potential_radius_ptr,
      multipole_moments_ptr, current_delta_v_hart_part_spl_tile_ptr, l_hartree_ptr, n_grid_ptr, n_cc_lm_ijk_ptr,
      centers_hartree_potential_ptr, center_to_atom_ptr, index_cc_ptr, index_ijk_max_cc_ptr, cc_ptr,
      coords_center_ptr, r_grid_min_ptr, log_r_grid_inc_ptr, scale_radial_ptr, partition_tab_ptr,
      delta_v_hartree_ptr);
}
