// Implementations of the reference-named entry points (reference_abi.hip), under internal names.
// The reference declares some of these names with C++ linkage (device.hpp:191-192,
// device_data.hpp:159-227, sum_up_whole_potential.h:50-95) and others with C linkage; the library
// exports both forms, each a thin definition over these functions: reference_abi.hip defines the C
// forms (include/frpa.h), reference_abi_cxx.hip the C++ ones.  This header includes nothing of the
// library so that the C++ forms can be defined next to it without the C declarations of frpa.h.
#pragma once

namespace frpa_ref {

void init_device_info();
void free_device_info();

void init_device_data(
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
    double *multipole_moments_ptr, double *current_delta_v_hart_part_spl_tile_ptr, double *delta_v_hartree_ptr);
void free_device_data();

// Both branches of the Hartree sum on the device arrays of init_device_data (the _cu_host_ form).
void sum_up_cu_host(
    int j_atom_begin, int j_atom_end, int n_full_points, int l_max_analytic_multipole, int index_cc_dim_0,
    int n_valid_points, int l_pot_max, int n_max_radial, int n_hartree_grid, int n_centers,
    int n_centers_hartree_potential, int n_atoms, int hartree_force_l_add, int n_species, const int *n_radial_ptr,
    const int *species_ptr, const int *species_center_ptr, const int *l_hartree_max_far_distance_ptr,
    const int *i_valid_point_2_i_full_points_map_ptr, const double *coord_points_ptr,
    const double *multipole_radius_sq_ptr, const double *outer_potential_radius_ptr,
    const double *multipole_moments_ptr, const double *current_delta_v_hart_part_spl_tile_ptr,
    const int *l_hartree_ptr, const int *n_grid_ptr, const int *n_cc_lm_ijk_ptr,
    const int *centers_hartree_potential_ptr, const int *center_to_atom_ptr, const int *index_cc_ptr,
    const int *index_ijk_max_cc_ptr, const double *cc_ptr, const double *coords_center_ptr,
    const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, const double *scale_radial_ptr,
    const double *partition_tab_ptr, double *delta_v_hartree_ptr);

} // namespace frpa_ref
