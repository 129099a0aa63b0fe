// Drop-in entry points with the reference's names and Fortran-ABI argument lists.
//
// The reference keeps device copies of every array in process-global singletons (devPs, devInfo;
// fast_rap_src_dfpt_gpu_eval/device_data.cu:6-8, device.cu:5), filled once by
// init_dfpt_device_data and used by the *_cu_host_ functions.  The same lifecycle is kept here:
// init_dfpt_device_info creates the stream, init_dfpt_device_data builds the rho1 and far-field
// plans and uploads the constant arrays; the *_cu_host_ functions upload their per-call arrays,
// run on the device and copy the results back.  Errors print and exit(1), as DEV_CHECK
// (device.hpp:19-42).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "farfield_common.hpp"
#include "rho1_common.hpp"

namespace {

using namespace frpa;

struct RefState {
  bool info = false;
  hipStream_t stream = nullptr;
  bool data = false;
  int T = 1;
  int n_batches = 0, n_full = 0, n_basis = 0, nmb = 0, n_valid_points = 0, n_atoms = 0;
  std::unique_ptr<Rho1Plan> rho1;
  std::unique_ptr<FarPlan> far;
  DevBuf<double> wave, gwave, P1, bias, gbias, rho, grad, dm;
  // per-call arrays of the first-order H entry point
  DevBuf<double> h_hw, h_vs, h_v2rr, h_v2rs, h_v2ss, h_grho, h_pref1, h_grho1, h_vhg, h_rho1, h_w, h_H;
  DevBuf<int> avn;  // atom_valid_n_compute_c_batches (n_atoms + 1, n_batches)
  DevBuf<double> coord_points;
  // far per-call buffers
  DevBuf<double> f_cp, f_mrsq, f_opr, f_M, f_dv, f_spl;
  DevBuf<int> f_map, f_lmfd;
};

RefState &ref() {
  static RefState s;
  return s;
}

[[noreturn]] void die(const char *where) {
  fprintf(stderr, "%s: %s\n", where, frpa_last_error());
  exit(1);
}

template <class F>
void run_or_die(const char *where, F &&f) {
  if (guarded(std::forward<F>(f)) != FRPA_OK)
    die(where);
}

// FRPA_RHO1_PRECISION: fp64 (default), ozaki (INT8 split, 5 digits), ozaki4 (4 digits).
// FRPA_RHO1_GRADIENT=0: first_order_gradient_rho is not computed (left zero; LDA callers).  The
// GGA gradient needs W in FP64, so with the gradient requested the FP64 kernel runs whatever the
// precision.
int rho1_precision_from_env() {
  const char *p = getenv("FRPA_RHO1_PRECISION");
  if (p && strcmp(p, "ozaki") == 0)
    return FRPA_PREC_OZAKI;
  if (p && strcmp(p, "ozaki4") == 0)
    return FRPA_PREC_OZAKI4;
  return FRPA_PREC_FP64;
}

bool rho1_gradient_from_env() {
  const char *g = getenv("FRPA_RHO1_GRADIENT");
  return !(g && strcmp(g, "0") == 0);
}

} // namespace

extern "C" {

void init_dfpt_device_info(void) {
  run_or_die("init_dfpt_device_info", [] {
    RefState &S = ref();
    int n = 0;
    FRPA_HIP(hipGetDeviceCount(&n));
    FRPA_REQUIRE(n > 0, FRPA_ERR_HIP, "no HIP device");
    FRPA_HIP(hipSetDevice(0));
    if (!S.stream)
      FRPA_HIP(hipStreamCreate(&S.stream));
    S.info = true;
  });
}

void free_dfpt_device_info(void) {
  RefState &S = ref();
  if (S.stream)
    (void)hipStreamDestroy(S.stream);
  S.stream = nullptr;
  S.info = false;
}

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
  run_or_die("init_dfpt_device_data", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.info, FRPA_ERR_STATE, "init_dfpt_device_info must be called first");
    const char *ts = getenv("FRPA_ATOM_TILE_SIZE");
    S.T = ts ? atoi(ts) : 1;
    FRPA_REQUIRE(S.T >= 1, FRPA_ERR_INVALID, "FRPA_ATOM_TILE_SIZE must be >= 1");
    const int nb = *n_my_batches_work_, nbI = *n_centers_basis_I_;
    S.n_batches = nb;
    S.n_full = *n_full_points_;
    S.n_basis = *n_basis_;
    S.nmb = *n_max_batch_size_;
    S.n_valid_points = *n_valid_points_;
    S.n_atoms = *n_atoms_;
    // dense (n_centers_basis_I, n_batches) 1-based lists -> CSR; int32 wave offsets -> int64
    std::vector<long long> ibp(nb + 1, 0), woff(nb + 1);
    std::vector<int> ib;
    for (int b = 0; b < nb; b++) {
      const int nc = n_compute_c_batches_ptr[b];
      FRPA_REQUIRE(nc >= 0 && nc <= nbI, FRPA_ERR_INVALID, "n_compute_c(%d) = %d", b, nc);
      ib.insert(ib.end(), i_basis_batches_ptr + (size_t)nbI * b, i_basis_batches_ptr + (size_t)nbI * b + nc);
      ibp[b + 1] = ibp[b] + nc;
    }
    for (int b = 0; b <= nb; b++)
      woff[b] = i_batch_2_wave_offset_ptr[b];
    S.rho1.reset(make_rho1_plan(nb, S.n_basis, S.n_full, S.nmb, n_point_batches_ptr,
                                    n_point_batches_prefix_sum_ptr, i_valid_point_2_i_full_points_map_ptr,
                                    n_compute_c_batches_ptr, ibp.data(), ib.data(), woff.data()));
    const size_t wl = nb ? (size_t)woff[nb] : 0;
    S.wave.upload(wave_batches_compress_ptr, wl);
    S.gwave.upload(gradient_wave_batches_compress_ptr, 3 * wl);
    S.avn.upload(atom_valid_n_compute_c_batches_ptr, (size_t)(S.n_atoms + 1) * nb);
    S.P1.alloc((size_t)S.n_basis * S.n_basis * S.T);
    S.bias.alloc((size_t)S.nmb * nb * S.T);
    S.gbias.alloc((size_t)3 * S.nmb * nb * S.T);
    S.rho.alloc((size_t)S.n_full * S.T);
    S.grad.alloc((size_t)3 * S.nmb * nb * S.T);
    S.coord_points.upload(coord_points_ptr, (size_t)3 * S.n_full);
    if (*n_centers_hartree_potential_ > 0 && *n_atoms_ > 0)
      S.far.reset(make_far_plan(*l_pot_max_, *l_max_analytic_multipole_, *index_cc_dim_0_, *n_centers_,
                                    *n_centers_hartree_potential_, *n_atoms_, *n_species_, *hartree_force_l_add_,
                                    species_ptr, l_hartree_ptr, centers_hartree_potential_ptr, center_to_atom_ptr,
                                    coords_center_ptr, index_cc_ptr, cc_ptr, n_cc_lm_ijk_ptr,
                                    index_ijk_max_cc_ptr));
    S.data = true;
  });
}

void free_dfpt_device_data(void) {
  RefState &S = ref();
  S.rho1.reset();
  S.far.reset();
  S.f_spl.release();
  S.wave.release();
  S.gwave.release();
  S.P1.release();
  S.bias.release();
  S.gbias.release();
  S.dm.release();
  S.avn.release();
  for (DevBuf<double> *b : {&S.h_hw, &S.h_vs, &S.h_v2rr, &S.h_v2rs, &S.h_v2ss, &S.h_grho, &S.h_pref1, &S.h_grho1,
                            &S.h_vhg, &S.h_rho1, &S.h_w, &S.h_H})
    b->release();
  S.rho.release();
  S.grad.release();
  S.coord_points.release();
  S.f_cp.release();
  S.f_mrsq.release();
  S.f_opr.release();
  S.f_M.release();
  S.f_dv.release();
  S.f_map.release();
  S.f_lmfd.release();
  S.data = false;
}

void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_density_matrix_ptr, double *first_order_rho_ptr,
    double *first_order_rho_bias_part2_ptr, double *first_order_gradient_rho_ptr,
    double *first_order_gradient_rho_bias_batches_atoms_ptr, int *basis_atom_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr) {
  run_or_die("evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.data && S.rho1, FRPA_ERR_STATE, "init_dfpt_device_data must be called first");
    FRPA_REQUIRE(*n_my_batches_work_ == S.n_batches && *n_full_points_ == S.n_full && *n_basis_ == S.n_basis &&
                     *n_max_batch_size_ == S.nmb,
                 FRPA_ERR_INVALID, "sizes differ from init_dfpt_device_data");
    const int n_tile = *j_atom_end_ - *j_atom_begin_ + 1;
    FRPA_REQUIRE(n_tile >= 1 && n_tile <= S.T, FRPA_ERR_INVALID,
                 "j_atom_end - j_atom_begin + 1 = %d outside [1, ATOM_TILE_SIZE = %d]", n_tile, S.T);
    hipStream_t s = S.stream;
    const size_t nbt = (size_t)S.nmb * S.n_batches * S.T;
    FRPA_HIP(hipMemcpyAsync(S.P1.ptr, first_order_density_matrix_ptr, S.P1.n * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.bias.ptr, first_order_rho_bias_part2_ptr, nbt * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.gbias.ptr, first_order_gradient_rho_bias_batches_atoms_ptr, 3 * nbt * 8,
                            hipMemcpyHostToDevice, s));
    const bool grad = rho1_gradient_from_env();
    const int prec = grad ? FRPA_PREC_FP64 : rho1_precision_from_env();
    const int rc = frpa_rho1_forward(reinterpret_cast<frpa_rho1_plan *>(S.rho1.get()), S.T, n_tile, prec,
                                     S.wave.ptr, S.P1.ptr, S.bias.ptr, S.rho.ptr, grad ? S.gwave.ptr : nullptr,
                                     grad ? S.gbias.ptr : nullptr, grad ? S.grad.ptr : nullptr, s);
    FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
    if (!grad)  // evaluate_first_order_rho.cpp:113-114 zeroes both outputs
      fill_f64(S.grad.ptr, 3 * nbt, 0.0, s);
    FRPA_HIP(hipMemcpyAsync(first_order_rho_ptr, S.rho.ptr, S.rho.n * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipMemcpyAsync(first_order_gradient_rho_ptr, S.grad.ptr, 3 * nbt * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipStreamSynchronize(s));
  });
}

void generate_first_order_rho_bias_part2_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_rho_bias_part2_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr) {
  run_or_die("generate_first_order_rho_bias_part2_", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.data && S.rho1, FRPA_ERR_STATE,
                 "init_dfpt_device_data must be called first (the batch data are used from the device)");
    FRPA_REQUIRE(*n_my_batches_work_ == S.n_batches && *n_basis_ == S.n_basis && *n_max_batch_size_ == S.nmb &&
                     *n_atoms_ == S.n_atoms,
                 FRPA_ERR_INVALID, "sizes differ from init_dfpt_device_data");
    const int n_tile = *j_atom_end_ - *j_atom_begin_ + 1;
    FRPA_REQUIRE(n_tile >= 1 && n_tile <= S.T, FRPA_ERR_INVALID,
                 "j_atom_end - j_atom_begin + 1 = %d outside [1, ATOM_TILE_SIZE = %d]", n_tile, S.T);
    hipStream_t s = S.stream;
    const size_t nbt = (size_t)S.nmb * S.n_batches * S.T;
    S.dm.alloc((size_t)S.n_basis * S.n_basis);
    FRPA_HIP(hipMemcpyAsync(S.dm.ptr, density_matrix_ptr, S.dm.n * 8, hipMemcpyHostToDevice, s));
    // entries the generator does not write (n_compute_c = 0 batches, points past n_points) keep the
    // caller's values, as with the reference's in-place CPU loop
    FRPA_HIP(hipMemcpyAsync(S.bias.ptr, first_order_rho_bias_part2_ptr, nbt * 8, hipMemcpyHostToDevice, s));
    const int rc = frpa_rho1_bias_part2(reinterpret_cast<frpa_rho1_plan *>(S.rho1.get()), S.T, *j_atom_begin_,
                                        n_tile, *j_coord_, S.n_atoms, S.avn.ptr, S.wave.ptr, S.gwave.ptr, S.dm.ptr,
                                        S.bias.ptr, s);
    FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
    FRPA_HIP(hipMemcpyAsync(first_order_rho_bias_part2_ptr, S.bias.ptr, nbt * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipStreamSynchronize(s));
  });
}

void evaluate_first_order_h_reduce_memory_c_v2_(
    int *j_atom_begin_, int *j_atom_end_, int *n_full_points_, int *j_coord_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_max_batch_size_, int *n_my_batches_work_, int *n_centers_basis_I_,
    int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, double *vrho_batches_ptr, double *vsigma_batches_ptr,
    double *gradient_rho_ptr, double *H_times_psi_batches_compress_ptr, int *i_batch_2_wave_offset_ptr,
    double *wave_batches_compress_ptr, double *gradient_wave_batches_compress_ptr,
    double *hessian_wave_batches_compress_ptr, double *first_order_H_tile_ptr, double *H_prefactor1_batches_tile_ptr,
    double *first_order_gradient_rho_batches_tile_ptr, double *v_hartree_gradient_ptr, double *first_order_rho_ptr,
    double *v2rho2_batches_ptr, double *v2rhosigma_batches_ptr, double *v2sigma2_batches_ptr,
    double *partition_tab_ptr, int *not_use) {
  run_or_die("evaluate_first_order_h_reduce_memory_c_v2_", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.data && S.rho1, FRPA_ERR_STATE,
                 "init_dfpt_device_data must be called first (the wave arrays are used from the device)");
    FRPA_REQUIRE(*n_my_batches_work_ == S.n_batches && *n_full_points_ == S.n_full && *n_basis_ == S.n_basis &&
                     *n_max_batch_size_ == S.nmb && *n_atoms_ == S.n_atoms,
                 FRPA_ERR_INVALID, "sizes differ from init_dfpt_device_data");
    const int n_tile = *j_atom_end_ - *j_atom_begin_ + 1;
    FRPA_REQUIRE(n_tile >= 1 && n_tile <= S.T, FRPA_ERR_INVALID,
                 "j_atom_end - j_atom_begin + 1 = %d outside [1, ATOM_TILE_SIZE = %d]", n_tile, S.T);
    hipStream_t s = S.stream;
    const size_t nb = S.n_batches, nmb = S.nmb, nf = S.n_full, T = S.T, nbas = S.n_basis;
    const size_t wl = S.wave.n;
    auto up = [&](DevBuf<double> &b, const double *h, size_t n) {
      b.alloc(std::max<size_t>(n, 1));
      if (n)
        FRPA_HIP(hipMemcpyAsync(b.ptr, h, n * 8, hipMemcpyHostToDevice, s));
    };
    up(S.h_hw, hessian_wave_batches_compress_ptr, 6 * wl);
    up(S.h_vs, vsigma_batches_ptr, 3 * nmb * nb);
    up(S.h_v2rr, v2rho2_batches_ptr, 3 * nmb * nb);
    up(S.h_v2rs, v2rhosigma_batches_ptr, 6 * nmb * nb);
    up(S.h_v2ss, v2sigma2_batches_ptr, 6 * nmb * nb);
    up(S.h_grho, gradient_rho_ptr, 3 * nf);
    up(S.h_pref1, H_prefactor1_batches_tile_ptr, nmb * nb * T);
    up(S.h_grho1, first_order_gradient_rho_batches_tile_ptr, 3 * nmb * nb * T);
    up(S.h_vhg, v_hartree_gradient_ptr, nf * T);
    up(S.h_rho1, first_order_rho_ptr, nf * T);
    up(S.h_w, partition_tab_ptr, nf);
    up(S.h_H, first_order_H_tile_ptr, nbas * nbas * T);
    const int rc = frpa_rho1_h1(reinterpret_cast<frpa_rho1_plan *>(S.rho1.get()), S.T, *j_atom_begin_, n_tile,
                                *j_coord_, S.n_atoms, S.avn.ptr, S.wave.ptr, S.gwave.ptr, S.h_hw.ptr, S.h_vs.ptr,
                                S.h_grho.ptr, S.h_pref1.ptr, S.h_grho1.ptr, S.h_vhg.ptr, S.h_rho1.ptr, S.h_v2rr.ptr,
                                S.h_v2rs.ptr, S.h_v2ss.ptr, S.h_w.ptr, S.h_H.ptr, s);
    FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
    FRPA_HIP(hipMemcpyAsync(first_order_H_tile_ptr, S.h_H.ptr, nbas * nbas * T * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipStreamSynchronize(s));
  });
}

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
  run_or_die("sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.data && S.far, FRPA_ERR_STATE, "init_dfpt_device_data (with hartree centres) must be called first");
    FarPlan &P = *S.far;
    FRPA_REQUIRE(l_pot_max == P.l_pot_max && n_atoms == P.n_atoms &&
                     n_centers_hartree_potential == P.n_hc,
                 FRPA_ERR_INVALID, "sizes differ from init_dfpt_device_data");
    const int n_tile = j_atom_end - j_atom_begin + 1;
    const int T = S.T;
    FRPA_REQUIRE(n_tile >= 1 && n_tile <= T, FRPA_ERR_INVALID, "j_atom tile %d outside [1, %d]", n_tile, T);
    hipStream_t s = S.stream;
    const int lp1 = l_pot_max + 1;
    S.f_cp.alloc((size_t)3 * n_full_points);
    S.f_map.alloc(std::max(n_valid_points, 1));
    S.f_lmfd.alloc((size_t)n_atoms * T);
    S.f_mrsq.alloc((size_t)n_atoms * T);
    S.f_opr.alloc((size_t)lp1 * n_atoms * T);
    S.f_M.alloc((size_t)lp1 * lp1 * n_atoms * T);
    S.f_dv.alloc((size_t)n_full_points * T);
    FRPA_HIP(hipMemcpyAsync(S.f_cp.ptr, coord_points_ptr, S.f_cp.n * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_map.ptr, i_valid_point_2_i_full_points_map_ptr, (size_t)n_valid_points * 4,
                            hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_lmfd.ptr, l_hartree_max_far_distance_ptr, S.f_lmfd.n * 4, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_mrsq.ptr, multipole_radius_sq_ptr, S.f_mrsq.n * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_opr.ptr, outer_potential_radius_ptr, S.f_opr.n * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_M.ptr, multipole_moments_ptr, S.f_M.n * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(S.f_dv.ptr, delta_v_hartree_ptr, S.f_dv.n * 8, hipMemcpyHostToDevice, s));
    // near field (sum_up_close_part_cu): this call's per-species log grids and spline tile
    const size_t n_spl = (size_t)lp1 * lp1 * 2 * n_hartree_grid * n_atoms * T;
    FRPA_REQUIRE(species_center_ptr && n_grid_ptr && r_grid_min_ptr && log_r_grid_inc_ptr &&
                     current_delta_v_hart_part_spl_tile_ptr && n_hartree_grid >= 2 && n_species == P.n_species,
                 FRPA_ERR_INVALID, "near field: grid arrays, spline tile and n_hartree_grid >= 2 are required");
    set_near_tables(P, n_species, species_center_ptr, n_grid_ptr, r_grid_min_ptr, log_r_grid_inc_ptr);
    S.f_spl.alloc(n_spl);
    FRPA_HIP(hipMemcpyAsync(S.f_spl.ptr, current_delta_v_hart_part_spl_tile_ptr, n_spl * 8, hipMemcpyHostToDevice, s));
    const char *v = getenv("FRPA_FAR_VARIANT");
    const int variant = (v && strcmp(v, "terms") == 0 && P.terms_ok) ? FRPA_FAR_TERMS : FRPA_FAR_HORNER;
    FarCall c;
    c.T = T;
    c.n_tile = n_tile;
    c.n_full = n_full_points;
    c.n_valid = n_valid_points;
    c.coord_points = S.f_cp.ptr;
    c.map = S.f_map.ptr;
    c.lmfd = S.f_lmfd.ptr;
    c.mrsq = S.f_mrsq.ptr;
    c.opr = S.f_opr.ptr;
    c.moments = S.f_M.ptr;
    c.spl = S.f_spl.ptr;
    c.n_hgrid = n_hartree_grid;
    near_forward(P, c, S.f_dv.ptr, s);
    far_forward(P, c, variant, S.f_dv.ptr, s);
    FRPA_HIP(hipMemcpyAsync(delta_v_hartree_ptr, S.f_dv.ptr, S.f_dv.n * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipStreamSynchronize(s));
  });
}

} // extern "C"
