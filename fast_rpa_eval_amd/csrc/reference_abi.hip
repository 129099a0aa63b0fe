// Drop-in entry points with the reference's names and Fortran-ABI argument lists.
//
// The reference keeps device copies of every array in process-global singletons (devPs, devInfo;
// fast_rap_src_dfpt_gpu_eval/device_data.cu:6-8, device.cu:5), filled once by
// init_dfpt_device_data and used by the *_cu_host_ functions.  The same lifecycle is kept here:
// init_dfpt_device_info creates the stream, init_dfpt_device_data builds the rho1 and far-field
// plans and uploads the constant arrays; the *_cu_host_ functions upload their per-call arrays,
// run on the device and copy the results back -- or, in the device-resident mode that the reference
// selects with ALL_KERNELS_ON_GPU (evaluate_first_order_rho.cu:17,259-263,363-367), the rho1
// function reads its per-call inputs from and leaves its outputs in device arrays that the caller
// reaches through frpa_ref_device_array.  The reference's CPU functions on the same path
// (evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_, sum_up_whole_potential_c_v3_
// atoms_full_points_j_atom_tile_, evaluate_first_order_h_reduce_memory_c_v2_,
// generate_first_order_rho_bias_part2_) run on the GPU too.  Errors print and exit(1), as DEV_CHECK
// (device.hpp:19-42).  The implementations live in namespace frpa_ref; the C-linkage symbols are
// defined below, the C++-linkage ones in reference_abi_cxx.hip.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "farfield_common.hpp"
#include "reference_impl.hpp"
#include "rho1_common.hpp"

namespace {

using namespace frpa;

struct RefState {
  bool info = false;
  hipStream_t stream = nullptr;
  bool data = false;
  bool resident = false;  // ALL_KERNELS_ON_GPU: rho1 per-call arrays stay on the device
  int T = 1;
  int n_batches = 0, n_full = 0, n_basis = 0, nmb = 0, n_valid_points = 0, n_atoms = 0;
  std::unique_ptr<Rho1Plan> rho1;
  std::unique_ptr<FarPlan> far;
  DevBuf<double> wave, gwave, P1, bias, gbias, rho, grad, dm;
  DevBuf<double> dm0;       // density_matrix of init_dfpt_device_data (device_data.cu:328)
  DevBuf<double> bias_inl;  // bias input plus the v6 inline bias
  // per-call arrays of the first-order H entry point
  DevBuf<double> h_hw, h_vs, h_v2rr, h_v2rs, h_v2ss, h_grho, h_pref1, h_grho1, h_vhg, h_rho1, h_w, h_H, h_vr;
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

// State of the GPU-backed CPU symbols, which need no init_dfpt_* call: their own stream, plans kept
// while the metadata they were built from is unchanged (a 64-bit hash of it), and buffers for the
// arrays passed on every call.
struct CpuSymState {
  hipStream_t stream = nullptr;
  uint64_t rho_key = 0, far_key = 0;
  std::unique_ptr<Rho1Plan> rho1;
  std::unique_ptr<FarPlan> far;
  DevBuf<double> wave, gwave, P1, bias, gbias, rho, grad, dm;
  DevBuf<int> avn;
  DevBuf<double> f_cp, f_mrsq, f_opr, f_M, f_dv, f_spl;
  DevBuf<int> f_map, f_lmfd;
};

CpuSymState &cpu_sym() {
  static CpuSymState s;
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

// FRPA_RHO1_PRECISION: ozaki (INT8 split with 5 digits, ~1e-12 of max|rho1| and of max|grad rho1|;
// the default), ozaki4 (4 digits, ~1e-9), fp64 (FP64 MFMA, the full product as the reference),
// fp64sym (FP64 MFMA on the symmetric half of P1 when the gradient is off, else the full product).
// FRPA_RHO1_GRADIENT=0: first_order_gradient_rho is not computed (left zero; LDA callers), which lets
// the INT8 split and fp64sym multiply only the symmetric half of P1.
int rho1_precision_from_env(bool want_grad) {
  const char *p = getenv("FRPA_RHO1_PRECISION");
  if (p && strcmp(p, "fp64") == 0)
    return FRPA_PREC_FP64;
  if (p && strcmp(p, "fp64sym") == 0)
    return want_grad ? FRPA_PREC_FP64 : FRPA_PREC_FP64_SYM;
  if (p && strcmp(p, "ozaki4") == 0)
    return FRPA_PREC_OZAKI4;
  return FRPA_PREC_OZAKI;
}

bool rho1_gradient_from_env() {
  const char *g = getenv("FRPA_RHO1_GRADIENT");
  return !(g && strcmp(g, "0") == 0);
}

// FRPA_RHO1_INLINE_BIAS=1: the rho1 entry points compute as the v6 variant of the reference's rho1
// function (task4_common_calculation_eliminated/version6_rho_with_bias_calculate/
// evaluate_first_order_rho.cpp:179-231, .cu.hpp:202-250), which adds to rho1 the inline bias
// -sum_{i in V} grad_{j_coord} phi_i sum_k DM_ik phi_k on top of the bias input, with the
// density_matrix of init_dfpt_device_data (GPU entry point) or of the call (CPU function's symbol).
bool rho1_inline_bias_from_env() {
  const char *v = getenv("FRPA_RHO1_INLINE_BIAS");
  return v && strcmp(v, "1") == 0;
}

// The reference's compile-time ATOM_TILE_SIZE (setting.h:19-20, default 1).
int tile_size_from_env() {
  const char *ts = getenv("FRPA_ATOM_TILE_SIZE");
  const int T = ts ? atoi(ts) : 1;
  FRPA_REQUIRE(T >= 1, FRPA_ERR_INVALID, "FRPA_ATOM_TILE_SIZE must be >= 1");
  return T;
}

// 64-bit hash of a byte range (multiply-xorshift over 8-byte words), to tell whether metadata
// passed again equals what a cached plan was built from.
uint64_t hash_bytes(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  auto mix = [](uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
  };
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, c + i, 8);
    h = mix(h ^ w) + 0x9e3779b97f4a7c15ULL;
  }
  uint64_t w = 0;
  memcpy(&w, c + i, n - i);
  return mix(h ^ w ^ (uint64_t)n);
}

// Dense 1-based (n_centers_basis_I, n_batches) basis lists and int32 wave offsets -> the plan's CSR
// lists and 64-bit offsets.
Rho1Plan *plan_from_reference_arrays(int nb, int n_basis, int n_full, int nmb, int nbI, const int *n_point_batches,
                                     const int *prefix, const int *map, const int *ncc, const int *i_basis_batches,
                                     const int *wave_offset) {
  std::vector<long long> ibp(nb + 1, 0), woff(nb + 1);
  std::vector<int> ib;
  for (int b = 0; b < nb; b++) {
    const int nc = ncc[b];
    FRPA_REQUIRE(nc >= 0 && nc <= nbI, FRPA_ERR_INVALID, "n_compute_c(%d) = %d", b, nc);
    ib.insert(ib.end(), i_basis_batches + (size_t)nbI * b, i_basis_batches + (size_t)nbI * b + nc);
    ibp[b + 1] = ibp[b] + nc;
  }
  for (int b = 0; b <= nb; b++)
    woff[b] = wave_offset[b];
  return make_rho1_plan(nb, n_basis, n_full, nmb, n_point_batches, prefix, map, ncc, ibp.data(), ib.data(),
                        woff.data());
}

// rho1 forward of a reference entry point: precision and gradient from the environment; without the
// gradient first_order_gradient_rho is zeroed (evaluate_first_order_rho.cpp:113-114).
void rho1_run(Rho1Plan &plan, int T, int n_tile, const double *wave, const double *gwave, const double *P1,
              const double *bias, const double *gbias, double *rho, double *grad, hipStream_t s) {
  const bool want_grad = rho1_gradient_from_env();
  const int prec = rho1_precision_from_env(want_grad);
  const int rc = frpa_rho1_forward(reinterpret_cast<frpa_rho1_plan *>(&plan), T, n_tile, prec, wave, P1, bias, rho,
                                   want_grad ? gwave : nullptr, want_grad ? gbias : nullptr,
                                   want_grad ? grad : nullptr, s);
  FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
  if (!want_grad)
    fill_f64(grad, (size_t)3 * plan.nmb * plan.n_batches * T, 0.0, s);
}

// The v6 inline bias: bias_out = bias_in + inline bias of DM for the tile's atoms (device arrays);
// returns bias_out, the bias input of the forward.
const double *rho1_inline_bias(Rho1Plan &plan, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms,
                               const int *avn, const double *wave, const double *gwave, const double *DM,
                               const double *bias_in, DevBuf<double> &bias_out, hipStream_t s) {
  const size_t nbt = (size_t)plan.nmb * plan.n_batches * T;
  bias_out.grow(nbt);
  FRPA_HIP(hipMemcpyAsync(bias_out.ptr, bias_in, nbt * 8, hipMemcpyDeviceToDevice, s));
  const int rc = frpa_rho1_bias_inline(reinterpret_cast<frpa_rho1_plan *>(&plan), T, j_atom_begin, n_tile, j_coord,
                                       n_atoms, avn, wave, gwave, DM, bias_out.ptr, s);
  FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
  return bias_out.ptr;
}

// Per-call far/near arrays of one Hartree-sum call in device buffers.
struct SumUpBufs {
  DevBuf<double> *cp, *mrsq, *opr, *M, *dv, *spl;
  DevBuf<int> *map, *lmfd;
};

// Both branches of the Hartree sum (sum_up_whole_potential.cpp:195-529 / .cu:548-770): uploads the
// per-call arrays, near-field spline sum then far-field multipole sum accumulated into
// delta_v_hartree, copies it back and synchronises.
void sumup_run(FarPlan &P, const SumUpBufs &B, int T, int n_tile, int n_full_points, int n_valid_points,
               int l_pot_max, int n_hartree_grid, int n_atoms, int n_species, const int *species_center_ptr,
               const int *l_hartree_max_far_distance_ptr, const int *i_valid_point_2_i_full_points_map_ptr,
               const double *coord_points_ptr, const double *multipole_radius_sq_ptr,
               const double *outer_potential_radius_ptr, const double *multipole_moments_ptr,
               const double *current_delta_v_hart_part_spl_tile_ptr, const int *n_grid_ptr,
               const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, double *delta_v_hartree_ptr,
               hipStream_t s) {
  FRPA_REQUIRE(n_tile >= 1 && n_tile <= T, FRPA_ERR_INVALID, "j_atom tile %d outside [1, %d]", n_tile, T);
  const int lp1 = l_pot_max + 1;
  B.cp->grow((size_t)3 * n_full_points);
  B.map->grow(std::max(n_valid_points, 1));
  B.lmfd->grow((size_t)n_atoms * T);
  B.mrsq->grow((size_t)n_atoms * T);
  B.opr->grow((size_t)lp1 * n_atoms * T);
  B.M->grow((size_t)lp1 * lp1 * n_atoms * T);
  B.dv->grow((size_t)n_full_points * T);
  FRPA_HIP(hipMemcpyAsync(B.cp->ptr, coord_points_ptr, (size_t)3 * n_full_points * 8, hipMemcpyHostToDevice, s));
  FRPA_HIP(hipMemcpyAsync(B.map->ptr, i_valid_point_2_i_full_points_map_ptr, (size_t)n_valid_points * 4,
                          hipMemcpyHostToDevice, s));
  FRPA_HIP(hipMemcpyAsync(B.lmfd->ptr, l_hartree_max_far_distance_ptr, (size_t)n_atoms * T * 4, hipMemcpyHostToDevice,
                          s));
  FRPA_HIP(hipMemcpyAsync(B.mrsq->ptr, multipole_radius_sq_ptr, (size_t)n_atoms * T * 8, hipMemcpyHostToDevice, s));
  FRPA_HIP(hipMemcpyAsync(B.opr->ptr, outer_potential_radius_ptr, (size_t)lp1 * n_atoms * T * 8,
                          hipMemcpyHostToDevice, s));
  FRPA_HIP(hipMemcpyAsync(B.M->ptr, multipole_moments_ptr, (size_t)lp1 * lp1 * n_atoms * T * 8,
                          hipMemcpyHostToDevice, s));
  FRPA_HIP(hipMemcpyAsync(B.dv->ptr, delta_v_hartree_ptr, (size_t)n_full_points * T * 8, hipMemcpyHostToDevice, s));
  // near field (sum_up_close_part_cu): this call's per-species log grids and spline tile
  const size_t n_spl = (size_t)lp1 * lp1 * 2 * n_hartree_grid * n_atoms * T;
  FRPA_REQUIRE(species_center_ptr && n_grid_ptr && r_grid_min_ptr && log_r_grid_inc_ptr &&
                   current_delta_v_hart_part_spl_tile_ptr && n_hartree_grid >= 2 && n_species == P.n_species,
               FRPA_ERR_INVALID, "near field: grid arrays, spline tile and n_hartree_grid >= 2 are required");
  set_near_tables(P, n_species, species_center_ptr, n_grid_ptr, r_grid_min_ptr, log_r_grid_inc_ptr);
  B.spl->grow(n_spl);
  FRPA_HIP(hipMemcpyAsync(B.spl->ptr, current_delta_v_hart_part_spl_tile_ptr, n_spl * 8, hipMemcpyHostToDevice, s));
  const char *v = getenv("FRPA_FAR_VARIANT");
  const int variant = (v && strcmp(v, "terms") == 0 && P.terms_ok) ? FRPA_FAR_TERMS : FRPA_FAR_HORNER;
  FarCall c;
  c.T = T;
  c.n_tile = n_tile;
  c.n_full = n_full_points;
  c.n_valid = n_valid_points;
  c.coord_points = B.cp->ptr;
  c.map = B.map->ptr;
  c.lmfd = B.lmfd->ptr;
  c.mrsq = B.mrsq->ptr;
  c.opr = B.opr->ptr;
  c.moments = B.M->ptr;
  c.spl = B.spl->ptr;
  c.n_hgrid = n_hartree_grid;
  near_forward(P, c, B.dv->ptr, s);
  far_forward(P, c, variant, B.dv->ptr, s);
  FRPA_HIP(hipMemcpyAsync(delta_v_hartree_ptr, B.dv->ptr, (size_t)n_full_points * T * 8, hipMemcpyDeviceToHost, s));
  FRPA_HIP(hipStreamSynchronize(s));
}

hipStream_t cpu_sym_stream() {
  CpuSymState &C = cpu_sym();
  if (!C.stream) {
    int n = 0;
    FRPA_HIP(hipGetDeviceCount(&n));
    FRPA_REQUIRE(n > 0, FRPA_ERR_HIP, "no HIP device");
    FRPA_HIP(hipStreamCreate(&C.stream));
  }
  return C.stream;
}

} // namespace

namespace frpa_ref {

void init_device_info() {
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

void free_device_info() {
  RefState &S = ref();
  if (S.stream)
    (void)hipStreamDestroy(S.stream);
  S.stream = nullptr;
  S.info = false;
}

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
    double *multipole_moments_ptr, double *current_delta_v_hart_part_spl_tile_ptr, double *delta_v_hartree_ptr) {
  run_or_die("init_dfpt_device_data", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.info, FRPA_ERR_STATE, "init_dfpt_device_info must be called first");
    S.T = tile_size_from_env();
    const char *res = getenv("FRPA_ALL_KERNELS_ON_GPU");
    if (res && res[0])
      S.resident = res[0] == '1';
    const int nb = *n_my_batches_work_, nbI = *n_centers_basis_I_;
    S.n_batches = nb;
    S.n_full = *n_full_points_;
    S.n_basis = *n_basis_;
    S.nmb = *n_max_batch_size_;
    S.n_valid_points = *n_valid_points_;
    S.n_atoms = *n_atoms_;
    S.rho1.reset(plan_from_reference_arrays(nb, S.n_basis, S.n_full, S.nmb, nbI, n_point_batches_ptr,
                                            n_point_batches_prefix_sum_ptr, i_valid_point_2_i_full_points_map_ptr,
                                            n_compute_c_batches_ptr, i_basis_batches_ptr, i_batch_2_wave_offset_ptr));
    const size_t wl = nb ? (size_t)i_batch_2_wave_offset_ptr[nb] : 0;
    S.wave.upload(wave_batches_compress_ptr, wl);
    S.gwave.upload(gradient_wave_batches_compress_ptr, 3 * wl);
    S.avn.upload(atom_valid_n_compute_c_batches_ptr, (size_t)(S.n_atoms + 1) * nb);
    S.P1.alloc((size_t)S.n_basis * S.n_basis * S.T);
    S.bias.alloc((size_t)S.nmb * nb * S.T);
    S.gbias.alloc((size_t)3 * S.nmb * nb * S.T);
    S.rho.alloc((size_t)S.n_full * S.T);
    S.grad.alloc((size_t)3 * S.nmb * nb * S.T);
    // as the reference's init (device_data.cu:185-187), the per-call arrays are allocated, not copied;
    // the density matrix is copied (device_data.cu:328), for the v6 inline bias
    if (density_matrix_ptr && S.n_basis > 0)
      S.dm0.upload(density_matrix_ptr, (size_t)S.n_basis * S.n_basis);
    S.coord_points.upload(coord_points_ptr, (size_t)3 * S.n_full);
    if (*n_centers_hartree_potential_ > 0 && *n_atoms_ > 0)
      S.far.reset(make_far_plan(*l_pot_max_, *l_max_analytic_multipole_, *index_cc_dim_0_, *n_centers_,
                                *n_centers_hartree_potential_, *n_atoms_, *n_species_, *hartree_force_l_add_,
                                species_ptr, l_hartree_ptr, centers_hartree_potential_ptr, center_to_atom_ptr,
                                coords_center_ptr, index_cc_ptr, cc_ptr, n_cc_lm_ijk_ptr, index_ijk_max_cc_ptr));
    S.data = true;
  });
}

void free_device_data() {
  RefState &S = ref();
  S.rho1.reset();
  S.far.reset();
  for (DevBuf<double> *b : {&S.f_spl, &S.wave, &S.gwave, &S.P1, &S.bias, &S.gbias, &S.dm, &S.dm0, &S.bias_inl,
                            &S.h_hw, &S.h_vs,
                            &S.h_v2rr, &S.h_v2rs, &S.h_v2ss, &S.h_grho, &S.h_pref1, &S.h_grho1, &S.h_vhg, &S.h_rho1,
                            &S.h_w, &S.h_H, &S.h_vr, &S.rho, &S.grad, &S.coord_points, &S.f_cp, &S.f_mrsq, &S.f_opr,
                            &S.f_M,
                            &S.f_dv})
    b->release();
  S.avn.release();
  S.f_map.release();
  S.f_lmfd.release();
  S.data = false;
}

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
    const double *partition_tab_ptr, double *delta_v_hartree_ptr) {
  run_or_die("sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_", [&] {
    RefState &S = ref();
    FRPA_REQUIRE(S.data && S.far, FRPA_ERR_STATE, "init_dfpt_device_data (with hartree centres) must be called first");
    FarPlan &P = *S.far;
    FRPA_REQUIRE(l_pot_max == P.l_pot_max && n_atoms == P.n_atoms && n_centers_hartree_potential == P.n_hc,
                 FRPA_ERR_INVALID, "sizes differ from init_dfpt_device_data");
    SumUpBufs B{&S.f_cp, &S.f_mrsq, &S.f_opr, &S.f_M, &S.f_dv, &S.f_spl, &S.f_map, &S.f_lmfd};
    sumup_run(P, B, S.T, j_atom_end - j_atom_begin + 1, n_full_points, n_valid_points, l_pot_max, n_hartree_grid,
              n_atoms, n_species, species_center_ptr, l_hartree_max_far_distance_ptr,
              i_valid_point_2_i_full_points_map_ptr, coord_points_ptr, multipole_radius_sq_ptr,
              outer_potential_radius_ptr, multipole_moments_ptr, current_delta_v_hart_part_spl_tile_ptr, n_grid_ptr,
              r_grid_min_ptr, log_r_grid_inc_ptr, delta_v_hartree_ptr, S.stream);
  });
}

} // namespace frpa_ref

extern "C" {

void init_dfpt_device_info(void) { frpa_ref::init_device_info(); }

void free_dfpt_device_info(void) { frpa_ref::free_device_info(); }

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

void free_dfpt_device_data(void) { frpa_ref::free_device_data(); }

void frpa_ref_set_all_kernels_on_gpu(int on) { ref().resident = on != 0; }

double *frpa_ref_device_array(const char *name) {
  RefState &S = ref();
  if (!S.data || !name)
    return nullptr;
  const struct {
    const char *n;
    DevBuf<double> *b;
  } tab[] = {{"first_order_density_matrix", &S.P1},
             {"first_order_rho_bias_part2", &S.bias},
             {"first_order_gradient_rho_bias_batches_atoms", &S.gbias},
             {"first_order_rho", &S.rho},
             {"first_order_gradient_rho", &S.grad},
             {"wave_batches_compress", &S.wave},
             {"gradient_wave_batches_compress", &S.gwave},
             {"density_matrix", &S.dm0},
             {"coord_points", &S.coord_points}};
  for (const auto &e : tab)
    if (strcmp(name, e.n) == 0)
      return e.b->ptr;
  return nullptr;
}

void *frpa_ref_stream(void) { return ref().stream; }

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
    if (!S.resident) {  // evaluate_first_order_rho.cu:259-263
      FRPA_HIP(hipMemcpyAsync(S.P1.ptr, first_order_density_matrix_ptr, S.P1.n * 8, hipMemcpyHostToDevice, s));
      FRPA_HIP(hipMemcpyAsync(S.bias.ptr, first_order_rho_bias_part2_ptr, nbt * 8, hipMemcpyHostToDevice, s));
      FRPA_HIP(hipMemcpyAsync(S.gbias.ptr, first_order_gradient_rho_bias_batches_atoms_ptr, 3 * nbt * 8,
                              hipMemcpyHostToDevice, s));
    }
    const double *bias = S.bias.ptr;
    if (rho1_inline_bias_from_env()) {
      FRPA_REQUIRE(S.dm0.ptr, FRPA_ERR_STATE, "FRPA_RHO1_INLINE_BIAS=1 needs the density_matrix of init_dfpt_device_data");
      bias = rho1_inline_bias(*S.rho1, S.T, *j_atom_begin_, n_tile, *j_coord_, S.n_atoms, S.avn.ptr, S.wave.ptr,
                              S.gwave.ptr, S.dm0.ptr, S.bias.ptr, S.bias_inl, s);
    }
    rho1_run(*S.rho1, S.T, n_tile, S.wave.ptr, S.gwave.ptr, S.P1.ptr, bias, S.gbias.ptr, S.rho.ptr, S.grad.ptr, s);
    if (!S.resident) {  // evaluate_first_order_rho.cu:363-367; resident: results stay, no synchronisation
      FRPA_HIP(hipMemcpyAsync(first_order_rho_ptr, S.rho.ptr, S.rho.n * 8, hipMemcpyDeviceToHost, s));
      FRPA_HIP(hipMemcpyAsync(first_order_gradient_rho_ptr, S.grad.ptr, 3 * nbt * 8, hipMemcpyDeviceToHost, s));
      FRPA_HIP(hipStreamSynchronize(s));
    }
  });
}

void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_density_matrix_ptr, double *first_order_rho_ptr,
    double *first_order_rho_bias_part2_ptr, double *first_order_gradient_rho_ptr,
    double *first_order_gradient_rho_bias_batches_atoms_ptr, int *basis_atom_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr) {
  run_or_die("evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_", [&] {
    CpuSymState &C = cpu_sym();
    hipStream_t s = cpu_sym_stream();
    const int nb = *n_my_batches_work_, n_full = *n_full_points_, n_basis = *n_basis_, nmb = *n_max_batch_size_;
    const int nbI = *n_centers_basis_I_;
    const int T = tile_size_from_env();
    const int n_tile = *j_atom_end_ - *j_atom_begin_ + 1;
    FRPA_REQUIRE(nb >= 0 && n_full >= 0 && n_basis >= 0 && nmb >= 0 && nbI >= 0, FRPA_ERR_INVALID, "bad sizes");
    FRPA_REQUIRE(n_tile >= 1 && n_tile <= T, FRPA_ERR_INVALID,
                 "j_atom_end - j_atom_begin + 1 = %d outside [1, ATOM_TILE_SIZE = %d]", n_tile, T);
    // the plan is rebuilt when the batch metadata differs from the previous call's
    uint64_t key = hash_bytes(0, n_point_batches_ptr, (size_t)nb * 4);
    const int sizes[5] = {nb, n_full, n_basis, nmb, nbI};
    key = hash_bytes(key, sizes, sizeof(sizes));
    key = hash_bytes(key, n_point_batches_prefix_sum_ptr, (size_t)(nb + 1) * 4);
    key = hash_bytes(key, n_compute_c_batches_ptr, (size_t)nb * 4);
    key = hash_bytes(key, i_batch_2_wave_offset_ptr, (size_t)(nb + 1) * 4);
    if (nb)
      key = hash_bytes(key, i_valid_point_2_i_full_points_map_ptr, (size_t)n_point_batches_prefix_sum_ptr[nb] * 4);
    for (int b = 0; b < nb; b++)
      key = hash_bytes(key, i_basis_batches_ptr + (size_t)nbI * b,
                       (size_t)std::max(0, std::min(n_compute_c_batches_ptr[b], nbI)) * 4);
    if (!C.rho1 || key != C.rho_key) {
      C.rho1.reset(plan_from_reference_arrays(nb, n_basis, n_full, nmb, nbI, n_point_batches_ptr,
                                              n_point_batches_prefix_sum_ptr, i_valid_point_2_i_full_points_map_ptr,
                                              n_compute_c_batches_ptr, i_basis_batches_ptr,
                                              i_batch_2_wave_offset_ptr));
      C.rho_key = key;
    }
    // every array is read afresh, as the CPU function reads it (evaluate_first_order_rho.cpp:113-209)
    const size_t wl = nb ? (size_t)i_batch_2_wave_offset_ptr[nb] : 0;
    const size_t nbt = (size_t)nmb * nb * T;
    C.wave.grow(wl);
    C.gwave.grow(3 * wl);
    C.P1.grow((size_t)n_basis * n_basis * T);
    C.bias.grow(nbt);
    C.gbias.grow(3 * nbt);
    C.rho.grow((size_t)n_full * T);
    C.grad.grow(3 * nbt);
    FRPA_HIP(hipMemcpyAsync(C.wave.ptr, wave_batches_compress_ptr, wl * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(C.gwave.ptr, gradient_wave_batches_compress_ptr, 3 * wl * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(C.P1.ptr, first_order_density_matrix_ptr, (size_t)n_basis * n_basis * T * 8,
                            hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(C.bias.ptr, first_order_rho_bias_part2_ptr, nbt * 8, hipMemcpyHostToDevice, s));
    FRPA_HIP(hipMemcpyAsync(C.gbias.ptr, first_order_gradient_rho_bias_batches_atoms_ptr, 3 * nbt * 8,
                            hipMemcpyHostToDevice, s));
    if (rho1_inline_bias_from_env()) {  // version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:179-231
      const int n_atoms = *n_atoms_;
      FRPA_REQUIRE(density_matrix_ptr && atom_valid_n_compute_c_batches_ptr && n_atoms >= 1, FRPA_ERR_INVALID,
                   "FRPA_RHO1_INLINE_BIAS=1 needs density_matrix and atom_valid_n_compute_c_batches");
      C.dm.grow((size_t)n_basis * n_basis);
      C.avn.grow((size_t)(n_atoms + 1) * nb);
      FRPA_HIP(hipMemcpyAsync(C.dm.ptr, density_matrix_ptr, (size_t)n_basis * n_basis * 8, hipMemcpyHostToDevice, s));
      FRPA_HIP(hipMemcpyAsync(C.avn.ptr, atom_valid_n_compute_c_batches_ptr, (size_t)(n_atoms + 1) * nb * 4,
                              hipMemcpyHostToDevice, s));
      // accumulated into C.bias, the device copy of this call's bias input
      const int rc = frpa_rho1_bias_inline(reinterpret_cast<frpa_rho1_plan *>(C.rho1.get()), T, *j_atom_begin_,
                                           n_tile, *j_coord_, n_atoms, C.avn.ptr, C.wave.ptr, C.gwave.ptr, C.dm.ptr,
                                           C.bias.ptr, s);
      FRPA_REQUIRE(rc == FRPA_OK, rc, "%s", frpa_last_error());
    }
    rho1_run(*C.rho1, T, n_tile, C.wave.ptr, C.gwave.ptr, C.P1.ptr, C.bias.ptr, C.gbias.ptr, C.rho.ptr, C.grad.ptr, s);
    FRPA_HIP(hipMemcpyAsync(first_order_rho_ptr, C.rho.ptr, (size_t)n_full * T * 8, hipMemcpyDeviceToHost, s));
    FRPA_HIP(hipMemcpyAsync(first_order_gradient_rho_ptr, C.grad.ptr, 3 * nbt * 8, hipMemcpyDeviceToHost, s));
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
      b.grow(n);
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

void evaluate_first_order_h_bias_part2_(
    int *j_atom_begin_, int *j_atom_end_, int *n_full_points_, int *j_coord_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_max_batch_size_, int *n_my_batches_work_, int *n_centers_basis_I_,
    int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, double *vrho_batches_ptr, double *vsigma_batches_ptr,
    double *gradient_rho_ptr, double *H_times_psi_batches_compress_ptr, int *i_batch_2_wave_offset_ptr,
    double *wave_batches_compress_ptr, double *gradient_wave_batches_compress_ptr, double *first_order_H_tile_ptr,
    double *H_prefactor1_batches_tile_ptr, double *first_order_gradient_rho_batches_tile_ptr,
    double *v_hartree_gradient_ptr, double *first_order_rho_ptr, double *v2rho2_batches_ptr,
    double *v2rhosigma_batches_ptr, double *v2sigma2_batches_ptr, double *partition_tab_ptr, int *not_use) {
  run_or_die("evaluate_first_order_h_bias_part2_", [&] {
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
    auto up = [&](DevBuf<double> &b, const double *h, size_t n) {
      b.grow(n);
      if (n)
        FRPA_HIP(hipMemcpyAsync(b.ptr, h, n * 8, hipMemcpyHostToDevice, s));
    };
    up(S.h_hw, H_times_psi_batches_compress_ptr, S.wave.n);  // the hessian-wave buffer holds H_times_psi here
    up(S.h_vr, vrho_batches_ptr, nmb * nb);
    up(S.h_vs, vsigma_batches_ptr, 3 * nmb * nb);
    up(S.h_grho, gradient_rho_ptr, 3 * nf);
    up(S.h_w, partition_tab_ptr, nf);
    up(S.h_H, first_order_H_tile_ptr, nbas * nbas * T);
    const int rc = frpa_rho1_h1_bias_part2(reinterpret_cast<frpa_rho1_plan *>(S.rho1.get()), S.T, *j_atom_begin_,
                                           n_tile, *j_coord_, S.n_atoms, S.avn.ptr, S.wave.ptr, S.gwave.ptr,
                                           S.h_hw.ptr, S.h_vr.ptr, S.h_vs.ptr, S.h_grho.ptr, S.h_w.ptr, S.h_H.ptr, s);
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

void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_(
    const int *j_atom_begin_, const int *j_atom_end_, const int *n_full_points_, const int *l_max_analytic_multipole_,
    const int *index_cc_dim_0_, const int *n_valid_points_, const int *l_pot_max_, const int *n_max_radial_,
    const int *n_hartree_grid_, const int *n_centers_, const int *n_centers_hartree_potential_, const int *n_atoms_,
    const int *hartree_force_l_add_, const int *n_species_, const int *n_radial_ptr, const int *species_ptr,
    const int *species_center_ptr, const int *l_hartree_max_far_distance_ptr,
    const int *i_valid_point_2_i_full_points_map_ptr, const double *coord_points_ptr,
    const double *multipole_radius_sq_ptr, const double *outer_potential_radius_ptr,
    const double *multipole_moments_ptr, const double *current_delta_v_hart_part_spl_tile_ptr,
    const int *l_hartree_ptr, const int *n_grid_ptr, const int *n_cc_lm_ijk_ptr,
    const int *centers_hartree_potential_ptr, const int *center_to_atom_ptr, const int *index_cc_ptr,
    const int *index_ijk_max_cc_ptr, const double *cc_ptr, const double *coords_center_ptr,
    const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, const double *scale_radial_ptr,
    const double *partition_tab_ptr, double *delta_v_hartree_ptr) {
  run_or_die("sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_", [&] {
    CpuSymState &C = cpu_sym();
    hipStream_t s = cpu_sym_stream();
    const int T = tile_size_from_env();
    const int l_pot_max = *l_pot_max_, lama = *l_max_analytic_multipole_, dim0 = *index_cc_dim_0_;
    const int n_centers = *n_centers_, n_hc = *n_centers_hartree_potential_, n_atoms = *n_atoms_;
    const int n_species = *n_species_, hfla = *hartree_force_l_add_;
    FRPA_REQUIRE(n_centers >= 0 && n_hc >= 0 && n_atoms >= 1 && n_species >= 1 && lama >= 0 && dim0 >= 1,
                 FRPA_ERR_INVALID, "bad sizes");
    // the plan (centre records, term tables) is rebuilt when the geometry or tables differ
    const int sizes[7] = {l_pot_max, lama, dim0, n_centers, n_hc, n_atoms, n_species};
    uint64_t key = hash_bytes(0, sizes, sizeof(sizes));
    key = hash_bytes(key, &hfla, 4);
    key = hash_bytes(key, species_ptr, (size_t)n_atoms * 4);
    key = hash_bytes(key, l_hartree_ptr, (size_t)n_species * 4);
    key = hash_bytes(key, centers_hartree_potential_ptr, (size_t)n_hc * 4);
    key = hash_bytes(key, center_to_atom_ptr, (size_t)n_centers * 4);
    key = hash_bytes(key, coords_center_ptr, (size_t)3 * n_centers * 8);
    key = hash_bytes(key, index_cc_ptr, (size_t)dim0 * 6 * 4);
    key = hash_bytes(key, cc_ptr, (size_t)dim0 * 8);
    key = hash_bytes(key, n_cc_lm_ijk_ptr, (size_t)(lama + 1) * 4);
    key = hash_bytes(key, index_ijk_max_cc_ptr, (size_t)3 * (lama + 1) * 4);
    if (!C.far || key != C.far_key) {
      C.far.reset(make_far_plan(l_pot_max, lama, dim0, n_centers, n_hc, n_atoms, n_species, hfla, species_ptr,
                                l_hartree_ptr, centers_hartree_potential_ptr, center_to_atom_ptr, coords_center_ptr,
                                index_cc_ptr, cc_ptr, n_cc_lm_ijk_ptr, index_ijk_max_cc_ptr));
      C.far_key = key;
    }
    SumUpBufs B{&C.f_cp, &C.f_mrsq, &C.f_opr, &C.f_M, &C.f_dv, &C.f_spl, &C.f_map, &C.f_lmfd};
    sumup_run(*C.far, B, T, *j_atom_end_ - *j_atom_begin_ + 1, *n_full_points_, *n_valid_points_, l_pot_max,
              *n_hartree_grid_, n_atoms, n_species, species_center_ptr, l_hartree_max_far_distance_ptr,
              i_valid_point_2_i_full_points_map_ptr, coord_points_ptr, multipole_radius_sq_ptr,
              outer_potential_radius_ptr, multipole_moments_ptr, current_delta_v_hart_part_spl_tile_ptr, n_grid_ptr,
              r_grid_min_ptr, log_r_grid_inc_ptr, delta_v_hartree_ptr, s);
  });
}

} // extern "C"
