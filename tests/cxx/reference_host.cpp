// A host program written against the reference's own declarations -- C++ linkage where the reference
// declares C++ linkage (fast_rap_src_dfpt_gpu_eval/device.hpp:191-192, device_data.hpp:159-227,
// sum_up_whole_potential.h:50-95), extern "C" where it declares C linkage
// (evaluate_first_order_rho.h:3-69, sum_up_whole_potential.h:3-48) -- linked against libfrpa.so.
// It runs the calls of the reference's direct-test runners
// (direct_test/evaluate_first_order_rho_run.cpp, sum_up_run.cpp) on inputs read from a directory and
// writes the outputs there for tests/test_gpu_reference_host.py to check against the oracle.
//
//   reference_host <dir>          run (needs a GPU)
//   reference_host --link-only    load and resolve every symbol, then exit (no GPU call)
//
// <dir> holds scalars.txt ("name value" lines) and one raw little-endian file per array, <name>.bin.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

// ---- the reference's declarations (its names, parameter types and linkage) ----
void init_dfpt_device_info();
void free_dfpt_device_info();
void init_dfpt_device_data(int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
                           int *, int *, int *, int *,                                             // 17 scalars
                           int *, int *, int *, int *, int *, int *, int *, int *, int *, double *, double *,
                           double *,                                                               // 12 arrays
                           int *, int *, int *, double *, int *, int *, int *, int *, int *, int *, int *,
                           double *, double *, double *, double *, double *, double *,             // 17 arrays
                           double *, double *, double *, double *, double *,                       // changeable
                           int *, double *, double *, double *, double *, double *);
void free_dfpt_device_data();
void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_(
    const int, const int, const int, const int, const int, const int, const int, const int, const int, const int,
    const int, const int, const int, const int, const int *, const int *, const int *, const int *, const int *,
    const double *, const double *, const double *, const double *, const double *, const int *, const int *,
    const int *, const int *, const int *, const int *, const int *, const double *, const double *, const double *,
    const double *, const double *, const double *, double *);
extern "C" void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_(
    int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
    int *, int *, double *, double *, double *, double *, double *, double *, double *, int *, double *, double *);
extern "C" void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_(
    int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *, int *,
    int *, int *, double *, double *, double *, double *, double *, double *, double *, int *, double *, double *);
extern "C" void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_(
    const int *, const int *, const int *, const int *, const int *, const int *, const int *, const int *,
    const int *, const int *, const int *, const int *, const int *, const int *, const int *, const int *,
    const int *, const int *, const int *, const double *, const double *, const double *, const double *,
    const double *, const int *, const int *, const int *, const int *, const int *, const int *, const int *,
    const double *, const double *, const double *, const double *, const double *, const double *, double *);
// ---- the library's device-resident mode (no counterpart declaration in the reference) ----
extern "C" void frpa_ref_set_all_kernels_on_gpu(int on);
extern "C" double *frpa_ref_device_array(const char *name);
extern "C" void *frpa_ref_stream(void);

namespace {

std::string g_dir;
std::map<std::string, long long> g_scalars;

int sc(const char *name) {
  auto it = g_scalars.find(name);
  if (it == g_scalars.end()) {
    fprintf(stderr, "scalar %s missing\n", name);
    exit(2);
  }
  return (int)it->second;
}

template <class T>
std::vector<T> load(const char *name) {
  std::ifstream f(g_dir + "/" + name + ".bin", std::ios::binary | std::ios::ate);
  if (!f) {
    fprintf(stderr, "array %s missing\n", name);
    exit(2);
  }
  const std::streamsize n = f.tellg();
  std::vector<T> v((size_t)n / sizeof(T) + 1);  // one spare element: never empty
  f.seekg(0);
  f.read(reinterpret_cast<char *>(v.data()), n);
  return v;
}

template <class T>
void save(const char *name, const std::vector<T> &v, size_t n) {
  std::ofstream f(g_dir + "/" + name + ".bin", std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(n * sizeof(T)));
}

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    exit(3);
  }
}

} // namespace

int main(int argc, char **argv) {
  if (argc == 2 && strcmp(argv[1], "--link-only") == 0) {
    // every function above is resolved at load time (linked with -z now); take their addresses
    const void *fns[] = {(void *)&init_dfpt_device_info, (void *)&free_dfpt_device_info,
                         (void *)&init_dfpt_device_data, (void *)&free_dfpt_device_data,
                         (void *)&sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_,
                         (void *)&evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_,
                         (void *)&evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_,
                         (void *)&sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_,
                         (void *)&frpa_ref_set_all_kernels_on_gpu, (void *)&frpa_ref_device_array,
                         (void *)&frpa_ref_stream};
    for (const void *f : fns)
      if (!f)
        return 1;
    printf("linked\n");
    return 0;
  }
  if (argc != 2) {
    fprintf(stderr, "usage: %s <dir> | --link-only\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  {
    std::ifstream f(g_dir + "/scalars.txt");
    std::string k;
    long long v;
    while (f >> k >> v)
      g_scalars[k] = v;
  }
  int n_batches = sc("n_my_batches_work"), n_full = sc("n_full_points"), n_valid = sc("n_valid_points");
  int n_basis = sc("n_basis"), n_atoms = sc("n_atoms"), nmch = sc("n_max_compute_ham");
  int nbI = sc("n_centers_basis_I"), nmb = sc("n_max_batch_size"), l_pot_max = sc("l_pot_max");
  int n_max_radial = sc("n_max_radial"), nhg = sc("n_hartree_grid"), n_centers = sc("n_centers");
  int n_hc = sc("n_centers_hartree_potential"), hfla = sc("hartree_force_l_add"), n_species = sc("n_species");
  int lama = sc("l_max_analytic_multipole"), dim0 = sc("index_cc_dim_0");
  const int T = sc("atom_tile_size"), jb = sc("j_atom_begin"), je = sc("j_atom_end"), jc = 1;
  const int f_full = sc("far_n_full_points"), f_valid = sc("far_n_valid_points");

  auto basis_atom = load<int>("basis_atom");
  auto batch_sizes = load<int>("batch_sizes");
  auto npb = load<int>("n_point_batches"), prefix = load<int>("n_point_batches_prefix_sum");
  auto map = load<int>("i_valid_point_2_i_full_points_map"), ncc = load<int>("n_compute_c_batches");
  auto ibasis = load<int>("i_basis_batches"), avn = load<int>("atom_valid_n_compute_c_batches");
  auto woff = load<int>("i_batch_2_wave_offset");
  auto wave = load<double>("wave_batches_compress"), gwave = load<double>("gradient_wave_batches_compress");
  auto dm = load<double>("density_matrix");
  auto n_radial = load<int>("n_radial"), species = load<int>("species"), spc = load<int>("species_center");
  auto coord_points = load<double>("coord_points");
  auto lh = load<int>("l_hartree"), n_grid = load<int>("n_grid");
  auto nccl = load<int>("n_cc_lm_ijk"), chp = load<int>("centers_hartree_potential");
  auto c2a = load<int>("center_to_atom"), icc = load<int>("index_cc"), ijk = load<int>("index_ijk_max_cc");
  auto cc = load<double>("cc"), cen = load<double>("coords_center"), rmin = load<double>("r_grid_min");
  auto linc = load<double>("log_r_grid_inc"), scale_radial = load<double>("scale_radial");
  auto ptab = load<double>("partition_tab"), P1 = load<double>("first_order_density_matrix");
  auto bias = load<double>("first_order_rho_bias_part2");
  auto gbias = load<double>("first_order_gradient_rho_bias_batches_atoms");
  auto lmfd = load<int>("l_hartree_max_far_distance");
  auto mrsq = load<double>("multipole_radius_sq");
  auto opr = load<double>("outer_potential_radius"), M = load<double>("multipole_moments");
  auto spl = load<double>("current_delta_v_hart_part_spl_tile"), dv0 = load<double>("delta_v_hartree");
  auto fcp = load<double>("far_coord_points");
  auto fmap = load<int>("far_i_valid_point_2_i_full_points_map");
  const size_t n_rho = (size_t)n_full * T, n_grad = (size_t)3 * nmb * n_batches * T, n_dv = (size_t)f_full * T;
  std::vector<double> rho(n_rho + 1, 0.0), grad(n_grad + 1, 0.0), dv = dv0;

  init_dfpt_device_info();
  init_dfpt_device_data(&n_batches, &n_full, &n_valid, &n_basis, &n_atoms, &nmch, &nbI, &nmb, &l_pot_max,
                        &n_max_radial, &nhg, &n_centers, &n_hc, &hfla, &n_species, &lama, &dim0, basis_atom.data(),
                        batch_sizes.data(), npb.data(), prefix.data(), map.data(), ncc.data(), ibasis.data(),
                        avn.data(), woff.data(), wave.data(), gwave.data(), dm.data(), n_radial.data(),
                        species.data(), spc.data(), coord_points.data(), lh.data(), n_grid.data(), nccl.data(),
                        chp.data(), c2a.data(), icc.data(), ijk.data(), cc.data(), cen.data(), rmin.data(),
                        linc.data(), scale_radial.data(), ptab.data(), P1.data(), rho.data(), bias.data(),
                        grad.data(), gbias.data(), lmfd.data(), mrsq.data(), opr.data(), M.data(), spl.data(),
                        dv.data());
  int jb_ = jb, je_ = je, jc_ = jc;
  auto rho1_args_call = [&](auto fn) {
    fn(&n_batches, &jb_, &je_, &jc_, &n_full, &n_basis, &n_atoms, &nmch, &nbI, &nmb, batch_sizes.data(), npb.data(),
       prefix.data(), map.data(), ncc.data(), ibasis.data(), avn.data(), woff.data(), wave.data(), gwave.data(),
       P1.data(), rho.data(), bias.data(), grad.data(), gbias.data(), basis_atom.data(), ptab.data(), dm.data());
  };
  // 1. GPU entry point, host arrays in and out
  rho1_args_call(evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_);
  save("out_rho_cu_host", rho, n_rho);
  save("out_grad_cu_host", grad, n_grad);
  // 2. the CPU function's symbol
  std::fill(rho.begin(), rho.end(), -7.0);
  std::fill(grad.begin(), grad.end(), -7.0);
  rho1_args_call(evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_);
  save("out_rho_cpu_symbol", rho, n_rho);
  save("out_grad_cpu_symbol", grad, n_grad);
  // 3. device-resident mode: the caller fills and reads the device arrays (as the reference's runner
  //    does through devPs, evaluate_first_order_rho_run.cpp:163-166,200-202)
  frpa_ref_set_all_kernels_on_gpu(1);
  hip_ok(hipMemcpy(frpa_ref_device_array("first_order_density_matrix"), P1.data(),
                   (size_t)n_basis * n_basis * T * sizeof(double), hipMemcpyHostToDevice),
         "upload P1");
  hip_ok(hipMemcpy(frpa_ref_device_array("first_order_rho_bias_part2"), bias.data(),
                   (size_t)nmb * n_batches * T * sizeof(double), hipMemcpyHostToDevice),
         "upload bias");
  hip_ok(hipMemcpy(frpa_ref_device_array("first_order_gradient_rho_bias_batches_atoms"), gbias.data(),
                   n_grad * sizeof(double), hipMemcpyHostToDevice),
         "upload gradient bias");
  std::fill(rho.begin(), rho.end(), -7.0);
  std::fill(grad.begin(), grad.end(), -7.0);
  rho1_args_call(evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_);
  bool untouched = true;  // host arrays are not written in this mode
  for (size_t i = 0; i < n_rho; i++)
    untouched &= rho[i] == -7.0;
  hip_ok(hipStreamSynchronize(reinterpret_cast<hipStream_t>(frpa_ref_stream())), "synchronise");
  hip_ok(hipMemcpy(rho.data(), frpa_ref_device_array("first_order_rho"), n_rho * sizeof(double),
                   hipMemcpyDeviceToHost),
         "download rho");
  hip_ok(hipMemcpy(grad.data(), frpa_ref_device_array("first_order_gradient_rho"), n_grad * sizeof(double),
                   hipMemcpyDeviceToHost),
         "download gradient");
  save("out_rho_resident", rho, n_rho);
  save("out_grad_resident", grad, n_grad);
  frpa_ref_set_all_kernels_on_gpu(0);
  // 4. Hartree sum, GPU entry point (C++ linkage, scalars by value), accumulating into delta_v
  dv = dv0;
  sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_(
      jb, je, f_full, lama, dim0, f_valid, l_pot_max, n_max_radial, nhg, n_centers, n_hc, n_atoms, hfla, n_species,
      n_radial.data(), species.data(), spc.data(), lmfd.data(), fmap.data(), fcp.data(), mrsq.data(), opr.data(),
      M.data(), spl.data(), lh.data(), n_grid.data(), nccl.data(), chp.data(), c2a.data(), icc.data(), ijk.data(),
      cc.data(), cen.data(), rmin.data(), linc.data(), scale_radial.data(), ptab.data(), dv.data());
  save("out_dv_cu_host", dv, n_dv);
  // 5. Hartree sum, the CPU function's symbol (scalars by pointer)
  dv = dv0;
  sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_(
      &jb, &je, &f_full, &lama, &dim0, &f_valid, &l_pot_max, &n_max_radial, &nhg, &n_centers, &n_hc, &n_atoms, &hfla,
      &n_species, n_radial.data(), species.data(), spc.data(), lmfd.data(), fmap.data(), fcp.data(), mrsq.data(),
      opr.data(), M.data(), spl.data(), lh.data(), n_grid.data(), nccl.data(), chp.data(), c2a.data(), icc.data(),
      ijk.data(), cc.data(), cen.data(), rmin.data(), linc.data(), scale_radial.data(), ptab.data(), dv.data());
  save("out_dv_cpu_symbol", dv, n_dv);
  free_dfpt_device_data();
  free_dfpt_device_info();
  printf("reference_host done; resident mode left the host arrays untouched: %d\n", (int)untouched);
  return untouched ? 0 : 4;
}
