/*
 * Oracle (test infrastructure only): restatement of generate_first_order_rho_bias_part2_
 * (fast_rap_src_dfpt_gpu_eval/generate_first_order_rho_bias_part2.cpp:9-167).
 *
 * For every batch b with n_compute_c > 0 and every atom j of the tile [j_atom_begin, j_atom_end]
 * (t = j - j_atom_begin), with V = compute rows [avn(j-1, b), avn(j, b)) of atom j and
 * DM(i, k) = density_matrix(i_basis(i) - 1, i_basis(k) - 1):
 *   phi'(k, p) = -grad_{j_coord} phi(k, p), k in V                                       (:108-112)
 *   bias(p, b, t)  = sum_i phi(i, p) sum_{k in V} DM(i, k) phi'(k, p)                    (:114-133)
 *                  + sum_{i in V} phi'(i, p) sum_k DM(i, k) phi(k, p)                     (:135-157)
 * and bias(:, b, t) = 0 when V is empty (:158-162); batches with n_compute_c = 0 are left as they
 * are.  The reference file fixes the tile extent of first_order_rho_bias_part2 at 4 (:7); here it
 * is the argument atom_tile_size, as everywhere in the oracle.
 * bias part 2 is the derivative of rho(p) = sum_ik phi_i DM_ik phi_k with respect to displacing
 * atom j along j_coord (the functions of atom j move with it), which is how it is pinned.
 */
#include <stdlib.h>

#include "oracle.h"

void oracle_generate_first_order_rho_bias_part2(
    int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end, int j_coord, int n_basis,
    int n_atoms, int n_centers_basis_I, int n_max_batch_size, const int *n_point_batches,
    const int *n_compute_c_batches, const int *i_basis_batches, const int *atom_valid_n_compute_c_batches,
    const int *i_batch_2_wave_offset, const double *wave_batches_compress,
    const double *gradient_wave_batches_compress, double *first_order_rho_bias_part2,
    const double *density_matrix) {
  (void)atom_tile_size;
#pragma omp parallel for schedule(dynamic, 4)
  for (int b = 0; b < n_my_batches_work; b++) {
    const int np = n_point_batches[b], nc = n_compute_c_batches[b];
    if (nc <= 0)
      continue;
    const int ld = (nc + 3) / 4 * 4;
    const double *wave = wave_batches_compress + i_batch_2_wave_offset[b];
    const double *gwave = gradient_wave_batches_compress + 3 * (size_t)i_batch_2_wave_offset[b];
    const int *ib = i_basis_batches + (size_t)n_centers_basis_I * b;
    double *dm = (double *)malloc(sizeof(double) * (size_t)nc * nc);
    for (int k = 0; k < nc; k++)
      for (int i = 0; i < nc; i++)
        dm[i + (size_t)nc * k] = density_matrix[(ib[i] - 1) + (size_t)n_basis * (ib[k] - 1)];
    for (int j = j_atom_begin; j <= j_atom_end; j++) {
      const int t = j - j_atom_begin;
      const int vs = atom_valid_n_compute_c_batches[(j - 1) + (size_t)(n_atoms + 1) * b];
      const int ve = atom_valid_n_compute_c_batches[j + (size_t)(n_atoms + 1) * b];
      double *out = first_order_rho_bias_part2 + (size_t)n_max_batch_size * (b + (size_t)n_my_batches_work * t);
      for (int p = 0; p < np; p++) {
        double acc = 0.0;
        if (ve > vs) {
          /* first term: sum_i phi(i) sum_{k in V} DM(i, k) phi'(k) */
          double a1 = 0.0;
          for (int i = 0; i < nc; i++) {
            double w = 0.0;
            for (int k = vs; k < ve; k++)
              w += dm[i + (size_t)nc * k] * -gwave[k + (size_t)ld * (j_coord - 1) + (size_t)3 * ld * p];
            a1 += wave[i + (size_t)ld * p] * w;
          }
          /* second term: sum_{i in V} phi'(i) sum_k DM(i, k) phi(k) */
          double a2 = 0.0;
          for (int i = vs; i < ve; i++) {
            double w = 0.0;
            for (int k = 0; k < nc; k++)
              w += dm[i + (size_t)nc * k] * wave[k + (size_t)ld * p];
            a2 += -gwave[i + (size_t)ld * (j_coord - 1) + (size_t)3 * ld * p] * w;
          }
          acc = a1 + a2;
        }
        out[p] = acc;
      }
    }
    free(dm);
  }
}

/*
 * Restatement of the inline bias of the v6 variant of evaluate_first_order_rho
 * (task4_common_calculation_eliminated/version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:
 * 179-231; its GPU copy evaluate_first_order_rho.cu.hpp:202-250): for every batch with n_compute_c > 0
 * and every atom j of the tile (t = j - j_atom_begin), with V the compute rows of atom j,
 *   work(i, p) = sum_k DM(i, k) phi(k, p), i in V                              (second dgemm, :189-201)
 *   acc_bias(p, t) = sum_{i in V} -grad_{j_coord} phi(i, p) work(i, p)            (:205-219)
 * which v6 adds to rho1 on top of first_order_rho_bias_part2 (:222).  Here it is ADDED into
 * first_order_rho_bias (n_max_batch_size, n_batches, atom_tile_size) for the tile's perturbations;
 * other entries are left as they are.  It is the second of the two terms of bias part 2 above, so
 * inline(DM) + inline(DM^T) = part2(DM).
 */
void oracle_rho_bias_inline(int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end, int j_coord,
                            int n_basis, int n_atoms, int n_centers_basis_I, int n_max_batch_size,
                            const int *n_point_batches, const int *n_compute_c_batches, const int *i_basis_batches,
                            const int *atom_valid_n_compute_c_batches, const int *i_batch_2_wave_offset,
                            const double *wave_batches_compress, const double *gradient_wave_batches_compress,
                            double *first_order_rho_bias, const double *density_matrix) {
  (void)atom_tile_size;
#pragma omp parallel for schedule(dynamic, 4)
  for (int b = 0; b < n_my_batches_work; b++) {
    const int np = n_point_batches[b], nc = n_compute_c_batches[b];
    if (nc <= 0)
      continue;
    const int ld = (nc + 3) / 4 * 4;
    const double *wave = wave_batches_compress + i_batch_2_wave_offset[b];
    const double *gwave = gradient_wave_batches_compress + 3 * (size_t)i_batch_2_wave_offset[b];
    const int *ib = i_basis_batches + (size_t)n_centers_basis_I * b;
    for (int j = j_atom_begin; j <= j_atom_end; j++) {
      const int t = j - j_atom_begin;
      const int vs = atom_valid_n_compute_c_batches[(j - 1) + (size_t)(n_atoms + 1) * b];
      const int ve = atom_valid_n_compute_c_batches[j + (size_t)(n_atoms + 1) * b];
      double *out = first_order_rho_bias + (size_t)n_max_batch_size * (b + (size_t)n_my_batches_work * t);
      for (int p = 0; p < np; p++) {
        double acc = 0.0;
        for (int i = vs; i < ve; i++) {
          double w = 0.0;
          for (int k = 0; k < nc; k++)
            w += density_matrix[(ib[i] - 1) + (size_t)n_basis * (ib[k] - 1)] * wave[k + (size_t)ld * p];
          acc += -gwave[i + (size_t)ld * (j_coord - 1) + (size_t)3 * ld * p] * w;
        }
        out[p] += acc;
      }
    }
  }
}
