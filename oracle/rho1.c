/*
 * Oracle (test infrastructure only): restatement of
 * evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_
 * (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.cpp:12-210).
 *
 * Per batch b with n_compute_c > 0:
 *   (1) gather   P1loc(t, i, j) = P1(i_basis(i) - 1, i_basis(j) - 1, t)          (:149-159)
 *   (2) GEMM     work1 = P1loc (T*nc x nc) * wave_b (nc x np, ld nc_pad)          (:161-175)
 *   (3) ddot     rho(p, t) = sum_ic wave(ic, p) work1(t, ic, p) + bias2(p, b, t)
 *                grad(j, p, b, t) = 2 sum_ic gradwave(ic, j, p) work1(t, ic, p) - gbias  (:177-200)
 *   (4) scatter  first_order_rho(map(p), t) = rho(p, t) for t < j_atom_end-j_atom_begin+1 (:202-207)
 * first_order_rho and first_order_gradient_rho are zeroed first (:113-114).
 * Unused arguments of the reference (basis_atom, partition_tab, density_matrix, batch_sizes,
 * atom_valid_n_compute_c_batches, j_coord, n_atoms, n_max_compute_ham) are accepted and ignored.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PB 8 /* points per register block of the GEMM loop */

static void batch_rho1(int T, int nc, int np, int n_basis, const int *ib, const double *wave, const double *gwave,
                       const double *P1, double *local_rho /* (np, T) */, double *acc_grad /* (3, np, T) */) {
  const int ncp = ((nc + 3) / 4) * 4;
  double *Ploc = (double *)malloc(sizeof(double) * (size_t)nc * nc);
  double *work = (double *)malloc(sizeof(double) * (size_t)nc * np);
  for (int t = 0; t < T; t++) {
    /* gather (one perturbation at a time; the reference's (T, nc, nc) order only changes strides) */
    for (int j = 0; j < nc; j++) {
      const size_t jb = (size_t)(ib[j] - 1);
      for (int i = 0; i < nc; i++)
        Ploc[i + (size_t)nc * j] = P1[(size_t)(ib[i] - 1) + (size_t)n_basis * (jb + (size_t)n_basis * t)];
    }
    /* work(i, p) = sum_j Ploc(i, j) wave(j, p), j ascending */
    for (int p0 = 0; p0 < np; p0 += PB) {
      const int pn = (np - p0) < PB ? (np - p0) : PB;
      for (int q = 0; q < pn; q++)
        memset(&work[(size_t)nc * (p0 + q)], 0, sizeof(double) * nc);
      for (int j = 0; j < nc; j++) {
        const double *pc = &Ploc[(size_t)nc * j];
        for (int q = 0; q < pn; q++) {
          const double w = wave[j + (size_t)ncp * (p0 + q)];
          double *wc = &work[(size_t)nc * (p0 + q)];
          for (int i = 0; i < nc; i++)
            wc[i] += pc[i] * w;
        }
      }
    }
    for (int p = 0; p < np; p++) {
      double acc = 0.0, ag0 = 0.0, ag1 = 0.0, ag2 = 0.0;
      const double *wc = &work[(size_t)nc * p];
      const double *phi = &wave[(size_t)ncp * p];
      const double *g = &gwave[(size_t)3 * ncp * p];
      for (int ic = 0; ic < nc; ic++) {
        acc += phi[ic] * wc[ic];
        ag0 += g[ic] * wc[ic];
        ag1 += g[ic + ncp] * wc[ic];
        ag2 += g[ic + 2 * ncp] * wc[ic];
      }
      local_rho[p + (size_t)np * t] = acc;
      acc_grad[0 + 3 * (p + (size_t)np * t)] = ag0;
      acc_grad[1 + 3 * (p + (size_t)np * t)] = ag1;
      acc_grad[2 + 3 * (p + (size_t)np * t)] = ag2;
    }
  }
  free(Ploc);
  free(work);
}

void oracle_evaluate_first_order_rho(
    int atom_tile_size, int n_my_batches_work, int j_atom_begin, int j_atom_end, int j_coord, int n_full_points,
    int n_basis, int n_atoms, int n_max_compute_ham, int n_centers_basis_I, int n_max_batch_size,
    const int *batch_sizes, const int *n_point_batches, const int *n_point_batches_prefix_sum,
    const int *i_valid_point_2_i_full_points_map, const int *n_compute_c_batches, const int *i_basis_batches,
    const int *atom_valid_n_compute_c_batches, const int *i_batch_2_wave_offset,
    const double *wave_batches_compress, const double *gradient_wave_batches_compress,
    const double *first_order_density_matrix, double *first_order_rho, const double *first_order_rho_bias_part2,
    double *first_order_gradient_rho, const double *first_order_gradient_rho_bias_batches_atoms) {
  (void)j_coord; (void)n_atoms; (void)n_max_compute_ham; (void)batch_sizes; (void)atom_valid_n_compute_c_batches;
  const int T = atom_tile_size;
  memset(first_order_rho, 0, sizeof(double) * (size_t)n_full_points * T);
  memset(first_order_gradient_rho, 0, sizeof(double) * (size_t)3 * n_max_batch_size * n_my_batches_work * T);
  const int n_tile = j_atom_end - j_atom_begin + 1;

#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < n_my_batches_work; b++) {
    const int np = n_point_batches[b];
    const int nc = n_compute_c_batches[b];
    if (nc <= 0)
      continue;
    const int *map = &i_valid_point_2_i_full_points_map[n_point_batches_prefix_sum[b]];
    const int *ib = &i_basis_batches[(size_t)n_centers_basis_I * b];
    const double *wave = wave_batches_compress + i_batch_2_wave_offset[b];
    const double *gwave = gradient_wave_batches_compress + (size_t)3 * i_batch_2_wave_offset[b];
    double *local_rho = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1) * T);
    double *acc_grad = (double *)malloc(sizeof(double) * (size_t)3 * (np > 0 ? np : 1) * T);
    batch_rho1(T, nc, np, n_basis, ib, wave, gwave, first_order_density_matrix, local_rho, acc_grad);
    for (int t = 0; t < T; t++) {
      for (int p = 0; p < np; p++) {
        const size_t bi = (size_t)p + (size_t)n_max_batch_size * (b + (size_t)n_my_batches_work * t);
        local_rho[p + (size_t)np * t] += first_order_rho_bias_part2[bi];
        for (int j = 0; j < 3; j++)
          first_order_gradient_rho[j + 3 * bi] =
              acc_grad[j + 3 * (p + (size_t)np * t)] * 2 - first_order_gradient_rho_bias_batches_atoms[j + 3 * bi];
      }
    }
    for (int t = 0; t < n_tile; t++)
      for (int p = 0; p < np; p++)
        first_order_rho[(size_t)map[p] + (size_t)n_full_points * t] = local_rho[p + (size_t)np * t];
    free(local_rho);
    free(acc_grad);
  }
}
