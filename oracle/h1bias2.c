/*
 * Oracle (test infrastructure only): restatement of evaluate_first_order_h_bias_part2_
 * (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H_bias_part2.cpp:13-218).
 *
 * For every batch b with n_compute_c > 0 and every atom j of [j_atom_begin, j_atom_end]
 * (t = j - j_atom_begin), with V = compute rows [avn(j-1, b), avn(j, b)) of atom j, points p of the
 * batch, q = i_valid_point_2_i_full_points_map(p) and w = partition_tab(q):
 *   C(i, p) = w H_times_psi(i, p) + w vrho(p) phi(i, p)
 *             + sum_c 2 w vsigma(0, p) gradient_rho(c, q) d_c phi(i, p)                 (:138-156)
 *   L(i, k) = sum_p C(i, p) (-d_{j_coord} phi(k, p)),  i < n_compute_c, k in V           (:160-191)
 *   first_order_H(ib_i, ib_k, t) += L(i, k),  first_order_H(ib_k, ib_i, t) += L(i, k)    (:193-200)
 * (a diagonal element i = k receives L(k, k) twice).  The reference's contract matrices are
 * per-point loops and a dgemm; here they are plain loops.  Arguments the reference declares but does
 * not read (H_prefactor1, first-order gradient, v_hartree_gradient, first_order_rho, the second
 * derivatives, not_use) are omitted.  Batches run in parallel chunks and are added in order.
 */
#include <stdlib.h>

#include "oracle.h"

static void batch_local(int b, int j_atom, int j_coord, int n_atoms, int n_max_batch_size,
                        const int *n_point_batches, const int *n_compute_c_batches,
                        const int *atom_valid_n_compute_c_batches, const int *n_point_batches_prefix_sum,
                        const int *i_valid_point_2_i_full_points_map, const double *vrho_batches,
                        const double *vsigma_batches, const double *gradient_rho,
                        const double *H_times_psi_batches_compress, const int *i_batch_2_wave_offset,
                        const double *wave_batches_compress, const double *gradient_wave_batches_compress,
                        const double *partition_tab, double *C, double *L, int *v0_out, int *v1_out) {
  const int nc = n_compute_c_batches[b], np = n_point_batches[b];
  const int ld = (nc + 3) / 4 * 4;
  const double *hpsi = H_times_psi_batches_compress + i_batch_2_wave_offset[b];
  const double *wave = wave_batches_compress + i_batch_2_wave_offset[b];
  const double *gw = gradient_wave_batches_compress + (size_t)3 * i_batch_2_wave_offset[b];
  const double *vrho = vrho_batches + (size_t)n_max_batch_size * b;
  const double *vsigma = vsigma_batches + (size_t)3 * n_max_batch_size * b;
  const int *map = i_valid_point_2_i_full_points_map + n_point_batches_prefix_sum[b];
  const int v0 = atom_valid_n_compute_c_batches[(j_atom - 1) + (size_t)(n_atoms + 1) * b];
  const int v1 = atom_valid_n_compute_c_batches[j_atom + (size_t)(n_atoms + 1) * b];
  *v0_out = v0;
  *v1_out = v1;
  for (int p = 0; p < np; p++) {
    const int q = map[p];
    const double w = partition_tab[q];
    const double f1 = w, f2 = w * vrho[p];
    double f3[3];
    for (int c = 0; c < 3; c++)
      f3[c] = w * 2.0 * vsigma[3 * p] * gradient_rho[c + 3 * (size_t)q];
    for (int i = 0; i < nc; i++) {
      double x = f1 * hpsi[i + (size_t)ld * p] + f2 * wave[i + (size_t)ld * p];
      for (int c = 0; c < 3; c++)
        x += f3[c] * gw[i + (size_t)ld * (c + 3 * (size_t)p)];
      C[i + (size_t)nc * p] = x;
    }
  }
  for (int k = v0; k < v1; k++)
    for (int i = 0; i < nc; i++) {
      double s = 0.0;
      for (int p = 0; p < np; p++)
        s += C[i + (size_t)nc * p] * -gw[k + (size_t)ld * ((j_coord - 1) + 3 * (size_t)p)];
      L[i + (size_t)nc * (k - v0)] = s;
    }
}

void oracle_evaluate_first_order_h_bias_part2(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int j_coord, int n_basis, int n_atoms,
    int n_max_batch_size, int n_my_batches_work, int n_centers_basis_I, const int *n_point_batches,
    const int *n_compute_c_batches, const int *i_basis_batches, const int *atom_valid_n_compute_c_batches,
    const int *n_point_batches_prefix_sum, const int *i_valid_point_2_i_full_points_map, const double *vrho_batches,
    const double *vsigma_batches, const double *gradient_rho, const double *H_times_psi_batches_compress,
    const int *i_batch_2_wave_offset, const double *wave_batches_compress,
    const double *gradient_wave_batches_compress, double *first_order_H_tile, const double *partition_tab) {
  (void)atom_tile_size;
  (void)n_full_points;
  const int CH = 64;
  for (int j_atom = j_atom_begin; j_atom <= j_atom_end; j_atom++) {
    const int t = j_atom - j_atom_begin;
    double *H = first_order_H_tile + (size_t)n_basis * n_basis * t;
    for (int b0 = 0; b0 < n_my_batches_work; b0 += CH) {
      const int b1 = b0 + CH < n_my_batches_work ? b0 + CH : n_my_batches_work;
      double *Ls[64];
      int v0s[64], v1s[64];
#pragma omp parallel for schedule(dynamic, 1)
      for (int b = b0; b < b1; b++) {
        Ls[b - b0] = NULL;
        v0s[b - b0] = v1s[b - b0] = 0;
        const int nc = n_compute_c_batches[b], np = n_point_batches[b];
        if (nc <= 0)
          continue;
        double *C = (double *)malloc(sizeof(double) * (size_t)nc * (np > 0 ? np : 1));
        double *L = (double *)malloc(sizeof(double) * (size_t)nc * (nc > 0 ? nc : 1));
        batch_local(b, j_atom, j_coord, n_atoms, n_max_batch_size, n_point_batches, n_compute_c_batches,
                    atom_valid_n_compute_c_batches, n_point_batches_prefix_sum, i_valid_point_2_i_full_points_map,
                    vrho_batches, vsigma_batches, gradient_rho, H_times_psi_batches_compress,
                    i_batch_2_wave_offset, wave_batches_compress, gradient_wave_batches_compress, partition_tab, C,
                    L, &v0s[b - b0], &v1s[b - b0]);
        free(C);
        Ls[b - b0] = L;
      }
      for (int b = b0; b < b1; b++) {
        double *L = Ls[b - b0];
        if (!L)
          continue;
        const int nc = n_compute_c_batches[b];
        const int *ib = i_basis_batches + (size_t)n_centers_basis_I * b;
        for (int k = v0s[b - b0]; k < v1s[b - b0]; k++)
          for (int i = 0; i < nc; i++) {
            const double l = L[i + (size_t)nc * (k - v0s[b - b0])];
            const size_t bi = (size_t)(ib[i] - 1), bk = (size_t)(ib[k] - 1);
            H[bi + (size_t)n_basis * bk] += l;
            H[bk + (size_t)n_basis * bi] += l;
          }
        free(L);
      }
    }
  }
}
