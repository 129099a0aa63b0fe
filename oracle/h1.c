/*
 * Oracle (test infrastructure only): restatement of evaluate_first_order_h_reduce_memory_c_v2_
 * (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H.cpp:13-222), LDA and GGA terms.
 *
 * For every batch b and perturbation j_atom (inner index t), with p over the batch's points and
 * q = i_valid_point_2_i_full_points_map(prefix(b) + p):
 *   sigma1 = 2 sum_i grad_rho1(i, p) grad_rho(i, q)                                  (:149-152)
 *   pref2_i = 2 w(q) grad_rho(i, q) (v2rhosigma(0,p) rho1(q, t) + v2sigma2(0,p) sigma1)
 *             + 2 w(q) vsigma(0, p) grad_rho1(i, p)                                   (:153-160)
 *   pref3 = w(q) (v_hartree_gradient(q, t) + v2rho2(0,p) rho1(q, t) + v2rhosigma(0,p) sigma1) (:161-164)
 *   pref5_i = 2 vsigma(0, p) w(q) grad_rho(i, q)                                      (:165-168)
 *   C(ic, p) = 0.5 (H_prefactor1(p, b, t) + pref3) wave(ic, p) + sum_i pref2_i gradwave(ic, i, p)
 *            - sum_i pref5_i hessian(ic, index_hessian[i][j_coord-1], p) for ic in the atom's rows
 *                                                                                     (:170-189)
 *   Hloc = C wave^T + wave C^T (lower triangle, dsyr2k)                               (:192-205)
 *   first_order_H(ib_i, ib_j, t) += Hloc for the full symmetric block (ACCUMULATED)   (:207-215)
 * The per-batch XC arrays are (k, n_max_batch_size) blocks at k * n_max_batch_size * b (:79-83).
 */
#include <stdlib.h>

#include "oracle.h"

static const int index_hessian[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

/* Local H of batch b, perturbation j_atom: Hl (nc x nc) lower triangle (:144-205).  Returns nc (0 =
 * nothing to add).  Hl is allocated here and freed by the caller. */
static int batch_h(int b, int j_atom, int j_atom_begin, int n_full_points, int j_coord, int n_atoms,
                   int n_max_batch_size, int n_my_batches_work, const int *n_point_batches,
                   const int *n_compute_c_batches, const int *atom_valid_n_compute_c_batches,
                   const int *n_point_batches_prefix_sum, const int *i_valid_point_2_i_full_points_map,
                   const double *vsigma_batches, const double *gradient_rho, const int *i_batch_2_wave_offset,
                   const double *wave_batches_compress, const double *gradient_wave_batches_compress,
                   const double *hessian_wave_batches_compress, const double *H_prefactor1_batches_tile,
                   const double *first_order_gradient_rho_batches_tile, const double *v_hartree_gradient,
                   const double *first_order_rho, const double *v2rho2_batches, const double *v2rhosigma_batches,
                   const double *v2sigma2_batches, const double *partition_tab, double **Hl_out) {
  *Hl_out = NULL;
  const double *vsigma = &vsigma_batches[(size_t)3 * n_max_batch_size * b];
  const double *v2rho2 = &v2rho2_batches[(size_t)3 * n_max_batch_size * b];
  const double *v2rhosigma = &v2rhosigma_batches[(size_t)6 * n_max_batch_size * b];
  const double *v2sigma2 = &v2sigma2_batches[(size_t)6 * n_max_batch_size * b];
  const int t = j_atom - j_atom_begin;
  const double *pref1 = &H_prefactor1_batches_tile[(size_t)n_max_batch_size * (b + (size_t)n_my_batches_work * t)];
  const double *grho1 =
      &first_order_gradient_rho_batches_tile[(size_t)3 * n_max_batch_size * (b + (size_t)n_my_batches_work * t)];
  const int np = n_point_batches[b];
  const int nc = n_compute_c_batches[b];
  const int ncp = ((nc + 3) / 4) * 4;
  if (nc <= 0)
    return 0;
  const double *wave = wave_batches_compress + i_batch_2_wave_offset[b];
  const double *gw = gradient_wave_batches_compress + (size_t)3 * i_batch_2_wave_offset[b];
  const double *hw = hessian_wave_batches_compress + (size_t)6 * i_batch_2_wave_offset[b];
  const int v0 = atom_valid_n_compute_c_batches[(j_atom - 1) + (n_atoms + 1) * (size_t)b];
  const int v1 = atom_valid_n_compute_c_batches[j_atom + (n_atoms + 1) * (size_t)b];
  const int *map = &i_valid_point_2_i_full_points_map[n_point_batches_prefix_sum[b]];
  double *C = (double *)malloc(sizeof(double) * (size_t)nc * (np > 0 ? np : 1));
  double *Hl = (double *)malloc(sizeof(double) * (size_t)nc * nc);
  for (int p = 0; p < np; p++) {
    const int q = map[p];
    const double w = partition_tab[q];
    const double rho1 = first_order_rho[(size_t)q + (size_t)n_full_points * t];
    double sig1 = 0.0;
    for (int i = 0; i < 3; i++)
      sig1 += grho1[i + 3 * p] * gradient_rho[i + 3 * (size_t)q];
    sig1 *= 2.0;
    double pref2[3], pref5[3];
    for (int i = 0; i < 3; i++) {
      pref2[i] = 2.0 * w * gradient_rho[i + 3 * (size_t)q] * (v2rhosigma[6 * p] * rho1 + v2sigma2[6 * p] * sig1) +
                 2.0 * w * vsigma[3 * p] * grho1[i + 3 * p];
      pref5[i] = 2.0 * vsigma[3 * p] * w * gradient_rho[i + 3 * (size_t)q];
    }
    const double pref3 = w * (v_hartree_gradient[(size_t)q + (size_t)n_full_points * t] + v2rho2[3 * p] * rho1 +
                              v2rhosigma[6 * p] * sig1);
    for (int ic = 0; ic < nc; ic++) {
      const double c1 = (pref1[p] + pref3) * wave[ic + (size_t)ncp * p];
      double c2 = 0.0;
      for (int i = 0; i < 3; i++)
        c2 += pref2[i] * gw[ic + (size_t)ncp * (i + 3 * (size_t)p)];
      C[ic + (size_t)nc * p] = 0.5 * c1 + c2;
    }
    for (int ic = v0; ic < v1; ic++) {
      double c5 = 0.0;
      for (int i = 0; i < 3; i++)
        c5 += pref5[i] * -hw[ic + (size_t)ncp * (index_hessian[i][j_coord - 1] + 6 * (size_t)p)];
      C[ic + (size_t)nc * p] += c5;
    }
  }
  for (int j = 0; j < nc; j++)
    for (int i = j; i < nc; i++) {
      double s = 0.0;
      for (int p = 0; p < np; p++)
        s += C[i + (size_t)nc * p] * wave[j + (size_t)ncp * p] + wave[i + (size_t)ncp * p] * C[j + (size_t)nc * p];
      Hl[i + (size_t)nc * j] = s;
    }
  free(C);
  *Hl_out = Hl;
  return nc;
}

/* Local H blocks of a chunk of batches are computed in parallel; they are added into H one batch
 * after another in the reference's order (batches outer, perturbations inner), so the sums do not
 * depend on the number of threads. */
#define H_CHUNK 32

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
    const double *v2sigma2_batches, const double *partition_tab) {
  (void)atom_tile_size; (void)n_max_compute_ham; (void)vrho_batches;
  const int n_tile = j_atom_end - j_atom_begin + 1;
  if (n_tile <= 0)
    return;
  double **Hl = (double **)calloc((size_t)H_CHUNK * n_tile, sizeof(double *));
  for (int b0 = 0; b0 < n_my_batches_work; b0 += H_CHUNK) {
    const int nb = (n_my_batches_work - b0) < H_CHUNK ? (n_my_batches_work - b0) : H_CHUNK;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int k = 0; k < nb; k++)
      for (int t = 0; t < n_tile; t++)
        (void)batch_h(b0 + k, j_atom_begin + t, j_atom_begin, n_full_points, j_coord, n_atoms, n_max_batch_size,
                      n_my_batches_work, n_point_batches, n_compute_c_batches, atom_valid_n_compute_c_batches,
                      n_point_batches_prefix_sum, i_valid_point_2_i_full_points_map, vsigma_batches, gradient_rho,
                      i_batch_2_wave_offset, wave_batches_compress, gradient_wave_batches_compress,
                      hessian_wave_batches_compress, H_prefactor1_batches_tile,
                      first_order_gradient_rho_batches_tile, v_hartree_gradient, first_order_rho, v2rho2_batches,
                      v2rhosigma_batches, v2sigma2_batches, partition_tab, &Hl[(size_t)k * n_tile + t]);
    for (int k = 0; k < nb; k++) {
      const int b = b0 + k;
      const int nc = n_compute_c_batches[b];
      const int *ib = &i_basis_batches[(size_t)n_centers_basis_I * b];
      for (int t = 0; t < n_tile; t++) {
        const double *L = Hl[(size_t)k * n_tile + t];
        if (!L)
          continue;
        double *H = &first_order_H_tile[(size_t)n_basis * n_basis * t];
        for (int i = 0; i < nc; i++) {
          const size_t bi = (size_t)(ib[i] - 1);
          H[bi + (size_t)n_basis * bi] += L[i + (size_t)nc * i];
          for (int j = i + 1; j < nc; j++) {
            const size_t bj = (size_t)(ib[j] - 1);
            H[bi + (size_t)n_basis * bj] += L[j + (size_t)nc * i];
            H[bj + (size_t)n_basis * bi] += L[j + (size_t)nc * i];
          }
        }
        free(Hl[(size_t)k * n_tile + t]);
        Hl[(size_t)k * n_tile + t] = NULL;
      }
    }
  }
  free(Hl);
}

#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_set_threads(int n) {
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (n > 0)
    omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}
