/*
 * Oracle (test infrastructure only): restatement of the far-field branch of
 * sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_
 * (fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cpp:195-529) and of
 * far_distance_real_hartree_potential_single_atom_p2_fused_fp_c (:91-187).
 *
 *   index_lm(m, l): 1-based running index over l = 0..l_pot_max, m = -l..l            (:349-356)
 *   multipole_c(n, a, t) = cc(n) M(index_lm(m_n, l_n) - 1, a, t) if l_n <= lmfd(a, t) and
 *                          |M| > 1e-10, else 0, for n < n_cc_lm_ijk(lmfd(a, t))          (:358-378)
 *   loop t, hartree centre ic (centre c = centers_hartree_potential(ic) - 1, atom a), point:
 *     d = r_p - R_c; l = l_hartree(species(a)); while outer_potential_radius(l, a, t) < d^2
 *     and l > 0: l--                                                                     (:382-407)
 *     far if d^2 >= multipole_radius_sq(a, t) and d^2 < 1e8                              (:411,478)
 *   far term: Fp(0) = 1/d, Fp(k) = Fp(k-1) (1 - 2k) / d^2 for k <= l + hartree_force_l_add;
 *     powers of dir up to max index_ijk_max_cc(:, l);
 *     dV += sum_{n < n_cc_lm_ijk(l)} x^i y^j z^k Fp(nn) multipole_c(n, atom)               (:141-186)
 *
 * As in the reference, the inner function is called with the LOOP index ic + 1 (:484), so the
 * monomials use coords_center(:, ic) and the coefficients use center_to_atom(ic); the distance and
 * the l reduction use the centre c = centers_hartree_potential(ic) - 1.  The two agree whenever
 * centers_hartree_potential is the identity, as in every cluster run.
 * Entries of multipole_c beyond n_cc_lm_ijk(lmfd(a, t)) are left NaN here (the reference leaves
 * them uninitialised); reading them would poison the result.
 * delta_v_hartree is ACCUMULATED into (+=).  The near branch (:411-477) is oracle_sum_up_near
 * (nearfield.c); the whole sum_up_whole_potential is the two applied one after the other.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

void oracle_sum_up_far(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int l_max_analytic_multipole,
    int index_cc_dim_0, int n_valid_points, int l_pot_max, int n_centers, int n_centers_hartree_potential,
    int n_atoms, int hartree_force_l_add,
    const int *species, const int *l_hartree_max_far_distance, const int *i_valid_point_2_i_full_points_map,
    const double *coord_points, const double *multipole_radius_sq, const double *outer_potential_radius,
    const double *multipole_moments, const int *l_hartree, const int *n_cc_lm_ijk,
    const int *centers_hartree_potential, const int *center_to_atom, const int *index_cc,
    const int *index_ijk_max_cc, const double *cc, const double *coords_center, double *delta_v_hartree) {
  (void)n_centers;
  const int T = atom_tile_size;
  const int n_tile = j_atom_end - j_atom_begin + 1;
  const int lp1 = l_pot_max + 1;
  const int l_pot_max_pow2 = lp1 * lp1;
  const int mc_dim0 = n_cc_lm_ijk[l_pot_max];

  int *index_lm = (int *)malloc(sizeof(int) * (size_t)(2 * l_pot_max + 1) * lp1);
  int i_index = 0;
  for (int l = 0; l <= l_pot_max; l++)
    for (int m = -l; m <= l; m++)
      index_lm[(m + l_pot_max) + (2 * l_pot_max + 1) * l] = ++i_index;

  double *multipole_c = (double *)malloc(sizeof(double) * (size_t)mc_dim0 * n_atoms * T);
  for (size_t q = 0; q < (size_t)mc_dim0 * n_atoms * T; q++)
    multipole_c[q] = NAN;
  for (int t = 0; t < n_tile; t++) {
    for (int a = 0; a < n_atoms; a++) {
      const int lm = l_hartree_max_far_distance[a + n_atoms * t];
      for (int n = 0; n < n_cc_lm_ijk[lm]; n++) {
        const int il = index_cc[n];
        const int im = index_cc[n + index_cc_dim_0] + l_pot_max;
        double v = 0.0;
        if (il <= lm) {
          const double M = multipole_moments[(index_lm[im + (2 * l_pot_max + 1) * il] - 1) +
                                             (size_t)l_pot_max_pow2 * (a + (size_t)n_atoms * t)];
          if (fabs(M) > 1e-10)
            v = cc[n] * M;
        }
        multipole_c[n + (size_t)mc_dim0 * (a + (size_t)n_atoms * t)] = v;
      }
    }
  }

#pragma omp parallel for schedule(static)
  for (int iv = 0; iv < n_valid_points; iv++) {
    const int p = i_valid_point_2_i_full_points_map[iv];
    const double *r = &coord_points[3 * (size_t)p];
    double Fp[64];
    double coord_c[3][64];
    for (int t = 0; t < n_tile; t++) {
      for (int ic = 0; ic < n_centers_hartree_potential; ic++) {
        const int cur = centers_hartree_potential[ic] - 1;
        const int atom = center_to_atom[cur] - 1;
        double dist_sq = 0.0;
        for (int c = 0; c < 3; c++) {
          const double d = r[c] - coords_center[c + 3 * (size_t)cur];
          dist_sq += d * d;
        }
        int l = l_hartree[species[atom] - 1];
        while (outer_potential_radius[l + lp1 * (atom + (size_t)n_atoms * t)] < dist_sq && l > 0)
          l--;
        if (dist_sq < multipole_radius_sq[atom + (size_t)n_atoms * t])
          continue; /* near field: oracle_sum_up_near */
        if (!(dist_sq < 1e8))
          continue;
        /* far_distance_real_hartree_potential_single_atom_p2_fused_fp_c(l, ..., ic + 1, sqrt(d2)) */
        const double dist = sqrt(dist_sq);
        const double dsq = dist * dist;
        int one_minus_2l = 1;
        Fp[0] = 1.0 / dist;
        for (int k = 1; k <= l + hartree_force_l_add; k++) {
          one_minus_2l -= 2;
          Fp[k] = Fp[k - 1] * one_minus_2l / dsq;
        }
        const int ic1 = ic; /* i_center - 1 with i_center = ic + 1 */
        double dir[3];
        for (int c = 0; c < 3; c++)
          dir[c] = r[c] - coords_center[c + 3 * (size_t)ic1];
        coord_c[0][0] = coord_c[1][0] = coord_c[2][0] = 1.0;
        int maxval = -1;
        for (int c = 0; c < 3; c++) {
          const int v = index_ijk_max_cc[c + 3 * l];
          maxval = maxval > v ? maxval : v;
        }
        for (int k = 1; k <= maxval; k++)
          for (int c = 0; c < 3; c++)
            coord_c[c][k] = dir[c] * coord_c[c][k - 1];
        const double *mc = &multipole_c[(size_t)mc_dim0 * ((center_to_atom[ic1] - 1) + (size_t)n_atoms * t)];
        double dpot = 0.0;
        for (int n = 0; n < n_cc_lm_ijk[l]; n++) {
          const int ii = index_cc[n + 2 * index_cc_dim_0];
          const int jj = index_cc[n + 3 * index_cc_dim_0];
          const int kk = index_cc[n + 4 * index_cc_dim_0];
          const int nn = index_cc[n + 5 * index_cc_dim_0];
          dpot += coord_c[0][ii] * coord_c[1][jj] * coord_c[2][kk] * Fp[nn] * mc[n];
        }
        delta_v_hartree[(size_t)p + (size_t)n_full_points * t] += dpot;
      }
    }
  }
  free(index_lm);
  free(multipole_c);
}
