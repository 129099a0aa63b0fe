/*
 * Oracle (test infrastructure only): restatement of the near-field branch of
 * sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_
 * (fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cpp:382-477), of the fused spline ddot
 * spline_vector_v2_n2_c_fused_ddot (:12-49) and of the real spherical harmonics SHEval_s
 * (fast_rap_src_dfpt_gpu_eval/sheval.hpp:6-343).
 *
 *   loop t, hartree centre ic (centre c = centers_hartree_potential(ic) - 1, atom a), point:
 *     d = r_p - R_c; l = l_hartree(species(a)); while outer_potential_radius(l, a, t) < d^2
 *     and l > 0: l--                                                                    (:386-411)
 *     near if d^2 < multipole_radius_sq(a, t)                                                (:411)
 *     r = sqrt(d^2), u = d / r, s = species_center(c);
 *     i_r_log = 1 + log(r / r_grid_min(s)) / log_r_grid_inc(s)                           (:415-425)
 *     (sin theta, cos theta, sin phi, cos phi) of u as tab_single_trigonom               (:427-450)
 *     Y_lm(u) for l' <= l                                                                (:452-467)
 *     dV(p, t) += sum_{lm < (l+1)^2} Y_lm S_lm(i_r_log; a, t)                            (:469-476)
 *   S_lm: cubic Hermite on the log-grid index, coefficients (value, d/di) of
 *     current_delta_v_hart_part_spl_tile (lm, 2, n_hartree_grid, a, t):
 *     i = clamp(int(i_r_log), 1, n_grid(s) - 1), t = i_r_log - i;
 *     S = c0(i-1)(t-1)^2(1+2t) + c1(i-1)(t-1)^2 t + c0(i) t^2(3-2t) + c1(i) t^2(t-1)      (:12-49)
 *
 * Y_lm convention (index l^2 + l + m, as SHEval): with x = sin(theta) cos(phi), y = sin(theta)
 * sin(phi), z = cos(theta), C_m + i S_m = (x + i y)^m and Pbar_l^m(z) = N_lm P_l^m(z) / sin^m(theta)
 * (P without the Condon-Shortley phase, N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)):
 *   Y_l0 = Pbar_l^0,  Y_lm = sqrt(2) (-1)^m Pbar_l^m C_m,  Y_l,-m = sqrt(2) Pbar_l^m S_m  (m > 0).
 * Pbar from Pbar_m^m = N_mm (2m-1)!!, Pbar_{m+1}^m = sqrt(2m+3) z Pbar_m^m and
 * Pbar_l^m = a_lm (z Pbar_{l-1}^m - b_lm Pbar_{l-2}^m), a_lm = sqrt((4l^2-1)/(l^2-m^2)),
 * b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)).  The reference evaluates l <= 6 from SHEval's
 * explicit polynomials (and calls the host program's increment_ylm beyond); this recurrence gives
 * the same functions for every l.
 * delta_v_hartree is ACCUMULATED into (+=).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define PI_D 3.14159265358979323846
#define SQRT2_D 1.41421356237309504880

void oracle_ylm(int lmax, double sintheta, double costheta, double sinphi, double cosphi, double *ylm) {
  const double x = sintheta * cosphi, y = sintheta * sinphi, z = costheta;
  double C[32], S[32], pmm[32];
  C[0] = 1.0;
  S[0] = 0.0;
  for (int m = 1; m <= lmax; m++) {
    C[m] = x * C[m - 1] - y * S[m - 1];
    S[m] = x * S[m - 1] + y * C[m - 1];
  }
  /* Pbar_m^m = N_mm (2m-1)!! = sqrt((2m+1)/(4 pi) / (2m)!) (2m-1)!! */
  pmm[0] = sqrt(1.0 / (4.0 * PI_D));
  for (int m = 1; m <= lmax; m++)
    pmm[m] = pmm[m - 1] * sqrt((2.0 * m + 1.0) / (2.0 * m));
  for (int m = 0; m <= lmax; m++) {
    double prev = 0.0, p = pmm[m]; /* Pbar_{l-1}^m, Pbar_l^m */
    for (int l = m; l <= lmax; l++) {
      if (l > m) {
        const double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        const double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        const double next = a * (z * p - b * prev);
        prev = p;
        p = next;
      }
      if (m == 0) {
        ylm[l * l + l] = p;
      } else {
        const double s = (m & 1) ? -SQRT2_D : SQRT2_D;
        ylm[l * l + l + m] = s * p * C[m];
        ylm[l * l + l - m] = SQRT2_D * p * S[m];
      }
    }
  }
}

/* tab_single_trigonom_p0_c (sum_up_whole_potential.cpp:427-450): trig[0..3] = sin theta,
 * cos theta, sin phi, cos phi of the unit vector u. */
static void trigonom(const double u[3], double trig[4]) {
  double ab;
  const double abmax = fmax(fabs(u[0]), fabs(u[1]));
  if (abmax > 1.0e-36) {
    ab = sqrt(u[0] * u[0] + u[1] * u[1]);
    trig[3] = u[0] / ab;
    trig[2] = u[1] / ab;
  } else {
    trig[3] = 1.0;
    trig[2] = 0.0;
    ab = 0.0;
  }
  const double abcmax = fmax(abmax, fabs(u[2]));
  if (abcmax > 1.0e-36) {
    const double abc = sqrt(ab * ab + u[2] * u[2]);
    trig[1] = u[2] / abc;
    trig[0] = ab / abc;
  } else {
    trig[1] = 0.0;
    trig[0] = 1.0;
  }
}

/* spline_vector_v2_n2_c_fused_ddot (sum_up_whole_potential.cpp:12-49) */
static double spline_ddot(double r_output, int n_vector, int n_l_dim, int n_points, const double *spl,
                          const double *ylm) {
  int i_spl = (int)r_output;
  i_spl = i_spl > 1 ? i_spl : 1;
  i_spl = i_spl < n_points - 1 ? i_spl : n_points - 1;
  const double t = r_output - i_spl;
  const double ta = (t - 1) * (t - 1) * (1 + 2 * t);
  const double tb = (t - 1) * (t - 1) * t;
  const double tc = t * t * (3 - 2 * t);
  const double td = t * t * (t - 1);
  const double *g0 = spl + (size_t)2 * n_l_dim * (i_spl - 1);
  const double *g1 = g0 + 2 * n_l_dim;
  double acc = 0.0;
  for (int i = 0; i < n_vector; i++)
    acc += (g0[i] * ta + g0[n_l_dim + i] * tb + g1[i] * tc + g1[n_l_dim + i] * td) * ylm[i];
  return acc;
}

void oracle_sum_up_near(
    int atom_tile_size, int j_atom_begin, int j_atom_end, int n_full_points, int n_valid_points, int l_pot_max,
    int n_hartree_grid, int n_centers_hartree_potential, int n_atoms,
    const int *species, const int *species_center, const int *i_valid_point_2_i_full_points_map,
    const double *coord_points, const double *multipole_radius_sq, const double *outer_potential_radius,
    const double *current_delta_v_hart_part_spl_tile, const int *l_hartree, const int *n_grid,
    const int *centers_hartree_potential, const int *center_to_atom, const double *coords_center,
    const double *r_grid_min, const double *log_r_grid_inc, double *delta_v_hartree) {
  const int T = atom_tile_size;
  const int n_tile = j_atom_end - j_atom_begin + 1;
  const int lp1 = l_pot_max + 1;
  const int L2 = lp1 * lp1;
  (void)T;
#pragma omp parallel for schedule(static)
  for (int iv = 0; iv < n_valid_points; iv++) {
    const int p = i_valid_point_2_i_full_points_map[iv];
    const double *r = &coord_points[3 * (size_t)p];
    double ylm[32 * 32];
    for (int t = 0; t < n_tile; t++) {
      for (int ic = 0; ic < n_centers_hartree_potential; ic++) {
        const int cur = centers_hartree_potential[ic] - 1;
        const int atom = center_to_atom[cur] - 1;
        double dir[3], dist_sq = 0.0;
        for (int c = 0; c < 3; c++) {
          dir[c] = r[c] - coords_center[c + 3 * (size_t)cur];
          dist_sq += dir[c] * dir[c];
        }
        int l = l_hartree[species[atom] - 1];
        while (outer_potential_radius[l + lp1 * (atom + (size_t)n_atoms * t)] < dist_sq && l > 0)
          l--;
        if (!(dist_sq < multipole_radius_sq[atom + (size_t)n_atoms * t]))
          continue; /* far field: oracle_sum_up_far */
        const double dist = sqrt(dist_sq);
        const double u[3] = {dir[0] / dist, dir[1] / dist, dir[2] / dist};
        const int sp = species_center[cur] - 1;
        const double i_r_log = 1.0 + log(dist / r_grid_min[sp]) / log_r_grid_inc[sp];
        double trig[4];
        trigonom(u, trig);
        oracle_ylm(l, trig[0], trig[1], trig[2], trig[3], ylm);
        const double *spl = current_delta_v_hart_part_spl_tile +
                            (size_t)L2 * 2 * n_hartree_grid * (atom + (size_t)n_atoms * t);
        delta_v_hartree[(size_t)p + (size_t)n_full_points * t] +=
            spline_ddot(i_r_log, (l + 1) * (l + 1), L2, n_grid[sp], spl, ylm);
      }
    }
  }
}
