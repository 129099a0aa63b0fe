// Far-field Hartree plan and the compile-time Cartesian term table.
#pragma once

#include <algorithm>
#include <utility>

#include "common.hpp"
#include "far_terms.inc"  // generated: FAR_TAB_LMAX, FAR_NCC[], FAR_TERMS[] (l, m, i, j, k, nn)

namespace frpa {

constexpr int FAR_LMAX = 10;  // largest l_pot_max the kernels are instantiated for
static_assert(FAR_TAB_LMAX >= FAR_LMAX, "generated term table too short");

// Per hartree centre ic, everything the kernels need (read with scalar loads: uniform per wave).
struct CenterRec {
  double dx, dy, dz;  // coords_center(:, centers_hartree_potential(ic) - 1): distance and l reduction
  double mx, my, mz;  // coords_center(:, ic): monomials (the reference passes the loop index)
  int atom;           // center_to_atom(centers_hartree_potential(ic) - 1) - 1
  int mc_atom;        // center_to_atom(ic) - 1: row of multipole_c used by the term sum
  int l_const;        // l_hartree(species(atom) - 1)
  int spc;            // species_center(centers_hartree_potential(ic)) - 1: log grid of the near field
};
static_assert(sizeof(CenterRec) == 64, "CenterRec layout");

// Log grid of one species for the near-field spline: i_r_log = 1 + (log r - log_rmin) * inv_inc.
struct NearSpecies {
  double log_rmin;  // log(r_grid_min)
  double inv_inc;   // 1 / log_r_grid_inc
  int n_points;     // n_grid: the spline interval index is clamped to [1, n_points - 1]
  int pad;
};

// 1 / sqrt(x) in FP64: hardware estimate refined by two Newton steps (error ~1 ulp).
__device__ __forceinline__ double rsqrt_f64(double x) {
  double r = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-x * r, r, 1.0);
    r = fma(0.5 * r, e, r);
  }
  return r;
}

// Number of merged (i, j, k, q) coefficients of the nested-Horner form for one l:
// sum_{q=0}^{l/2} (l - 2q + 1)^2.
constexpr int horner_ncoef(int L) {
  int n = 0;
  for (int q = 0; 2 * q <= L; q++)
    n += (L - 2 * q + 1) * (L - 2 * q + 1);
  return n;
}
constexpr int horner_prefix(int L) {  // offset of the l = L table inside one (atom, t) block
  int n = 0;
  for (int l = 0; l < L; l++)
    n += horner_ncoef(l);
  return n;
}
// Index of coefficient (q, i, j, k) inside the table of L.  Order: q, then i, j, k ascending with
// i <= min(1, D), j <= D - i, k <= D - i - j, D = L - 2q.
constexpr int horner_index(int L, int q, int i, int j, int k) {
  int n = 0;
  for (int qq = 0; qq < q; qq++)
    n += (L - 2 * qq + 1) * (L - 2 * qq + 1);
  const int D = L - 2 * q;
  for (int ii = 0; ii < i; ii++)
    for (int jj = 0; jj <= D - ii; jj++)
      n += D - ii - jj + 1;
  for (int jj = 0; jj < j; jj++)
    n += D - i - jj + 1;
  return n + k;
}

struct FarPlan {
  int l_pot_max = 0, lama = 0, dim0 = 0, n_centers = 0, n_hc = 0, n_atoms = 0, n_species = 0, hfla = 1;
  int mc_dim0 = 0;          // n_cc_lm_ijk(l_pot_max)
  bool terms_ok = false;    // caller's table equals the compiled term table (term-by-term kernel usable)
  int coef_stride = 0;      // horner_prefix(l_pot_max + 1)
  std::vector<int> h_ncc;
  DevBuf<CenterRec> d_centers;
  DevBuf<int> d_index_cc;   // (dim0, 6)
  DevBuf<double> d_cc;
  DevBuf<int> d_ncc;
  // merged-coefficient gather plan: entry e of a (atom, t) block sums terms hc_term[hc_ptr[e]..hc_ptr[e+1])
  DevBuf<int> d_hc_ptr, d_hc_term;
  DevBuf<double> d_hc_w;    // (-1)^nn (2nn-1)!! cc(n) per listed term
  DevBuf<int> d_atom_ic_ptr, d_atom_ic;  // hartree centres ic whose multipole_c row is atom a
  // near field: per-species log grids (set_near_tables); centre records carry the species
  std::vector<CenterRec> h_centers;
  std::vector<int> h_chp;   // centers_hartree_potential (1-based)
  bool near_ok = false;
  int near_max_points = 0;  // largest n_grid over the species used by the centres
  DevBuf<NearSpecies> d_nspec;
  std::vector<NearSpecies> h_nspec;
  // per-call scratch
  DevBuf<double> d_mc;      // (mc_dim0, n_atoms, T) reference multipole_c
  DevBuf<double> d_coef;    // (coef_stride, n_atoms, T)
  DevBuf<double> d_partial; // adjoint partial sums
};

struct FarCall {
  int T = 1, n_tile = 1, n_full = 0, n_valid = 0;
  const double *coord_points = nullptr;
  const int *map = nullptr, *lmfd = nullptr;
  const double *mrsq = nullptr, *opr = nullptr, *moments = nullptr;
  // near field: current_delta_v_hart_part_spl_tile (lm, 2, n_hartree_grid, n_atoms, T)
  const double *spl = nullptr;
  int n_hgrid = 0;
};

FarPlan *make_far_plan(int l_pot_max, int lama, int dim0, int n_centers, int n_hc, int n_atoms, int n_species,
                       int hfla, const int *species, const int *l_hartree, const int *chp, const int *center_to_atom,
                       const double *coords_center, const int *index_cc, const double *cc, const int *n_cc_lm_ijk,
                       const int *index_ijk_max_cc);
// Per-call coefficient table of a forward: multipole_c (FRPA_FAR_TERMS) or the merged Horner
// coefficients, (stride, n_atoms, T) doubles.
size_t far_coef_doubles(const FarPlan &P, int variant, int T);
// coef_buf: where the per-call coefficient table goes (at least far_coef_doubles doubles); nullptr =
// the plan's own buffer, allocated on first use and grown when a call needs more.
void far_forward(FarPlan &P, const FarCall &c, int variant, double *dv, hipStream_t s, double *coef_buf = nullptr);
// far_forward with the coefficient table in a fixed device array (no allocation; see frpa.h)
void far_forward_noalloc(FarPlan &P, const FarCall &c, int variant, double *dv, hipStream_t s);
void far_adjoint(FarPlan &P, const FarCall &c, const double *gV, double *gM, hipStream_t s);
void set_near_tables(FarPlan &P, int n_species, const int *species_center, const int *n_grid,
                     const double *r_grid_min, const double *log_r_grid_inc);
void near_forward(FarPlan &P, const FarCall &c, double *dv, hipStream_t s);

} // namespace frpa
