// Near-field Hartree potential: the spline branch of sum_up_whole_potential.
// Reference: sum_up_close_part_cu (fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cu:168-371),
// SHEval_spline_vector_v2_n2_ddot_fused (sheval.hpp:350-563), CPU form
// sum_up_whole_potential.cpp:382-477 with spline_vector_v2_n2_c_fused_ddot (:12-49).
//
// For every (valid point p, hartree centre c) pair with d^2 = |r_p - R_c|^2 < multipole_radius_sq:
//   l = l_hartree(species(atom)), lowered while outer_potential_radius(l) < d^2;
//   i_r_log = 1 + log(|d| / r_grid_min(s)) / log_r_grid_inc(s), s = species_center(c);
//   dV(p, t) += sum_{lm < (l+1)^2} Y_lm(d / |d|) S_lm(i_r_log),
// S_lm the cubic Hermite interpolant on the log-grid index of the tile (lm, 2, n_hartree_grid, atom, t)
// (value and d/di at each grid point), interval index clamped to [1, n_grid(s) - 1].
//
// Work layout: a wave owns 64 consecutive valid points (batches are spatially compact, so the
// wave's points are close together).  It scans the centres 64 at a time against the bounding box
// of its points (one centre per lane, ballot), and for each candidate centre appends the pairs that
// are near (d^2 < multipole_radius_sq, the reference's test) to a pair queue in LDS, centre-major.
// Whenever 64 pairs are queued, each lane evaluates one pair -- Y_lm by the row recurrence in
// registers, fused with the four spline dot products over the contiguous 4 (l_pot_max+1)^2-double
// coefficient block of its interval -- and adds the value to its point's LDS accumulator.
// Neighbouring lanes then share a centre and mostly nearby grid intervals, so the coefficient
// gathers of one load instruction fall on few cache lines, and every lane does useful work whatever
// the number of near centres per point.  One wave owns each point, so the accumulation is
// deterministic; at the end each point's sum is added to delta_v_hartree (the output is accumulated
// into, as in the reference).
//
// Y_lm (SHEval convention, index l^2 + l + m): with u = d / |d|, C_m + i S_m = (u_x + i u_y)^m and
// Pbar_l^m(z) = N_lm P_l^m(z) / sin^m (no Condon-Shortley phase, N_lm = sqrt((2l+1)/(4 pi)
// (l-m)!/(l+m)!)):  Y_l0 = Pbar_l^0, Y_lm = sqrt(2) (-1)^m Pbar_l^m C_m, Y_l,-m = sqrt(2) Pbar_l^m S_m.
// Pbar_m^m = N_mm (2m-1)!!, Pbar_{m+1}^m = sqrt(2m+3) z Pbar_m^m,
// Pbar_l^m = a_lm (z Pbar_{l-1}^m - b_lm Pbar_{l-2}^m); the constants are compile-time tables.
#include "farfield_common.hpp"

namespace frpa {
namespace {

constexpr int NEAR_WAVES = 4;   // waves per block
constexpr int NEAR_Q = 128;     // pair-queue capacity per wave (a flush happens once 64 are queued)

constexpr double csqrt(double x) {  // compile-time square root (Newton from above; x >= 0)
  if (x <= 0.0)
    return 0.0;
  double r = x > 1.0 ? x : 1.0;
  for (int it = 0; it < 200; it++) {
    const double n = 0.5 * (r + x / r);
    if (n >= r)
      break;
    r = n;
  }
  return r;
}

struct ShTab {
  double pmm[FAR_LMAX + 1];                  // Pbar_m^m
  double a[FAR_LMAX + 1][FAR_LMAX + 1];      // a_lm
  double b[FAR_LMAX + 1][FAR_LMAX + 1];      // b_lm
};

constexpr ShTab make_sh_tab() {
  ShTab t{};
  t.pmm[0] = csqrt(1.0 / (4.0 * 3.14159265358979323846));
  for (int m = 1; m <= FAR_LMAX; m++)
    t.pmm[m] = t.pmm[m - 1] * csqrt((2.0 * m + 1.0) / (2.0 * m));
  for (int l = 0; l <= FAR_LMAX; l++)
    for (int m = 0; m < l; m++) {
      t.a[l][m] = csqrt((4.0 * l * l - 1.0) / (double(l) * l - double(m) * m));
      t.b[l][m] = csqrt((double(l - 1) * (l - 1) - double(m) * m) / (4.0 * (l - 1) * (l - 1) - 1.0));
    }
  return t;
}
__device__ constexpr ShTab SH = make_sh_tab();
constexpr double SQRT2 = 1.41421356237309504880;

// Row l of the recurrence and its 2l + 1 terms: s_k += Y_lm coefficient_k(lm), k = 0..3 (the four
// coefficient planes of the interval), then row l + 1 while l + 1 <= lp.  Compile-time l and m make
// every table constant and coefficient offset an immediate.
template <int LMAX, int l>
__device__ __forceinline__ void near_row(const double (&C)[LMAX + 1], const double (&S)[LMAX + 1], double z,
                                         const double (&prev)[LMAX + 1], const double (&cur)[LMAX + 1], int lp,
                                         const double *__restrict__ g, double (&s)[4]) {
  constexpr int L2 = (LMAX + 1) * (LMAX + 1);
  double row[LMAX + 1];
#pragma unroll
  for (int m = 0; m <= l; m++) {
    if (m == l)
      row[m] = SH.pmm[m];
    else if (m == l - 1)
      row[m] = SH.a[l][m] * z * cur[m];
    else
      row[m] = SH.a[l][m] * (z * cur[m] - SH.b[l][m] * prev[m]);
  }
  const double *__restrict__ q = g + l * l + l;
#pragma unroll
  for (int m = -l; m <= l; m++) {
    double Y;
    if (m == 0)
      Y = row[0];
    else if (m > 0)
      Y = ((m & 1) ? -SQRT2 : SQRT2) * row[m] * C[m];
    else
      Y = SQRT2 * row[-m] * S[-m];
#pragma unroll
    for (int k = 0; k < 4; k++)
      s[k] = fma(Y, q[m + k * L2], s[k]);
  }
  if constexpr (l < LMAX) {
    if (l + 1 <= lp)
      near_row<LMAX, l + 1>(C, S, z, cur, row, lp, g, s);
  }
}

// One near pair: sum_{lm < (lp+1)^2} Y_lm(u) S_lm(i_r_log).  blk = the (lm, 2, grid) tile of the pair's
// (atom, t), lm extent (LMAX + 1)^2.
template <int LMAX>
__device__ __forceinline__ double near_pair(double dx, double dy, double dz, double d2, int lp,
                                            const NearSpecies sp, const double *__restrict__ blk) {
  constexpr int L2 = (LMAX + 1) * (LMAX + 1);
  const double rinv = rsqrt_f64(d2);
  const double x = dx * rinv, y = dy * rinv, z = dz * rinv;
  const double irl = 1.0 + (0.5 * log(d2) - sp.log_rmin) * sp.inv_inc;
  // int(i_r_log) clamped to [1, n_points - 1]; clamping in FP64 first keeps the conversion in range
  // for any i_r_log (d = 0 gives -inf and lands on interval 1, as max(1, int(i_r_log)) does)
  const int i = (int)fmin(fmax(irl, 1.0), (double)(sp.n_points - 1));
  const double t = irl - (double)i;
  const double tm = t - 1.0;
  // coefficient block of interval i: c0(:, i-1), c1(:, i-1), c0(:, i), c1(:, i), each L2 long
  const double *__restrict__ g = blk + (size_t)(2 * L2) * (i - 1);
  double C[LMAX + 1], S[LMAX + 1];
  C[0] = 1.0;
  S[0] = 0.0;
#pragma unroll
  for (int m = 1; m <= LMAX; m++) {
    C[m] = x * C[m - 1] - y * S[m - 1];
    S[m] = x * S[m - 1] + y * C[m - 1];
  }
  double zero[LMAX + 1];
#pragma unroll
  for (int m = 0; m <= LMAX; m++)
    zero[m] = 0.0;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  near_row<LMAX, 0>(C, S, z, zero, zero, lp, g, s);
  const double ta = tm * tm * (1.0 + 2.0 * t);
  const double tb = tm * tm * t;
  const double tc = t * t * (3.0 - 2.0 * t);
  const double td = t * t * tm;
  return ta * s[0] + tb * s[1] + tc * s[2] + td * s[3];
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LMAX>
__global__ __launch_bounds__(64 * NEAR_WAVES) void near_kernel(
    const CenterRec *__restrict__ centers, int n_hc, const NearSpecies *__restrict__ nspec,
    const double *__restrict__ coord_points, const int *__restrict__ map, int n_valid,
    const double *__restrict__ opr, int lp1, const double *__restrict__ mrsq, int n_atoms,
    const double *__restrict__ spl, int n_hgrid, double *__restrict__ dv, int n_full) {
  constexpr int L2 = (LMAX + 1) * (LMAX + 1);
  __shared__ double s_pt[NEAR_WAVES][3][64];
  __shared__ double s_acc[NEAR_WAVES][64];
  __shared__ int s_q[NEAR_WAVES][NEAR_Q];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.y;
  const long long iv = ((long long)blockIdx.x * NEAR_WAVES + w) * 64 + lane;
  const bool valid = iv < n_valid;
  const int p = valid ? map[iv] : 0;
  double px = 0.0, py = 0.0, pz = 0.0;
  if (valid) {
    px = coord_points[3 * (size_t)p];
    py = coord_points[3 * (size_t)p + 1];
    pz = coord_points[3 * (size_t)p + 2];
  }
  if (__ballot(valid) == 0ull)
    return;  // whole wave past the end (uniform)
  s_pt[w][0][lane] = px;
  s_pt[w][1][lane] = py;
  s_pt[w][2][lane] = pz;
  s_acc[w][lane] = 0.0;
  const double inf = __builtin_huge_val();
  const double lox = wave_min(valid ? px : inf), hix = wave_max(valid ? px : -inf);
  const double loy = wave_min(valid ? py : inf), hiy = wave_max(valid ? py : -inf);
  const double loz = wave_min(valid ? pz : inf), hiz = wave_max(valid ? pz : -inf);
  const double *__restrict__ mr = mrsq + (size_t)n_atoms * t;
  const double *__restrict__ op = opr + (size_t)lp1 * n_atoms * t;
  const double *__restrict__ sp_t = spl + (size_t)L2 * 2 * n_hgrid * n_atoms * t;
  const unsigned long long below = (1ull << lane) - 1ull;
  wave_sync();

  // evaluate queue entries 0 .. cnt-1 (one per lane) and add them to their points
  auto flush = [&](int cnt) {
    if (lane < cnt) {
      const int e = s_q[w][lane];
      const int ic = e >> 6, pl = e & 63;
      const CenterRec c = centers[ic];
      const double dx = s_pt[w][0][pl] - c.dx, dy = s_pt[w][1][pl] - c.dy, dz = s_pt[w][2][pl] - c.dz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      int l = min(c.l_const, LMAX);
      while (l > 0 && op[l + (size_t)lp1 * c.atom] < d2)
        l--;
      const double v = near_pair<LMAX>(dx, dy, dz, d2, l, nspec[c.spc], sp_t + (size_t)L2 * 2 * n_hgrid * c.atom);
      atomicAdd(&s_acc[w][pl], v);
    }
    wave_sync();
  };

  // fill the queue from the candidate centres until 64 pairs wait or the centres are exhausted,
  // then evaluate up to 64 of them (one call site of the pair evaluation)
  int qn = 0, cb = 0, cbase = 0;
  unsigned long long cm = 0ull;
  for (;;) {
    while (qn < 64) {
      if (cm == 0ull) {
        if (cb >= n_hc)
          break;
        const int icl = cb + lane;
        bool cand = false;
        if (icl < n_hc) {
          const CenterRec c = centers[icl];
          const double ex = fmax(fmax(lox - c.dx, c.dx - hix), 0.0);
          const double ey = fmax(fmax(loy - c.dy, c.dy - hiy), 0.0);
          const double ez = fmax(fmax(loz - c.dz, c.dz - hiz), 0.0);
          // the box distance bounds every point's distance from below (the slack covers the rounding
          // of both sums), so no near pair of this wave is dropped
          cand = (ex * ex + ey * ey + ez * ez) * (1.0 - 1e-12) < mr[c.atom];
        }
        cm = __ballot(cand);
        cbase = cb;
        cb += 64;
        continue;
      }
      const int ic = cbase + __builtin_ctzll(cm);
      cm &= cm - 1ull;
      const CenterRec c = centers[ic];
      const double dx = px - c.dx, dy = py - c.dy, dz = pz - c.dz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      const bool near = valid && d2 < mr[c.atom];
      const unsigned long long nm = __ballot(near);
      if (near)
        s_q[w][qn + __popcll(nm & below)] = (ic << 6) | lane;
      qn += __popcll(nm);
    }
    const int cnt = min(qn, 64);
    if (cnt == 0)
      break;
    wave_sync();
    flush(cnt);
    int rest = 0;
    if (lane < qn - cnt)
      rest = s_q[w][cnt + lane];
    wave_sync();
    if (lane < qn - cnt)
      s_q[w][lane] = rest;
    wave_sync();
    qn -= cnt;
  }
  if (valid)
    dv[(size_t)p + (size_t)n_full * t] += s_acc[w][lane];
}

template <int LMAX>
void launch_near(FarPlan &P, const FarCall &c, double *dv, hipStream_t s) {
  const long long blocks = ((long long)c.n_valid + 64 * NEAR_WAVES - 1) / (64 * NEAR_WAVES);
  FRPA_REQUIRE(blocks <= 0x7fffffff, FRPA_ERR_INVALID, "near field: too many points");
  dim3 grid((unsigned)blocks, c.n_tile);
  hipLaunchKernelGGL(near_kernel<LMAX>, grid, dim3(64 * NEAR_WAVES), 0, s, P.d_centers.ptr, P.n_hc, P.d_nspec.ptr,
                     c.coord_points, c.map, c.n_valid, c.opr, P.l_pot_max + 1, c.mrsq, P.n_atoms, c.spl, c.n_hgrid,
                     dv, c.n_full);
}

} // namespace

void set_near_tables(FarPlan &P, int n_species, const int *species_center, const int *n_grid,
                     const double *r_grid_min, const double *log_r_grid_inc) {
  FRPA_REQUIRE(n_species > 0 && species_center && n_grid && r_grid_min && log_r_grid_inc, FRPA_ERR_INVALID,
               "near field tables: null array or n_species %d", n_species);
  std::vector<NearSpecies> sp(n_species);
  for (int s = 0; s < n_species; s++) {
    FRPA_REQUIRE(r_grid_min[s] > 0.0 && log_r_grid_inc[s] > 0.0 && n_grid[s] >= 2, FRPA_ERR_INVALID,
                 "species %d: r_grid_min %g, log_r_grid_inc %g, n_grid %d (need > 0, > 0, >= 2)", s + 1,
                 r_grid_min[s], log_r_grid_inc[s], n_grid[s]);
    sp[s] = NearSpecies{std::log(r_grid_min[s]), 1.0 / log_r_grid_inc[s], n_grid[s], 0};
  }
  // centre ic's grid: species_center(centers_hartree_potential(ic))
  std::vector<CenterRec> cen = P.h_centers;
  int maxp = 0;
  for (size_t ic = 0; ic < cen.size(); ic++) {
    const int cur = P.h_chp[ic] - 1;
    const int s = species_center[cur] - 1;
    FRPA_REQUIRE(s >= 0 && s < n_species, FRPA_ERR_INVALID, "species_center(%d) = %d outside [1, %d]", cur + 1, s + 1,
                 n_species);
    cen[ic].spc = s;
    maxp = std::max(maxp, n_grid[s]);
  }
  bool same = P.near_ok && P.h_nspec.size() == sp.size();
  for (size_t q = 0; same && q < sp.size(); q++)
    same = sp[q].log_rmin == P.h_nspec[q].log_rmin && sp[q].inv_inc == P.h_nspec[q].inv_inc &&
           sp[q].n_points == P.h_nspec[q].n_points;
  for (size_t ic = 0; same && ic < cen.size(); ic++)
    same = cen[ic].spc == P.h_centers[ic].spc;
  if (same)
    return;  // the drop-in entry point passes the same tables on every call
  P.d_centers.upload(cen);
  P.h_centers = cen;
  P.d_nspec.upload(sp);
  P.h_nspec = sp;
  P.near_max_points = maxp;
  P.near_ok = true;
}

void near_forward(FarPlan &P, const FarCall &c, double *dv, hipStream_t s) {
  FRPA_REQUIRE(P.near_ok, FRPA_ERR_STATE, "near field: species grids not set (frpa_far_plan_set_near)");
  FRPA_REQUIRE(c.n_hgrid >= P.near_max_points, FRPA_ERR_INVALID,
               "near field: n_hartree_grid %d < n_grid %d of a centre's species", c.n_hgrid, P.near_max_points);
  if (c.n_valid == 0 || P.n_hc == 0)
    return;
  FRPA_REQUIRE(c.spl, FRPA_ERR_INVALID, "near field: null spline tile");
  switch (P.l_pot_max) {
#define FRPA_NEAR_CASE(L)                                                                                          \
  case L:                                                                                                          \
    launch_near<L>(P, c, dv, s);                                                                                   \
    break;
    FRPA_NEAR_CASE(0)
    FRPA_NEAR_CASE(1)
    FRPA_NEAR_CASE(2)
    FRPA_NEAR_CASE(3)
    FRPA_NEAR_CASE(4)
    FRPA_NEAR_CASE(5)
    FRPA_NEAR_CASE(6)
    FRPA_NEAR_CASE(7)
    FRPA_NEAR_CASE(8)
    FRPA_NEAR_CASE(9)
    FRPA_NEAR_CASE(10)
#undef FRPA_NEAR_CASE
  default:
    FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "near field: l_pot_max %d > %d", P.l_pot_max, FAR_LMAX);
  }
  check_launch("near_kernel");
}

} // namespace frpa
