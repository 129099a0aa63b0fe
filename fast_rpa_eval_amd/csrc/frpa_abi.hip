// C-ABI of libfrpa: plan construction (host-side batch/centre preprocessing), the frpa_* entry
// points, and the reference's Fortran-ABI entry points (declared in include/frpa.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>

#include "farfield_common.hpp"
#include "rho1_common.hpp"

namespace frpa {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}
void clear_error() { g_last_error.clear(); }

__global__ void fill_f64_kernel(double *p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_i32_kernel(int *p, size_t n, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void fill_i32(int *p, size_t n, int v, hipStream_t s) {
  if (n == 0)
    return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  check_launch("fill_i32_kernel");
}

LaunchArgs launch_args(Rho1Plan &P, const LaunchCfg &cfg, int T, int blocks_per_cu, hipStream_t s, int n_pos) {
  if (n_pos < 0)
    n_pos = P.n_valid;
  LaunchArgs a{nullptr, nullptr, dim3(n_pos, T)};
  if (cfg.order == 1)
    a.order = P.d_order_spatial.ptr;
  else if (cfg.order == 2)
    a.order = P.d_order_items.ptr;
  const int grid = cfg.persistent >= N_XCD ? cfg.persistent : std::max(1, P.n_cu) * blocks_per_cu;
  if (cfg.persistent && grid >= N_XCD && (long long)grid < (long long)n_pos * T) {
    fill_i32(P.d_counters.ptr, N_XCD, 0, s);
    a.counters = P.d_counters.ptr;
    a.grid = dim3(grid / N_XCD * N_XCD, 1);
  }
  return a;
}

void build_item_order(Rho1Plan &P, int pch) {
  if (P.items_pch == pch)
    return;
  FRPA_REQUIRE(P.max_np <= 255 * pch, FRPA_ERR_UNSUPPORTED, "order 2: a batch of %d points has more than 255 passes",
               P.max_np);
  // items in the caller's batch order, the passes of a batch consecutive; cut into N_XCD contiguous
  // ranges and interleaved like the spatial order, so a batch's passes are claimed back to back on
  // one XCD and stream the same P' tiles through its L2
  std::vector<int> caller(P.n_valid);
  std::iota(caller.begin(), caller.end(), 0);
  std::stable_sort(caller.begin(), caller.end(), [&](int a, int b) { return P.h_desc[a].b < P.h_desc[b].b; });
  std::vector<int> seq;
  for (int di : caller)
    for (int q = 0; q * pch < P.h_desc[di].np; q++)
      seq.push_back((di << 8) | q);
  const int n = (int)seq.size();
  std::vector<int> order(std::max(n, 1), 0);
  int start = 0;
  for (int x = 0; x < N_XCD; x++) {
    const int len = n / N_XCD + (x < n % N_XCD ? 1 : 0);
    for (int k = 0; k < len; k++)
      order[N_XCD * k + x] = seq[start + k];
    start += len;
  }
  P.d_order_items.upload(order);
  P.n_items = n;
  P.items_pch = pch;
}

TileSched &build_tile_items(Rho1Plan &P, int pch, bool grad) {
  TileSched &ts = P.tiles[grad ? 1 : 0];
  if (ts.pch == pch)
    return ts;
  // row tiles grouped into contiguous ranges of about `work` K steps (a row tile I of nT costs
  // nT - I steps, nT with the gradient, plus ~2 for staging its diagonal chunk and its epilogue)
  int work = 16;
  if (const char *e = std::getenv("FRPA_TILE_WORK"))
    work = std::max(1, std::atoi(e));
  const int C = grad ? 4 : 1;
  std::vector<int> caller(P.n_valid);
  std::iota(caller.begin(), caller.end(), 0);
  std::stable_sort(caller.begin(), caller.end(), [&](int a, int b) { return P.h_desc[a].b < P.h_desc[b].b; });
  std::vector<TileItem> seq;
  std::vector<PassRec> recs;
  std::vector<long long> qoff;
  long long poff = 0;
  for (int di : caller) {
    const BatchDesc &d = P.h_desc[di];
    const int nT = (d.nc + 63) / 64, passes = (d.np + pch - 1) / pch;
    qoff.assign(passes, 0);
    for (int q = 0; q < passes; q++) {
      const int npc = std::min(pch, d.np - q * pch);
      recs.push_back(PassRec{di, q, nT, npc, poff});
      qoff[q] = poff;
      poff += (long long)nT * C * npc;
    }
    // ranges of row tiles, each pass of a range an item: a batch's items are consecutive, the
    // passes of one range adjacent (they read the same P' tiles at about the same time)
    int I0 = 0, w = 0;
    for (int I = 0; I < nT; I++) {
      w += (grad ? nT : nT - I) + 2;
      if (w >= work || I == nT - 1) {
        for (int q = 0; q < passes; q++)
          seq.push_back(TileItem{di, q, I0, I + 1, qoff[q]});
        I0 = I + 1;
        w = 0;
      }
    }
  }
  const int n = (int)seq.size();
  std::vector<TileItem> items(std::max(n, 1), TileItem{0, 0, 0, 0, 0});
  int start = 0;
  for (int x = 0; x < N_XCD; x++) {
    const int len = n / N_XCD + (x < n % N_XCD ? 1 : 0);
    for (int k = 0; k < len; k++)
      items[N_XCD * k + x] = seq[start + k];
    start += len;
  }
  ts.d_items.upload(items);
  if (recs.empty())
    recs.push_back(PassRec{0, 0, 0, 0, 0});
  ts.d_recs.upload(recs);
  ts.n_items = n;
  ts.n_recs = (int)recs.size();
  ts.part_doubles = poff;
  ts.pch = pch;
  return ts;
}

void fill_f64(double *p, size_t n, double v, hipStream_t s) {
  if (n == 0)
    return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  check_launch("fill_f64_kernel");
}

Rho1Plan::~Rho1Plan() {
  if (ozaki)
    ozaki_release(ozaki);
  if (adj)
    adj_release(adj);
  if (adjoz)
    adjoz_release(adjoz);
  if (lowp)
    lowp_release(lowp);
  delete relabel;
}

const Relabel &get_relabel(Rho1Plan &P) {
  if (P.relabel)
    return *P.relabel;
  std::unique_ptr<Relabel> R(new Relabel());
  const int nb = P.n_basis;
  std::vector<int> order(P.n_valid);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return P.h_desc[a].b < P.h_desc[b].b; });
  R->h_pi.assign(nb, -1);
  int next = 0;
  for (int di : order) {
    const BatchDesc &d = P.h_desc[di];
    for (int i = 0; i < d.nc; i++) {
      const int mu = P.h_ibasis0[d.ib_off + i];
      if (R->h_pi[mu] < 0)
        R->h_pi[mu] = next++;
    }
  }
  for (int mu = 0; mu < nb; mu++)
    if (R->h_pi[mu] < 0)
      R->h_pi[mu] = next++;
  R->h_pinv.resize(nb);
  for (int mu = 0; mu < nb; mu++)
    R->h_pinv[R->h_pi[mu]] = mu;
  R->h_srow.resize(P.h_ibasis0.size());
  R->h_sib.resize(P.h_ibasis0.size());
  R->h_spos.resize(P.h_ibasis0.size());
  for (const BatchDesc &d : P.h_desc) {
    const int *ib = P.h_ibasis0.data() + d.ib_off;
    int *sr = R->h_srow.data() + d.ib_off;
    std::iota(sr, sr + d.nc, 0);
    std::sort(sr, sr + d.nc, [&](int a, int b) { return R->h_pi[ib[a]] < R->h_pi[ib[b]]; });
    for (int k = 0; k < d.nc; k++) {
      R->h_sib[d.ib_off + k] = R->h_pi[ib[sr[k]]];
      R->h_spos[d.ib_off + sr[k]] = k;
    }
  }
  R->d_pi.upload(R->h_pi);
  R->d_pinv.upload(R->h_pinv);
  R->d_srow.upload(R->h_srow.empty() ? std::vector<int>{0} : R->h_srow);
  R->d_sib.upload(R->h_sib.empty() ? std::vector<int>{0} : R->h_sib);
  R->d_spos.upload(R->h_spos.empty() ? std::vector<int>{0} : R->h_spos);
  P.relabel = R.release();
  return *P.relabel;
}

// ---------------------------------------------------------------------------------------------
// rho1 plan
// ---------------------------------------------------------------------------------------------
Rho1Plan *make_rho1_plan(int n_batches, int n_basis, int n_full_points, int n_max_batch_size,
                                const int *n_point_batches, const int *prefix, const int *map, const int *ncc,
                                const long long *ibasis_ptr, const int *ibasis, const long long *wave_offset) {
  FRPA_REQUIRE(n_batches >= 0 && n_basis > 0 && n_full_points >= 0 && n_max_batch_size >= 0, FRPA_ERR_INVALID,
               "rho1 plan: bad sizes (n_batches %d, n_basis %d, n_full_points %d, n_max_batch_size %d)", n_batches,
               n_basis, n_full_points, n_max_batch_size);
  FRPA_REQUIRE(n_batches == 0 || (n_point_batches && prefix && ncc && ibasis_ptr && wave_offset), FRPA_ERR_INVALID,
               "rho1 plan: null metadata array");
  std::unique_ptr<Rho1Plan> P(new Rho1Plan());
  P->n_batches = n_batches;
  P->n_basis = n_basis;
  P->n_full = n_full_points;
  P->nmb = n_max_batch_size;
  P->wave_len = n_batches ? wave_offset[n_batches] : 0;
  const int n_valid_points = n_batches ? prefix[n_batches] : 0;
  FRPA_REQUIRE(n_valid_points >= 0 && n_valid_points <= n_full_points, FRPA_ERR_INVALID,
               "rho1 plan: %d valid points for %d full points", n_valid_points, n_full_points);
  for (int i = 0; i < n_valid_points; i++)
    FRPA_REQUIRE(map[i] >= 0 && map[i] < n_full_points, FRPA_ERR_INVALID, "rho1 plan: map[%d] = %d out of range",
                 i, map[i]);
  std::vector<BatchDesc> desc;
  for (int b = 0; b < n_batches; b++) {
    const int nc = ncc[b], np = n_point_batches[b];
    FRPA_REQUIRE(nc >= 0 && np >= 0 && np <= n_max_batch_size, FRPA_ERR_INVALID,
                 "rho1 plan: batch %d has n_compute_c %d, n_points %d (n_max_batch_size %d)", b, nc, np,
                 n_max_batch_size);
    FRPA_REQUIRE(prefix[b + 1] - prefix[b] == np, FRPA_ERR_INVALID, "rho1 plan: prefix sum disagrees at batch %d", b);
    FRPA_REQUIRE(ibasis_ptr[b + 1] - ibasis_ptr[b] == nc, FRPA_ERR_INVALID,
                 "rho1 plan: batch %d lists %lld basis functions for n_compute_c %d", b,
                 ibasis_ptr[b + 1] - ibasis_ptr[b], nc);
    const long long ncp = (nc + 3) / 4 * 4;
    FRPA_REQUIRE(wave_offset[b + 1] - wave_offset[b] >= ncp * np, FRPA_ERR_INVALID,
                 "rho1 plan: wave block of batch %d too short", b);
    if (nc == 0 || np == 0)
      continue;  // the reference skips n_compute_c == 0; an empty batch writes nothing
    for (int i = 0; i < nc; i++) {
      const int v = ibasis[ibasis_ptr[b] + i];
      FRPA_REQUIRE(v >= 1 && v <= n_basis, FRPA_ERR_INVALID, "rho1 plan: i_basis %d of batch %d out of range", v, b);
    }
    desc.push_back(BatchDesc{wave_offset[b], 0, nc, np, prefix[b], b});
  }
  // BatchDesc order: descending work nc^2 np (longest first); the XCD-aware spatial order is a
  // separate position map (d_order_spatial)
  std::stable_sort(desc.begin(), desc.end(), [](const BatchDesc &x, const BatchDesc &y) {
    return (double)x.nc * x.nc * x.np > (double)y.nc * y.nc * y.np;
  });
  // compact 0-based basis lists, in processing order
  long long off = 0;
  for (auto &d : desc) {
    const int b = d.b;
    d.ib_off = off;
    for (int i = 0; i < d.nc; i++)
      P->h_ibasis0.push_back(ibasis[ibasis_ptr[b] + i] - 1);
    off += d.nc;
    P->max_np = std::max(P->max_np, d.np);
    P->max_nc = std::max(P->max_nc, d.nc);
  }
  P->n_valid = (int)desc.size();
  P->wave_off_quads = std::all_of(desc.begin(), desc.end(), [](const BatchDesc &d) { return d.wave_off % 4 == 0; });
  P->h_desc = desc;
  P->d_desc.upload(desc);
  {
    // XCD-aware spatial order: caller order cut into N_XCD contiguous ranges, position 8 k + x
    // takes the k-th batch of range x (range lengths differ by at most one)
    std::vector<int> caller(P->n_valid);
    std::iota(caller.begin(), caller.end(), 0);
    std::stable_sort(caller.begin(), caller.end(), [&](int a, int b) { return desc[a].b < desc[b].b; });
    std::vector<int> order(std::max(P->n_valid, 1), 0);
    const int n = P->n_valid;
    int start = 0;
    for (int x = 0; x < N_XCD; x++) {
      const int len = n / N_XCD + (x < n % N_XCD ? 1 : 0);
      for (int k = 0; k < len; k++)
        order[N_XCD * k + x] = caller[start + k];
      start += len;
    }
    P->d_order_spatial.upload(order);
    P->d_counters.alloc(N_XCD);
    int dev = 0;
    FRPA_HIP(hipGetDevice(&dev));
    FRPA_HIP(hipDeviceGetAttribute(&P->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  P->d_ibasis0.upload(P->h_ibasis0);
  P->d_map.upload(map, (size_t)n_valid_points);
  // default launches: what the autotuner picks on the 60k-batch, 1024-atom workload (FP64: XCD-aware
  // spatial order, persistent, 13 point groups per wave; INT8 split: row-tile ranges of point passes
  // as work items, persistent, four-wave workgroups with direct global -> LDS staging; the FP32 /
  // BF16x3 comparison paths share that launch)
  P->cfg[FRPA_PREC_FP64] = LaunchCfg{1, 1, 13};
  P->cfg[FRPA_PREC_OZAKI] = LaunchCfg{3, 1, 1};
  P->cfg[FRPA_PREC_OZAKI4] = LaunchCfg{3, 1, 1};
  P->cfg[FRPA_PREC_FP32] = LaunchCfg{3, 1, 1};
  P->cfg[FRPA_PREC_BF16X3] = LaunchCfg{3, 1, 1};
  P->cfg[FRPA_PREC_FP64_SYM] = LaunchCfg{3, 1, 1};
  return P.release();
}

// ---------------------------------------------------------------------------------------------
// far-field plan
// ---------------------------------------------------------------------------------------------
static int odd_double_factorial(int n) {  // (2n-1)!!, (-1)!! = 1
  int r = 1;
  for (int v = 2 * n - 1; v > 0; v -= 2)
    r *= v;
  return r;
}

FarPlan *make_far_plan(int l_pot_max, int lama, int dim0, int n_centers, int n_hc, int n_atoms, int n_species,
                              int hfla, const int *species, const int *l_hartree, const int *chp,
                              const int *center_to_atom, const double *coords_center, const int *index_cc,
                              const double *cc, const int *n_cc_lm_ijk, const int *index_ijk_max_cc) {
  (void)index_ijk_max_cc;
  FRPA_REQUIRE(l_pot_max >= 0 && l_pot_max <= FAR_LMAX, FRPA_ERR_UNSUPPORTED, "l_pot_max %d outside [0, %d]",
               l_pot_max, FAR_LMAX);
  FRPA_REQUIRE(lama >= l_pot_max && n_centers >= 0 && n_hc >= 0 && n_atoms > 0 && n_species > 0 && dim0 > 0,
               FRPA_ERR_INVALID, "far plan: bad sizes");
  std::unique_ptr<FarPlan> P(new FarPlan());
  P->l_pot_max = l_pot_max;
  P->lama = lama;
  P->dim0 = dim0;
  P->n_centers = n_centers;
  P->n_hc = n_hc;
  P->n_atoms = n_atoms;
  P->n_species = n_species;
  P->hfla = hfla;
  P->h_ncc.assign(n_cc_lm_ijk, n_cc_lm_ijk + lama + 1);
  P->mc_dim0 = n_cc_lm_ijk[l_pot_max];
  FRPA_REQUIRE(P->h_ncc[lama] <= dim0, FRPA_ERR_INVALID, "n_cc_lm_ijk(l_max_analytic_multipole) > index_cc_dim_0");
  for (int s = 0; s < n_species; s++)
    FRPA_REQUIRE(l_hartree[s] >= 0 && l_hartree[s] <= l_pot_max, FRPA_ERR_INVALID,
                 "l_hartree(%d) = %d outside [0, l_pot_max]", s + 1, l_hartree[s]);
  // term table rows n < n_cc(l_pot_max): (l, m, i, j, k, nn)
  auto col = [&](int n, int c) { return index_cc[n + (size_t)dim0 * c]; };
  for (int n = 0; n < P->h_ncc[lama]; n++)
    FRPA_REQUIRE(col(n, 0) >= 0 && col(n, 0) <= lama && std::abs(col(n, 1)) <= col(n, 0), FRPA_ERR_INVALID,
                 "index_cc row %d: (l, m) = (%d, %d) invalid", n, col(n, 0), col(n, 1));
  bool same = true;
  for (int l = 0; l <= l_pot_max; l++)
    same = same && P->h_ncc[l] == FAR_NCC[l];
  for (int n = 0; same && n < P->mc_dim0; n++) {
    const FarTerm &T = FAR_TERMS[n];
    same = col(n, 0) == T.l && col(n, 1) == T.m && col(n, 2) == T.i && col(n, 3) == T.j && col(n, 4) == T.k &&
           col(n, 5) == T.nn;
  }
  P->terms_ok = same;
  // centres
  std::vector<CenterRec> cen(n_hc);
  std::vector<std::vector<int>> atom_ic(n_atoms);
  for (int ic = 0; ic < n_hc; ic++) {
    const int cur = chp[ic] - 1;
    FRPA_REQUIRE(cur >= 0 && cur < n_centers && ic < n_centers, FRPA_ERR_INVALID,
                 "centers_hartree_potential(%d) = %d out of range", ic + 1, chp[ic]);
    const int atom = center_to_atom[cur] - 1;
    const int mca = center_to_atom[ic] - 1;
    FRPA_REQUIRE(atom >= 0 && atom < n_atoms && mca >= 0 && mca < n_atoms, FRPA_ERR_INVALID,
                 "center_to_atom out of range at centre %d", ic + 1);
    const int sp = species[atom] - 1;
    FRPA_REQUIRE(sp >= 0 && sp < n_species, FRPA_ERR_INVALID, "species(%d) out of range", atom + 1);
    cen[ic] = CenterRec{coords_center[3 * cur], coords_center[3 * cur + 1], coords_center[3 * cur + 2],
                        coords_center[3 * ic], coords_center[3 * ic + 1], coords_center[3 * ic + 2],
                        atom, mca, l_hartree[sp], -1};
    atom_ic[mca].push_back(ic);
  }
  // merged nested-Horner gather plan: for each l and each (i, j, k, q) the terms n < n_cc(l) feeding it
  P->coef_stride = horner_prefix(l_pot_max + 1);
  std::vector<std::vector<std::pair<int, double>>> lists(P->coef_stride);
  bool horner_ok = true;
  for (int L = 0; L <= l_pot_max; L++) {
    for (int n = 0; n < P->h_ncc[L]; n++) {
      const int i = col(n, 2), j = col(n, 3), k = col(n, 4), nn = col(n, 5), tl = col(n, 0);
      const int d = i + j + k, q = nn - d;
      if (i < 0 || i > 1 || j < 0 || k < 0 || q < 0 || d + 2 * q > L || tl > L) {
        horner_ok = false;
        continue;
      }
      const double sf = ((nn & 1) ? -1.0 : 1.0) * odd_double_factorial(nn);
      lists[horner_prefix(L) + horner_index(L, q, i, j, k)].push_back({n, sf * cc[n]});
    }
  }
  FRPA_REQUIRE(horner_ok, FRPA_ERR_UNSUPPORTED,
               "index_cc rows are not of the form x^i y^j z^k Fp(nn), i <= 1, nn = (l + i + j + k) / 2");
  std::vector<int> hp(1, 0), ht;
  std::vector<double> hw;
  for (auto &v : lists) {
    for (auto &pr : v) {
      ht.push_back(pr.first);
      hw.push_back(pr.second);
    }
    hp.push_back((int)ht.size());
  }
  std::vector<int> aip(1, 0), ai;
  for (auto &v : atom_ic) {
    ai.insert(ai.end(), v.begin(), v.end());
    aip.push_back((int)ai.size());
  }
  P->d_centers.upload(cen);
  P->h_centers = cen;
  P->h_chp.assign(chp, chp + n_hc);
  P->d_index_cc.upload(index_cc, (size_t)dim0 * 6);
  std::vector<double> ccp(std::max<size_t>(dim0, FAR_NCC[FAR_LMAX]), 0.0);
  std::copy(cc, cc + dim0, ccp.begin());
  P->d_cc.upload(ccp);
  P->d_ncc.upload(P->h_ncc);
  P->d_hc_ptr.upload(hp);
  P->d_hc_term.upload(ht.empty() ? std::vector<int>{0} : ht);
  P->d_hc_w.upload(hw.empty() ? std::vector<double>{0.0} : hw);
  P->d_atom_ic_ptr.upload(aip);
  P->d_atom_ic.upload(ai.empty() ? std::vector<int>{0} : ai);
  return P.release();
}

} // namespace frpa

using namespace frpa;

// =============================================================================================
// frpa_* C-ABI
// =============================================================================================
extern "C" {

const char *frpa_last_error(void) { return g_last_error.c_str(); }
const char *frpa_version(void) { return "fast_rpa_eval_amd 0.1 (gfx950)"; }

int frpa_rho1_plan_create(frpa_rho1_plan **plan, int n_batches, int n_basis, int n_full_points, int n_max_batch_size,
                          const int *n_point_batches, const int *n_point_batches_prefix_sum,
                          const int *i_valid_point_2_i_full_points_map, const int *n_compute_c_batches,
                          const long long *ibasis_ptr, const int *ibasis, const long long *wave_offset) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "plan is null");
    *plan = reinterpret_cast<frpa_rho1_plan *>(make_rho1_plan(
        n_batches, n_basis, n_full_points, n_max_batch_size, n_point_batches, n_point_batches_prefix_sum,
        i_valid_point_2_i_full_points_map, n_compute_c_batches, ibasis_ptr, ibasis, wave_offset));
  });
}

int frpa_rho1_plan_destroy(frpa_rho1_plan *plan) {
  return guarded([&] { delete reinterpret_cast<Rho1Plan *>(plan); });
}

int frpa_rho1_forward(frpa_rho1_plan *plan, int atom_tile_size, int n_tile, int precision, const double *wave,
                      const double *density_matrix, const double *rho_bias, double *first_order_rho,
                      const double *gradient_wave, const double *grad_bias, double *first_order_gradient_rho,
                      void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "plan is null");
    Rho1Plan &P = *reinterpret_cast<Rho1Plan *>(plan);
    FRPA_REQUIRE(atom_tile_size >= 1 && n_tile >= 0 && n_tile <= atom_tile_size, FRPA_ERR_INVALID,
                 "atom_tile_size %d, n_tile %d", atom_tile_size, n_tile);
    FRPA_REQUIRE(first_order_rho && density_matrix && (wave || P.n_valid == 0), FRPA_ERR_INVALID,
                 "null device array");
    FRPA_REQUIRE(!gradient_wave || first_order_gradient_rho, FRPA_ERR_INVALID,
                 "gradient_wave given without first_order_gradient_rho");
    hipStream_t s = as_stream(stream);
    Rho1Call c;
    c.T = atom_tile_size;
    c.n_tile = n_tile;
    c.wave = wave;
    c.gwave = gradient_wave;
    c.P1 = density_matrix;
    c.bias = rho_bias;
    c.gbias = grad_bias;
    c.rho = first_order_rho;
    c.grad = gradient_wave ? first_order_gradient_rho : nullptr;
    // evaluate_first_order_rho.cpp:113-114: both outputs are zeroed first
    fill_f64(first_order_rho, (size_t)P.n_full * atom_tile_size, 0.0, s);
    if (first_order_gradient_rho)
      fill_f64(first_order_gradient_rho, (size_t)3 * P.nmb * P.n_batches * atom_tile_size, 0.0, s);
    if (precision == FRPA_PREC_FP64) {
      rho1_forward_fp64(P, c, s);
    } else if (precision == FRPA_PREC_OZAKI || precision == FRPA_PREC_OZAKI4) {
      rho1_forward_ozaki(P, c, precision == FRPA_PREC_OZAKI ? 5 : 4, s);
    } else if (precision == FRPA_PREC_FP32 || precision == FRPA_PREC_BF16X3 || precision == FRPA_PREC_FP64_SYM) {
      rho1_forward_lowp(P, c, precision, s);
    } else {
      FRPA_REQUIRE(false, FRPA_ERR_INVALID, "unknown precision %d", precision);
    }
  });
}

int frpa_rho1_adjoint(frpa_rho1_plan *plan, int atom_tile_size, int precision, const double *wave, const double *g,
                      double *first_order_H, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && g && first_order_H && atom_tile_size >= 1, FRPA_ERR_INVALID, "bad adjoint arguments");
    Rho1Plan &P = *reinterpret_cast<Rho1Plan *>(plan);
    rho1_adjoint(P, atom_tile_size, precision, wave, g, first_order_H, as_stream(stream));
  });
}

int frpa_rho1_set_launch(frpa_rho1_plan *plan, int precision, int order, int persistent, int points_per_wave) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    FRPA_REQUIRE(precision >= FRPA_PREC_FP64 && precision <= FRPA_PREC_FP64_SYM, FRPA_ERR_INVALID,
                 "unknown precision %d", precision);
    FRPA_REQUIRE(order == 0 || order == 1 || ((order == 2 || order == 3) && precision != FRPA_PREC_FP64),
                 FRPA_ERR_INVALID, "order must be 0 or 1 (or 2, 3 for the INT8 split, FP32 and BF16x3)");
    FRPA_REQUIRE(persistent == 0 || persistent == 1 || persistent >= N_XCD, FRPA_ERR_INVALID,
                 "persistent must be 0, 1 or a grid of at least %d workgroups", N_XCD);
    const bool ok = points_per_wave == 0 ||
                    (precision == FRPA_PREC_FP64 && (points_per_wave == 4 || points_per_wave == 8 ||
                                                     points_per_wave == 13)) ||
                    (precision == FRPA_PREC_OZAKI &&
                     (points_per_wave == 4 || points_per_wave == 2 || points_per_wave == 1)) ||
                    (precision == FRPA_PREC_OZAKI4 && (points_per_wave == 4 || points_per_wave == 5 ||
                                                       points_per_wave == 2 || points_per_wave == 1)) ||
                    ((precision == FRPA_PREC_FP32 || precision == FRPA_PREC_BF16X3 ||
                      precision == FRPA_PREC_FP64_SYM) &&
                     points_per_wave == 1);
    FRPA_REQUIRE(ok, FRPA_ERR_INVALID, "%d point groups per wave not available for precision %d", points_per_wave,
                 precision);
    reinterpret_cast<Rho1Plan *>(plan)->cfg[precision] = LaunchCfg{order, persistent, points_per_wave};
  });
}

int frpa_rho1_autotune(frpa_rho1_plan *plan, int atom_tile_size, int precision, const double *wave,
                       const double *density_matrix, double *first_order_rho, int reps, char *choice,
                       int choice_len, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && density_matrix && first_order_rho && atom_tile_size >= 1 && reps >= 1, FRPA_ERR_INVALID,
                 "bad autotune arguments");
    FRPA_REQUIRE(precision >= FRPA_PREC_FP64 && precision <= FRPA_PREC_FP64_SYM, FRPA_ERR_INVALID,
                 "unknown precision %d", precision);
    Rho1Plan &P = *reinterpret_cast<Rho1Plan *>(plan);
    if (P.n_valid == 0)
      return;
    hipStream_t s = as_stream(stream);
    std::vector<int> npgs;
    if (precision == FRPA_PREC_FP64)
      npgs = {4, 8, 13};
    else if (precision == FRPA_PREC_OZAKI4)
      npgs = {4, 5, 2, 1};
    else if (precision == FRPA_PREC_OZAKI)
      npgs = {0, 2, 1};
    else
      npgs = {1};  // FP32 / BF16x3 / FP64 symmetric half: one block shape
    std::vector<LaunchCfg> cands;
    // INT8, FP32, BF16x3: order 2, point passes as items; order 3, row-tile ranges of point passes as items
    const int n_orders = precision == FRPA_PREC_FP64 ? 2 : 4;
    for (int order = 0; order < n_orders; order++)
      for (int pers = 0; pers < 2; pers++)
        for (int npg : npgs)
          cands.push_back(LaunchCfg{order, pers, npg});
    hipEvent_t e0, e1;
    FRPA_HIP(hipEventCreate(&e0));
    FRPA_HIP(hipEventCreate(&e1));
    std::vector<float> ms(cands.size());
    try {
      for (size_t i = 0; i < cands.size(); i++) {
        P.cfg[precision] = cands[i];
        auto run = [&] {
          FRPA_CHECK_OK(frpa_rho1_forward(plan, atom_tile_size, atom_tile_size, precision, wave, density_matrix,
                                          nullptr, first_order_rho, nullptr, nullptr, nullptr, stream));
        };
        run();
        FRPA_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++)
          run();
        FRPA_HIP(hipEventRecord(e1, s));
        FRPA_HIP(hipEventSynchronize(e1));
        FRPA_HIP(hipEventElapsedTime(&ms[i], e0, e1));
        ms[i] /= reps;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const size_t best = std::min_element(ms.begin(), ms.end()) - ms.begin();
    P.cfg[precision] = cands[best];
    if (choice && choice_len > 0) {
      std::string out;
      auto desc = [&](size_t i) {
        char buf[96];
        snprintf(buf, sizeof(buf), "order=%d persistent=%d npg=%d ms=%.3f", cands[i].order, cands[i].persistent,
                 cands[i].npg, ms[i]);
        return std::string(buf);
      };
      out = desc(best);
      for (size_t i = 0; i < cands.size(); i++)
        if (i != best)
          out += "; " + desc(i);
      snprintf(choice, (size_t)choice_len, "%s", out.c_str());
    }
  });
}

int frpa_rho1_prepare_wave(frpa_rho1_plan *plan, int precision, const double *wave, long long *bytes,
                           void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && wave, FRPA_ERR_INVALID, "bad arguments");
    FRPA_REQUIRE(precision != FRPA_PREC_FP64 && precision >= FRPA_PREC_OZAKI && precision <= FRPA_PREC_BF16X3,
                 FRPA_ERR_INVALID, "prepare_wave applies to the INT8-split, FP32 and BF16x3 precisions");
    Rho1Plan &P = *reinterpret_cast<Rho1Plan *>(plan);
    if (precision == FRPA_PREC_FP32 || precision == FRPA_PREC_BF16X3) {
      const long long b = lowp_prepare_wave(P, wave, precision, as_stream(stream));
      if (bytes)
        *bytes = b;
      return;
    }
    const int S = precision == FRPA_PREC_OZAKI ? 5 : 4;
    long long b = ozaki_prepare_wave(P, wave, S, as_stream(stream));
    // the adjoint keeps digits of its own: rebuild them too once the adjoint has made them
    if (adjoz_built(P)) {
      adjoz_prepare(P, S, wave, as_stream(stream));
      b += adjoz_bytes(P);
    }
    if (bytes)
      *bytes = b;
  });
}

int frpa_rho1_set_adjoint_wait_limit(frpa_rho1_plan *plan, int polls) {
  return guarded([&] {
    FRPA_REQUIRE(plan && polls >= 0, FRPA_ERR_INVALID, "bad arguments");
    reinterpret_cast<Rho1Plan *>(plan)->adj_spin_limit = polls;
  });
}

int frpa_rho1_adjoint_status(frpa_rho1_plan *plan, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    const int fails = adjoz_take_wait_failures(*reinterpret_cast<Rho1Plan *>(plan), as_stream(stream));
    FRPA_REQUIRE(fails == 0, FRPA_ERR_HIP,
                 "INT8 adjoint: %d slot waits gave up since the last check; the first_order_H of those calls is "
                 "incomplete",
                 fails);
  });
}

int frpa_synth_wave(frpa_rho1_plan *plan, int n_atoms, const double *atom_xyz, const int *basis_atom,
                    const int *basis_l, const int *basis_m, const double *basis_alpha, const double *coord_points,
                    double *wave, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && atom_xyz && basis_atom && basis_l && basis_m && basis_alpha && coord_points && wave,
                 FRPA_ERR_INVALID, "null argument");
    const Rho1Plan &P = *reinterpret_cast<Rho1Plan *>(plan);
    synth_wave(P, n_atoms, atom_xyz, basis_atom, basis_l, basis_m, basis_alpha, P.n_basis, coord_points, wave,
               as_stream(stream));
  });
}

int frpa_far_plan_create(frpa_far_plan **plan, int l_pot_max, int l_max_analytic_multipole, int index_cc_dim_0,
                         int n_centers, int n_centers_hartree_potential, int n_atoms, int n_species,
                         int hartree_force_l_add, const int *species, const int *l_hartree,
                         const int *centers_hartree_potential, const int *center_to_atom,
                         const double *coords_center, const int *index_cc, const double *cc,
                         const int *n_cc_lm_ijk, const int *index_ijk_max_cc) {
  return guarded([&] {
    FRPA_REQUIRE(plan && species && l_hartree && centers_hartree_potential && center_to_atom && coords_center &&
                     index_cc && cc && n_cc_lm_ijk,
                 FRPA_ERR_INVALID, "null argument");
    *plan = reinterpret_cast<frpa_far_plan *>(make_far_plan(
        l_pot_max, l_max_analytic_multipole, index_cc_dim_0, n_centers, n_centers_hartree_potential, n_atoms,
        n_species, hartree_force_l_add, species, l_hartree, centers_hartree_potential, center_to_atom, coords_center,
        index_cc, cc, n_cc_lm_ijk, index_ijk_max_cc));
  });
}

int frpa_far_plan_destroy(frpa_far_plan *plan) {
  return guarded([&] { delete reinterpret_cast<FarPlan *>(plan); });
}

static FarCall far_call(int T, int n_tile, int n_full, int n_valid, const double *cp, const int *map, const int *lmfd,
                        const double *mrsq, const double *opr, const double *M) {
  FRPA_REQUIRE(T >= 1 && n_tile >= 0 && n_tile <= T && n_full >= 0 && n_valid >= 0, FRPA_ERR_INVALID,
               "far: bad sizes (T %d, n_tile %d, n_full %d, n_valid %d)", T, n_tile, n_full, n_valid);
  FRPA_REQUIRE(n_valid == 0 || (cp && map && lmfd && mrsq && opr && M), FRPA_ERR_INVALID, "far: null array");
  FarCall c;
  c.T = T;
  c.n_tile = n_tile;
  c.n_full = n_full;
  c.n_valid = n_valid;
  c.coord_points = cp;
  c.map = map;
  c.lmfd = lmfd;
  c.mrsq = mrsq;
  c.opr = opr;
  c.moments = M;
  return c;
}

int frpa_far_forward(frpa_far_plan *plan, int atom_tile_size, int n_tile, int variant, int n_full_points,
                     int n_valid_points, const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                     const int *l_hartree_max_far_distance, const double *multipole_radius_sq,
                     const double *outer_potential_radius, const double *multipole_moments,
                     double *delta_v_hartree, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && delta_v_hartree, FRPA_ERR_INVALID, "null argument");
    FarCall c = far_call(atom_tile_size, n_tile, n_full_points, n_valid_points, coord_points,
                         i_valid_point_2_i_full_points_map, l_hartree_max_far_distance, multipole_radius_sq,
                         outer_potential_radius, multipole_moments);
    far_forward(*reinterpret_cast<FarPlan *>(plan), c, variant, delta_v_hartree, as_stream(stream));
  });
}

int frpa_far_forward_noalloc(frpa_far_plan *plan, int atom_tile_size, int n_tile, int variant, int n_full_points,
                             int n_valid_points, const double *coord_points,
                             const int *i_valid_point_2_i_full_points_map, const int *l_hartree_max_far_distance,
                             const double *multipole_radius_sq, const double *outer_potential_radius,
                             const double *multipole_moments, double *delta_v_hartree, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && delta_v_hartree, FRPA_ERR_INVALID, "null argument");
    FarCall c = far_call(atom_tile_size, n_tile, n_full_points, n_valid_points, coord_points,
                         i_valid_point_2_i_full_points_map, l_hartree_max_far_distance, multipole_radius_sq,
                         outer_potential_radius, multipole_moments);
    far_forward_noalloc(*reinterpret_cast<FarPlan *>(plan), c, variant, delta_v_hartree, as_stream(stream));
  });
}

int frpa_far_adjoint(frpa_far_plan *plan, int atom_tile_size, int n_tile, int n_full_points, int n_valid_points,
                     const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                     const int *l_hartree_max_far_distance, const double *multipole_radius_sq,
                     const double *outer_potential_radius, const double *multipole_moments, const double *gV,
                     double *grad_moments, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && gV && grad_moments, FRPA_ERR_INVALID, "null argument");
    FarCall c = far_call(atom_tile_size, n_tile, n_full_points, n_valid_points, coord_points,
                         i_valid_point_2_i_full_points_map, l_hartree_max_far_distance, multipole_radius_sq,
                         outer_potential_radius, multipole_moments);
    far_adjoint(*reinterpret_cast<FarPlan *>(plan), c, gV, grad_moments, as_stream(stream));
  });
}

int frpa_rho1_bias_part2(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                         int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                         const double *gradient_wave, const double *density_matrix,
                         double *first_order_rho_bias_part2, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    rho1_bias_part2(*reinterpret_cast<Rho1Plan *>(plan), atom_tile_size, j_atom_begin, n_tile, j_coord, n_atoms,
                    atom_valid_n_compute_c_batches, wave, gradient_wave, density_matrix, first_order_rho_bias_part2,
                    as_stream(stream));
  });
}

int frpa_rho1_bias_inline(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                          int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                          const double *gradient_wave, const double *density_matrix, double *first_order_rho_bias,
                          void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    rho1_bias_part2(*reinterpret_cast<Rho1Plan *>(plan), atom_tile_size, j_atom_begin, n_tile, j_coord, n_atoms,
                    atom_valid_n_compute_c_batches, wave, gradient_wave, density_matrix, first_order_rho_bias,
                    as_stream(stream), 1, true);
  });
}

int frpa_rho1_h1_bias_part2(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                            int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                            const double *gradient_wave, const double *H_times_psi, const double *vrho_batches,
                            const double *vsigma_batches, const double *gradient_rho, const double *partition_tab,
                            double *first_order_H_tile, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    rho1_h1_bias_part2(*reinterpret_cast<Rho1Plan *>(plan), atom_tile_size, j_atom_begin, n_tile, j_coord, n_atoms,
                       atom_valid_n_compute_c_batches, wave, gradient_wave, H_times_psi, vrho_batches, vsigma_batches,
                       gradient_rho, partition_tab, first_order_H_tile, as_stream(stream));
  });
}

int frpa_rho1_h1(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord, int n_atoms,
                 const int *atom_valid_n_compute_c_batches, const double *wave, const double *gradient_wave,
                 const double *hessian_wave, const double *vsigma_batches, const double *gradient_rho,
                 const double *H_prefactor1_batches_tile, const double *first_order_gradient_rho_batches_tile,
                 const double *v_hartree_gradient, const double *first_order_rho, const double *v2rho2_batches,
                 const double *v2rhosigma_batches, const double *v2sigma2_batches, const double *partition_tab,
                 double *first_order_H_tile, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    rho1_h1_gga(*reinterpret_cast<Rho1Plan *>(plan), atom_tile_size, j_atom_begin, n_tile, j_coord, n_atoms,
                atom_valid_n_compute_c_batches, wave, gradient_wave, hessian_wave, vsigma_batches, gradient_rho,
                H_prefactor1_batches_tile, first_order_gradient_rho_batches_tile, v_hartree_gradient, first_order_rho,
                v2rho2_batches, v2rhosigma_batches, v2sigma2_batches, partition_tab, first_order_H_tile,
                as_stream(stream));
  });
}

int frpa_far_plan_set_near(frpa_far_plan *plan, int n_species, const int *species_center, const int *n_grid,
                           const double *r_grid_min, const double *log_r_grid_inc) {
  return guarded([&] {
    FRPA_REQUIRE(plan, FRPA_ERR_INVALID, "null plan");
    set_near_tables(*reinterpret_cast<FarPlan *>(plan), n_species, species_center, n_grid, r_grid_min,
                    log_r_grid_inc);
  });
}

int frpa_near_forward(frpa_far_plan *plan, int atom_tile_size, int n_tile, int n_full_points, int n_valid_points,
                      int n_hartree_grid, const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                      const double *multipole_radius_sq, const double *outer_potential_radius,
                      const double *current_delta_v_hart_part_spl_tile, double *delta_v_hartree, void *stream) {
  return guarded([&] {
    FRPA_REQUIRE(plan && delta_v_hartree, FRPA_ERR_INVALID, "null argument");
    FRPA_REQUIRE(atom_tile_size >= 1 && n_tile >= 0 && n_tile <= atom_tile_size && n_full_points >= 0 &&
                     n_valid_points >= 0 && n_hartree_grid >= 0,
                 FRPA_ERR_INVALID, "near: bad sizes (T %d, n_tile %d, n_full %d, n_valid %d, n_hartree_grid %d)",
                 atom_tile_size, n_tile, n_full_points, n_valid_points, n_hartree_grid);
    FRPA_REQUIRE(n_valid_points == 0 || (coord_points && i_valid_point_2_i_full_points_map && multipole_radius_sq &&
                                         outer_potential_radius && current_delta_v_hart_part_spl_tile),
                 FRPA_ERR_INVALID, "near: null array");
    FarCall c;
    c.T = atom_tile_size;
    c.n_tile = n_tile;
    c.n_full = n_full_points;
    c.n_valid = n_valid_points;
    c.coord_points = coord_points;
    c.map = i_valid_point_2_i_full_points_map;
    c.mrsq = multipole_radius_sq;
    c.opr = outer_potential_radius;
    c.spl = current_delta_v_hart_part_spl_tile;
    c.n_hgrid = n_hartree_grid;
    near_forward(*reinterpret_cast<FarPlan *>(plan), c, delta_v_hartree, as_stream(stream));
  });
}

} // extern "C"
