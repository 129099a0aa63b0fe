// Adjoint of rho1 with respect to P1 = the LDA first-order H integration:
//   first_order_H(mu, nu, t) += sum_batches sum_{p in batch} g(map(p), t) phi_mu(p) phi_nu(p)
// for every pair (mu, nu) of the batch's compute set.
//
// Reference: evaluate_first_order_H.cpp:144-215 builds per batch contract = 0.5 g phi, a dsyr2k
// H_loc = contract phi^T + phi contract^T, and scatter-adds H_loc into the dense n_basis^2 matrix.
// A per-batch GPU scatter would issue sum_b nc_b^2 atomic adds (3e10 for 60k batches of ~650
// functions), so the work is reorganised by output ownership instead: H is cut into 64 x 64 basis
// tiles; each workgroup owns one tile pair (I <= J) and loops over exactly the batches whose
// compute sets touch both tiles, accumulating
//   H_IJ += sum_b Phi_b[rows in I] diag(g) Phi_b[rows in J]^T
// in FP64 MFMA registers (K = the batch's points), then writes the tile and, for I < J, its
// transpose.  Every H element is written once per call, in a fixed batch order: deterministic,
// no atomics.
// That tile-owner kernel is the deterministic mode (FRPA_ADJOINT=tiles).  Its tiles are only as
// full as the compute sets are contiguous in the basis order, and compact 3-D compute sets never are
// in any 1-D order (25-40% fill measured on the synthetic clusters), so the default mode instead
// computes each batch's local H densely (64 x 64 tiles of its own rows, I <= J) and adds it with
// hardware FP64 atomics (global_atomic_add_f64): every MFMA flop is useful; the summation order
// across batches, and so the last bits, can vary from run to run.  The atomics are the cost of that
// mode, so they are made few and line-friendly: the rows are taken in the relabelled order
// (get_relabel), only the upper triangle x <= y of the relabelled H is accumulated, into a zeroed
// scratch matrix Hs, and the MFMA operands are swapped so that the 16 lanes of an atomic run along
// 16 consecutive rows (one 128-byte line when the labels are contiguous); a final pass adds
// Hs(min(x, y), max(x, y)) into H(pinv x, pinv y) for every element.
// Tiles of both modes are cut in the relabelled order pi (first appearance over the caller's batch
// order), so that functions that share batches share tiles even when the caller's basis numbering is
// not spatially ordered; results are written back at the caller's indices.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rho1_common.hpp"

namespace frpa {

typedef double v4d __attribute__((ext_vector_type(4)));

struct AdjEntry {
  int d;       // index into the plan's BatchDesc array
  int i0, i1;  // positions in the batch's pi-sorted row list that fall in tile I
  int j0, j1;  // ... in tile J
};
struct AdjTask {
  int I, J;
  int e0, e1;  // entries [e0, e1)
};

struct AdjState {
  bool tiles_built = false;  // work lists of the tile-owner mode
  int n_tasks = 0;
  DevBuf<AdjEntry> d_entries;
  DevBuf<AdjTask> d_tasks;
  DevBuf<double> d_hs;       // push mode: upper triangle of the relabelled H, (n_basis, n_basis, T)
  DevBuf<double> d_gga_cp;   // full first-order H: per point, C then phi of the batch's sorted rows
  // full first-order H work items: (BatchDesc index, first row tile, end row tile), the items of a
  // batch adjacent on one XCD (launch position 8 k + x takes item k of range x)
  int n_gga_items = 0;
  DevBuf<int4> d_gga_items;
};

void adj_release(AdjState *st) { delete st; }

namespace {

constexpr int TILE = 64;
constexpr int KCH = 16;

__global__ __launch_bounds__(256) void h1_tile_kernel(const AdjTask *__restrict__ tasks,
                                                      const AdjEntry *__restrict__ entries,
                                                      const BatchDesc *__restrict__ descs,
                                                      const int *__restrict__ ibasis0, const int *__restrict__ srow,
                                                      const int *__restrict__ pi, const int *__restrict__ pinv,
                                                      const int *__restrict__ map,
                                                      const double *__restrict__ wave, const double *__restrict__ g,
                                                      int n_full, int n_basis, double *__restrict__ H) {
  constexpr int LDT = TILE + 1;
  __shared__ double As[KCH * LDT];
  __shared__ double Bs[KCH * LDT];
  __shared__ int invI[TILE], invJ[TILE];
  __shared__ double gk[KCH];

  const AdjTask task = tasks[blockIdx.x];
  const int t = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wm = 32 * (w >> 1), wn = 32 * (w & 1);  // this wave's 32 x 32 sub-tile
  const double *gt = g + (size_t)n_full * t;

  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
      acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int e = task.e0; e < task.e1; e++) {
    const AdjEntry en = entries[e];
    const BatchDesc d = descs[en.d];
    const int ld = (d.nc + 3) & ~3;
    const double *W = wave + d.wave_off;
    const int *ib = ibasis0 + d.ib_off;
    const int *sr = srow + d.ib_off;
    __syncthreads();
    if (tid < TILE) {
      invI[tid] = -1;
      invJ[tid] = -1;
    }
    __syncthreads();
    for (int k = en.i0 + tid; k < en.i1; k += 256) {
      const int i = sr[k];
      invI[pi[ib[i]] - TILE * task.I] = i;
    }
    for (int k = en.j0 + tid; k < en.j1; k += 256) {
      const int i = sr[k];
      invJ[pi[ib[i]] - TILE * task.J] = i;
    }
    for (int p0 = 0; p0 < d.np; p0 += KCH) {
      __syncthreads();
      if (tid < KCH) {
        const int p = p0 + tid;
        gk[tid] = p < d.np ? gt[map[d.pstart + p]] : 0.0;
      }
      __syncthreads();
      for (int q = tid; q < TILE * KCH; q += 256) {
        const int m = q % TILE, k = q / TILE;
        const int p = p0 + k;
        const int ia = invI[m], ja = invJ[m];
        As[k * LDT + m] = (ia >= 0 && p < d.np) ? W[ia + (size_t)ld * p] * gk[k] : 0.0;
        Bs[k * LDT + m] = (ja >= 0 && p < d.np) ? W[ja + (size_t)ld * p] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < KCH; kk += 4) {
        const double a0 = As[(kk + lk) * LDT + wm + lr];
        const double a1 = As[(kk + lk) * LDT + wm + 16 + lr];
        const double b0 = Bs[(kk + lk) * LDT + wn + lr];
        const double b1 = Bs[(kk + lk) * LDT + wn + 16 + lr];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
    }
  }

  // v_mfma_f64_16x16x4f64 leaves D(4 r + lane / 16, lane % 16) in acc[r]:
  // acc[a][b][r] = H_tile(wm + 16 a + 4 r + lk, wn + 16 b + lr)
  double *Ht = H + (size_t)n_basis * n_basis * t;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int pm = TILE * task.I + wm + 16 * a + 4 * r + lk;
        const int pn = TILE * task.J + wn + 16 * b + lr;
        if (pm < n_basis && pn < n_basis) {
          const int mu = pinv[pm], nu = pinv[pn];
          const double v = acc[a][b][r];
          Ht[mu + (size_t)n_basis * nu] += v;
          if (task.I != task.J)
            Ht[nu + (size_t)n_basis * mu] += v;
        }
      }
}

// Default mode: a work item is a range of row tiles of one batch (its local tile pairs I <= J of the
// sorted compute rows, I in the range), the items of a batch adjacent on one XCD; atomics into the
// upper triangle of the relabelled scratch Hs.  The operands of the next (pair, point chunk) step
// are loaded into registers while the current step's MFMAs run.
__global__ __launch_bounds__(256) void h1_push_kernel(const BatchDesc *__restrict__ descs,
                                                      const int4 *__restrict__ items,
                                                      const int *__restrict__ sib_all, const int *__restrict__ srow_all,
                                                      const int *__restrict__ map, const double *__restrict__ wave,
                                                      const double *__restrict__ g, int n_full, int n_basis,
                                                      double *__restrict__ Hs) {
  constexpr int LDT = TILE + 1;
  __shared__ double As[KCH * LDT];  // rows of tile I times g
  __shared__ double Bs[KCH * LDT];  // rows of tile J
  const int4 item = items[blockIdx.x];  // row tiles [item.y, item.z) of batch item.x
  const BatchDesc d = descs[item.x];
  const int t = blockIdx.y;
  const int nc = d.nc, np = d.np, ld = (nc + 3) & ~3;
  const double *W = wave + d.wave_off;
  const int *sib = sib_all + d.ib_off;
  const int *srow = srow_all + d.ib_off;
  const double *gt = g + (size_t)n_full * t;
  double *Ht = Hs + (size_t)n_basis * n_basis * t;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wm = 32 * (w >> 1), wn = 32 * (w & 1);
  const int nT = (nc + TILE - 1) / TILE;
  const int nck = (np + KCH - 1) / KCH;
  const int m = tid % TILE;  // this thread's staged row within a tile
  constexpr int QN = TILE * KCH / 256;
  double nx[QN][2];
  auto load_step = [&](int I, int J, int c) {
    const int i_st = TILE * I + m, j_st = TILE * J + m;
    const int ri = i_st < nc ? srow[i_st] : -1, rj = j_st < nc ? srow[j_st] : -1;
#pragma unroll
    for (int u = 0; u < QN; u++) {
      const int k = (tid + 256 * u) / TILE, p = KCH * c + k;
      double a = 0.0, b = 0.0;
      if (p < np) {
        if (ri >= 0)
          a = W[ri + (size_t)ld * p] * gt[map[d.pstart + p]];
        if (rj >= 0)
          b = W[rj + (size_t)ld * p];
      }
      nx[u][0] = a;
      nx[u][1] = b;
    }
  };
  auto advance = [&](int &I, int &J, int &c) {
    if (++c == nck) {
      c = 0;
      if (++J == nT) {
        ++I;
        J = I;
      }
    }
  };
  v4d acc[2][2];
  const int I1 = min(item.z, nT);
  int I = item.y, J = item.y, c = 0;
  if (I < I1)
    load_step(I, J, c);
  while (I < I1) {
    __syncthreads();  // the previous step's MFMAs are done with the staged operands
#pragma unroll
    for (int u = 0; u < QN; u++) {
      const int k = (tid + 256 * u) / TILE;
      As[k * LDT + m] = nx[u][0];
      Bs[k * LDT + m] = nx[u][1];
    }
    __syncthreads();
    int nI = I, nJ = J, ncx = c;
    advance(nI, nJ, ncx);
    if (nI < I1)
      load_step(nI, nJ, ncx);  // in flight during the MFMAs below
    if (c == 0) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
          acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int kk = 0; kk < KCH; kk += 4) {
      const double a0 = As[(kk + lk) * LDT + wm + lr];
      const double a1 = As[(kk + lk) * LDT + wm + 16 + lr];
      const double b0 = Bs[(kk + lk) * LDT + wn + lr];
      const double b1 = Bs[(kk + lk) * LDT + wn + 16 + lr];
      // D = (J rows) x (I rows): the I row runs along lane % 16 in the result
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
    }
    if (c == nck - 1) {
      // acc[a][b][r] = Hloc(i = 64 I + wm + 16 a + lr, j = 64 J + wn + 16 b + 4 r + lk)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int i = TILE * I + wm + 16 * a + lr;
        const size_t x = i < nc ? (size_t)sib[i] : 0;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int j = TILE * J + wn + 16 * b + 4 * r + lk;
            if (i < nc && j < nc && i <= j)
              atomicAdd(&Ht[x + (size_t)n_basis * sib[j]], acc[a][b][r]);
          }
      }
    }
    I = nI;
    J = nJ;
    c = ncx;
  }
}

// H(pinv x, pinv y) += Hs(min(x, y), max(x, y)) for every element; 64 x 64 tiles of the relabelled
// indices, the transposed tile read through LDS so that both reads run down columns.  With bound
// non-null Hs holds int64 fixed point with the scale fixed_scale(bound[t]) (the INT8-split adjoint).
__device__ __forceinline__ double hs_value(const double *Hst, size_t e, double inv) {
  return inv > 0.0 ? (double)reinterpret_cast<const long long *>(Hst)[e] * inv : Hst[e];
}

__global__ __launch_bounds__(256) void h1_unrelabel_kernel(const double *__restrict__ Hs, const int *__restrict__ pinv,
                                                           int n_basis, const double *__restrict__ bound,
                                                           double *__restrict__ H) {
  __shared__ double tr[64][65];
  const int x0 = 64 * blockIdx.x, y0 = 64 * blockIdx.y, t = blockIdx.z;
  const size_t nn = (size_t)n_basis * n_basis;
  const double *Hst = Hs + nn * t;
  const double inv = bound ? 1.0 / fixed_scale(bound[t]) : 0.0;  // 0: Hs holds FP64
  double *Ht = H + nn * t;
  const int a = threadIdx.x & 63, b0 = threadIdx.x >> 6;
  if (x0 > y0 || x0 == y0) {
    // tr[i][j] = Hs(y0 + j, x0 + i)
    const int y = y0 + a;
#pragma unroll 4
    for (int u = 0; u < 16; u++) {
      const int i = b0 + 4 * u, x = x0 + i;
      tr[i][a] = (x < n_basis && y < n_basis) ? hs_value(Hst, y + (size_t)n_basis * x, inv) : 0.0;
    }
  }
  __syncthreads();
  const int x = x0 + a;
  if (x >= n_basis)
    return;
  const size_t px = pinv[x];
#pragma unroll 4
  for (int u = 0; u < 16; u++) {
    const int j = b0 + 4 * u, y = y0 + j;
    if (y >= n_basis)
      break;
    const double v = x <= y ? hs_value(Hst, x + (size_t)n_basis * y, inv) : tr[a][j];
    Ht[px + (size_t)n_basis * pinv[y]] += v;
  }
}

void build_tiles(Rho1Plan &P, AdjState &st) {
  const int nb = P.n_basis;
  const int nt = (nb + TILE - 1) / TILE;
  const Relabel &R = get_relabel(P);
  const std::vector<int> &pi = R.h_pi;
  const std::vector<int> &srow = R.h_srow;
  std::vector<std::vector<AdjEntry>> per(size_t(nt) * nt);
  for (int di = 0; di < P.n_valid; di++) {
    const BatchDesc &d = P.h_desc[di];
    const int *ib = P.h_ibasis0.data() + d.ib_off;
    const int *sr = srow.data() + d.ib_off;
    // tile runs of the pi-sorted row list
    std::vector<int> tl, r0, r1;
    for (int k = 0; k < d.nc;) {
      const int tile = pi[ib[sr[k]]] / TILE;
      int j = k;
      while (j < d.nc && pi[ib[sr[j]]] / TILE == tile)
        j++;
      tl.push_back(tile);
      r0.push_back(k);
      r1.push_back(j);
      k = j;
    }
    for (size_t x = 0; x < tl.size(); x++)
      for (size_t y = x; y < tl.size(); y++)
        per[size_t(tl[x]) * nt + tl[y]].push_back(AdjEntry{di, r0[x], r1[x], r0[y], r1[y]});
  }
  std::vector<AdjEntry> entries;
  std::vector<AdjTask> tasks;
  std::vector<double> work;
  for (int I = 0; I < nt; I++)
    for (int J = I; J < nt; J++) {
      auto &v = per[size_t(I) * nt + J];
      if (v.empty())
        continue;
      double wk = 0;
      for (auto &e : v)
        wk += (double)P.h_desc[e.d].np;
      tasks.push_back(AdjTask{I, J, (int)entries.size(), (int)(entries.size() + v.size())});
      work.push_back(wk);
      entries.insert(entries.end(), v.begin(), v.end());
    }
  // largest work lists first
  std::vector<int> ord(tasks.size());
  for (size_t i = 0; i < ord.size(); i++)
    ord[i] = (int)i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work[a] > work[b]; });
  std::vector<AdjTask> sorted(tasks.size());
  for (size_t i = 0; i < ord.size(); i++)
    sorted[i] = tasks[ord[i]];
  st.n_tasks = (int)sorted.size();
  st.d_entries.upload(entries.empty() ? std::vector<AdjEntry>{AdjEntry{0, 0, 0, 0, 0}} : entries);
  st.d_tasks.upload(sorted.empty() ? std::vector<AdjTask>{AdjTask{0, 0, 0, 0}} : sorted);
  st.tiles_built = true;
}

// Full first-order H (LDA and GGA terms), evaluate_first_order_H.cpp:144-215: for batch b and
// perturbation t (atom j = j_atom_begin + t, its compute rows V = [avn(j-1, b), avn(j, b))), per point
//   sigma1 = 2 grad_rho1 . grad_rho,  a = (H_prefactor1 + pref3) / 2,
//   pref3 = w (v_hartree_gradient + v2rho2 rho1 + v2rhosigma sigma1),
//   pref2_c = 2 w grad_rho_c (v2rhosigma rho1 + v2sigma2 sigma1) + 2 w vsigma grad_rho1_c,
//   pref5_c = 2 vsigma w grad_rho_c,
//   C(i, p) = a phi(i, p) + sum_c pref2_c dphi_c(i, p) - [i in V] sum_c pref5_c d2phi_{c, j_coord}(i, p),
//   Hloc = C phi^T + phi C^T,
// with the push kernel's structure: tile pairs I <= J of the label-sorted rows, the points as the
// contraction index in chunks of KCH (prefactors per chunk in LDS, C built while staging), both
// products on FP64 MFMA into one accumulator, the upper triangle added with FP64 atomics into Hs.
// A work item is a range of row tiles of one batch (its pairs I <= J for I in the range), the items
// of a batch adjacent on one XCD; the operands of the next (pair, chunk) step are loaded into
// registers while the current step's MFMAs run.
struct GgaIn {
  const double *gwave, *hwave;                            // (ncp, 3, np) / (ncp, 6, np) per batch
  const int *avn;                                         // (n_atoms + 1, n_batches)
  const double *vsigma, *v2rho2, *v2rhosigma, *v2sigma2;  // (3 | 3 | 6 | 6, nmb) per batch
  const double *gradient_rho;                             // (3, n_full)
  const double *pref1;                                    // H_prefactor1 (nmb, n_batches, T)
  const double *grho1;                                    // first_order_gradient_rho (3, nmb, n_batches, T)
  const double *vhg, *rho1;                               // (n_full, T)
  const double *wpart;                                    // partition_tab (n_full)
  int n_atoms, j_atom_begin, j_coord, nmb, n_batches;
};

// Full first-order H, pass 1: per batch, the staged operands of both products in the batch's sorted
// row order, CP[2 wave_off + 2 ld p + k] = C(k, p) and CP[... + ld + k] = phi(k, p), k < nc, with
//   C(r, p) = a(p) phi(r, p) + sum_c pref2_c(p) d_c phi(r, p) - [r in V] sum_c pref5_c(p) d_c d_jc phi(r, p)
// and the per-point prefactors of evaluate_first_order_H.cpp:144-168 (computed once per point).
__global__ __launch_bounds__(256) void h1_gga_prep_kernel(const BatchDesc *__restrict__ descs,
                                                          const int *__restrict__ srow_all,
                                                          const int *__restrict__ map, const double *__restrict__ wave,
                                                          const GgaIn in, int t, int n_full, double *__restrict__ CP) {
  __shared__ double pk[256][8];  // a, pref2[3], pref5[3] of the pass's points
  const BatchDesc d = descs[blockIdx.x];
  const int nc = d.nc, np = d.np, ld = (nc + 3) & ~3;
  const double *W = wave + d.wave_off;
  const double *GW = in.gwave + 3 * d.wave_off;
  const double *HW = in.hwave + 6 * d.wave_off;
  const int *srow = srow_all + d.ib_off;
  double *out = CP + 2 * d.wave_off;
  const int *av = in.avn + (size_t)(in.n_atoms + 1) * d.b;
  const int v0 = min(max(av[in.j_atom_begin - 1 + t], 0), nc), v1 = min(max(av[in.j_atom_begin + t], v0), nc);
  // hessian component (c, j_coord) in the packed (xx, xy, xz, yy, yz, zz) order
  const int jc = in.j_coord - 1;
  const int h0 = jc, h1 = jc == 0 ? 1 : jc == 1 ? 3 : 4, h2 = jc == 0 ? 2 : jc == 1 ? 4 : 5;
  const int tid = threadIdx.x;
  const size_t bt = (size_t)d.b + (size_t)in.n_batches * t;
  for (int p0 = 0; p0 < np; p0 += 256) {
    const int npp = min(256, np - p0);
    __syncthreads();  // the previous pass is done with pk
    if (tid < npp) {
      const int p = p0 + tid;
      const int q = map[d.pstart + p];
      const double wq = in.wpart[q];
      const double r1 = in.rho1[(size_t)q + (size_t)n_full * t];
      const double *gr = in.gradient_rho + 3 * (size_t)q;
      const double *g1 = in.grho1 + 3 * ((size_t)in.nmb * bt + p);
      const double sig1 = 2.0 * (g1[0] * gr[0] + g1[1] * gr[1] + g1[2] * gr[2]);
      const size_t pb = (size_t)in.nmb * d.b + p;
      const double vs = in.vsigma[3 * pb], vrs = in.v2rhosigma[6 * pb], vss = in.v2sigma2[6 * pb];
      const double pref3 = wq * (in.vhg[(size_t)q + (size_t)n_full * t] + in.v2rho2[3 * pb] * r1 + vrs * sig1);
      pk[tid][0] = 0.5 * (in.pref1[(size_t)in.nmb * bt + p] + pref3);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        pk[tid][1 + c] = 2.0 * wq * gr[c] * (vrs * r1 + vss * sig1) + 2.0 * wq * vs * g1[c];
        pk[tid][4 + c] = 2.0 * vs * wq * gr[c];
      }
    }
    __syncthreads();
    for (int e = tid; e < nc * npp; e += 256) {
      const int k = e % nc, pl = e / nc, p = p0 + pl;
      const int r = srow[k];
      const double phi = W[r + (size_t)ld * p];
      const double *g = GW + r + (size_t)3 * ld * p;
      double c = pk[pl][0] * phi + pk[pl][1] * g[0] + pk[pl][2] * g[ld] + pk[pl][3] * g[2 * ld];
      if (r >= v0 && r < v1) {
        const double *h = HW + r + (size_t)6 * ld * p;
        c -= pk[pl][4] * h[(size_t)ld * h0] + pk[pl][5] * h[(size_t)ld * h1] + pk[pl][6] * h[(size_t)ld * h2];
      }
      double *col = out + (size_t)2 * ld * p;
      col[k] = c;
      col[ld + k] = phi;
    }
  }
}

// Full first-order H, pass 2: Hloc = C phi^T + phi C^T on the tile pairs I <= J of the batch's sorted
// rows, points as the contraction index in chunks of KCH (both products in one accumulator), the
// upper triangle added with FP64 atomics into the relabelled scratch of perturbation t.
__global__ __launch_bounds__(256) void h1_gga_kernel(const BatchDesc *__restrict__ descs,
                                                     const int4 *__restrict__ items,
                                                     const int *__restrict__ sib_all, const double *__restrict__ CP,
                                                     int n_basis, double *__restrict__ Ht) {
  constexpr int LDT = TILE + 1;
  __shared__ double Ci[KCH * LDT], Pi[KCH * LDT], Cj[KCH * LDT], Pj[KCH * LDT];
  const int4 item = items[blockIdx.x];  // rows tiles [item.y, item.z) of batch item.x
  const BatchDesc d = descs[item.x];
  const int nc = d.nc, np = d.np, ld = (nc + 3) & ~3;
  const double *CPb = CP + 2 * d.wave_off;
  const int *sib = sib_all + d.ib_off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wm = 32 * (w >> 1), wn = 32 * (w & 1);
  const int nT = (nc + TILE - 1) / TILE;
  const int nck = (np + KCH - 1) / KCH;
  const int m = tid % TILE;
  // steps (I, J, chunk) over the tile pairs I <= J and the point chunks; the operands of the next
  // step are loaded into registers while this step's MFMAs run
  constexpr int QN = TILE * KCH / 256;  // staged elements per thread and array
  double nx[QN][4];
  auto load_step = [&](int I, int J, int c) {
    const int i_st = TILE * I + m, j_st = TILE * J + m;
#pragma unroll
    for (int u = 0; u < QN; u++) {
      const int k = (tid + 256 * u) / TILE, p = KCH * c + k;
      double ci = 0.0, fi = 0.0, cj = 0.0, fj = 0.0;
      if (p < np) {
        const double *col = CPb + (size_t)2 * ld * p;
        if (i_st < nc) {
          ci = col[i_st];
          fi = col[ld + i_st];
        }
        if (j_st < nc) {
          cj = col[j_st];
          fj = col[ld + j_st];
        }
      }
      nx[u][0] = ci;
      nx[u][1] = fi;
      nx[u][2] = cj;
      nx[u][3] = fj;
    }
  };
  auto advance = [&](int &I, int &J, int &c) {
    if (++c == nck) {
      c = 0;
      if (++J == nT) {
        ++I;
        J = I;
      }
    }
  };
  v4d acc[2][2];
  const int I1 = min(item.z, nT);
  int I = item.y, J = item.y, c = 0;
  if (I < I1)
    load_step(I, J, c);
  while (I < I1) {
    __syncthreads();  // the previous step's MFMAs are done with the staged operands
#pragma unroll
    for (int u = 0; u < QN; u++) {
      const int k = (tid + 256 * u) / TILE;
      Ci[k * LDT + m] = nx[u][0];
      Pi[k * LDT + m] = nx[u][1];
      Cj[k * LDT + m] = nx[u][2];
      Pj[k * LDT + m] = nx[u][3];
    }
    __syncthreads();
    int nI = I, nJ = J, ncx = c;
    advance(nI, nJ, ncx);
    if (nI < I1)
      load_step(nI, nJ, ncx);  // in flight during the MFMAs below
    if (c == 0) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
          acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int kk = 0; kk < KCH; kk += 4) {
      const int o = (kk + lk) * LDT;
      const double ca0 = Ci[o + wm + lr], ca1 = Ci[o + wm + 16 + lr];
      const double pa0 = Pi[o + wm + lr], pa1 = Pi[o + wm + 16 + lr];
      const double cb0 = Cj[o + wn + lr], cb1 = Cj[o + wn + 16 + lr];
      const double pb0 = Pj[o + wn + lr], pb1 = Pj[o + wn + 16 + lr];
      // D = (J rows) x (I rows): C(i) phi(j) + phi(i) C(j)
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(pb0, ca0, acc[0][0], 0, 0, 0);
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cb0, pa0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(pb1, ca0, acc[0][1], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cb1, pa0, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(pb0, ca1, acc[1][0], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cb0, pa1, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(pb1, ca1, acc[1][1], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cb1, pa1, acc[1][1], 0, 0, 0);
    }
    if (c == nck - 1) {
      // acc[a][b][r] = Hloc(i = 64 I + wm + 16 a + lr, j = 64 J + wn + 16 b + 4 r + lk)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int i = TILE * I + wm + 16 * a + lr;
        const size_t x = i < nc ? (size_t)sib[i] : 0;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int j = TILE * J + wn + 16 * b + 4 * r + lk;
            if (i < nc && j < nc && i <= j)
              atomicAdd(&Ht[x + (size_t)n_basis * sib[j]], acc[a][b][r]);
          }
      }
    }
    I = nI;
    J = nJ;
    c = ncx;
  }
}

} // namespace

// Work items of the FP64 adjoint and the full first-order H: per batch (in the caller's order, so that neighbouring
// batches follow each other), ranges of row tiles of about `work` tile pairs (tile I has nT - I); the
// items of a batch are adjacent in one XCD's range, so they read its staged operands from that L2.
void build_gga_items(Rho1Plan &P, AdjState &st) {
  int work = 8;
  if (const char *e = std::getenv("FRPA_GGA_TILE_WORK"))
    work = std::max(1, std::atoi(e));
  std::vector<int> caller(P.n_valid);
  for (int i = 0; i < P.n_valid; i++)
    caller[i] = i;
  std::stable_sort(caller.begin(), caller.end(), [&](int a, int b) { return P.h_desc[a].b < P.h_desc[b].b; });
  std::vector<int4> seq;
  for (int di : caller) {
    const int nT = (P.h_desc[di].nc + TILE - 1) / TILE;
    int I0 = 0, w = 0;
    for (int I = 0; I < nT; I++) {
      w += nT - I;
      if (w >= work || I == nT - 1) {
        seq.push_back(make_int4(di, I0, I + 1, 0));
        I0 = I + 1;
        w = 0;
      }
    }
  }
  const int n = (int)seq.size();
  std::vector<int4> items(std::max(n, 1), make_int4(0, 0, 0, 0));
  int start = 0;
  for (int x = 0; x < N_XCD; x++) {
    const int len = n / N_XCD + (x < n % N_XCD ? 1 : 0);
    for (int k = 0; k < len; k++)
      items[N_XCD * k + x] = seq[start + k];
    start += len;
  }
  st.d_gga_items.upload(items);
  st.n_gga_items = n;
}

void rho1_adjoint(Rho1Plan &P, int T, int precision, const double *wave, const double *g, double *H, hipStream_t s) {
  FRPA_REQUIRE(wave || P.n_valid == 0, FRPA_ERR_INVALID, "null wave");
  if (P.n_valid == 0)
    return;
  FRPA_REQUIRE(precision == FRPA_PREC_FP64 || precision == FRPA_PREC_OZAKI || precision == FRPA_PREC_OZAKI4,
               FRPA_ERR_INVALID, "unknown precision %d", precision);
  if (!P.adj)
    P.adj = new AdjState();
  AdjState &st = *P.adj;
  const Relabel &R = get_relabel(P);
  const int n = P.n_basis;
  const char *mode = getenv("FRPA_ADJOINT");
  if (precision != FRPA_PREC_FP64 || !(mode && strcmp(mode, "tiles") == 0)) {
    const size_t nn = (size_t)n * n * T;
    st.d_hs.grow(nn);
    fill_f64(st.d_hs.ptr, nn, 0.0, s);  // +0.0 and int64 0 have the same bits
    const double *bound = nullptr;      // non-null: Hs holds int64 fixed point (INT8 split)
    if (precision == FRPA_PREC_FP64) {
      if (st.n_gga_items == 0)
        build_gga_items(P, st);
      hipLaunchKernelGGL(h1_push_kernel, dim3(st.n_gga_items, T), dim3(256), 0, s, P.d_desc.ptr, st.d_gga_items.ptr,
                         R.d_sib.ptr, R.d_srow.ptr, P.d_map.ptr, wave, g, P.n_full, n, st.d_hs.ptr);
      check_launch("h1_push_kernel");
    } else {
      bound = h1_ozaki_push(P, T, precision == FRPA_PREC_OZAKI ? 5 : 4, wave, g,
                            reinterpret_cast<unsigned long long *>(st.d_hs.ptr), s);
      // a slot wait that gave up (here or in a graph replay since the last check) means an
      // incomplete H: report it instead of adding it (not while the stream is being captured;
      // frpa_rho1_adjoint_status covers graph replays)
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      FRPA_HIP(hipStreamIsCapturing(s, &cap));
      if (cap == hipStreamCaptureStatusNone) {
        const int fails = adjoz_take_wait_failures(P, s);
        FRPA_REQUIRE(fails == 0, FRPA_ERR_HIP,
                     "INT8 adjoint: %d slot waits gave up (in this call or a graph replay since the last check); "
                     "first_order_H was not updated by this call",
                     fails);
      }
    }
    hipLaunchKernelGGL(h1_unrelabel_kernel, dim3((n + 63) / 64, (n + 63) / 64, T), dim3(256), 0, s, st.d_hs.ptr,
                       R.d_pinv.ptr, n, bound, H);
    check_launch("h1_unrelabel_kernel");
    return;
  }
  if (!st.tiles_built)
    build_tiles(P, st);
  if (st.n_tasks == 0)
    return;
  hipLaunchKernelGGL(h1_tile_kernel, dim3(st.n_tasks, T), dim3(256), 0, s, st.d_tasks.ptr, st.d_entries.ptr,
                     P.d_desc.ptr, P.d_ibasis0.ptr, R.d_srow.ptr, R.d_pi.ptr, R.d_pinv.ptr, P.d_map.ptr, wave, g,
                     P.n_full, n, H);
  check_launch("h1_tile_kernel");
}

void rho1_h1_gga(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                 const double *wave, const double *gwave, const double *hwave, const double *vsigma,
                 const double *gradient_rho, const double *pref1, const double *grho1, const double *vhg,
                 const double *rho1, const double *v2rho2, const double *v2rhosigma, const double *v2sigma2,
                 const double *partition_tab, double *H, hipStream_t s) {
  FRPA_REQUIRE(n_tile >= 1 && n_tile <= T, FRPA_ERR_INVALID, "H1: n_tile %d outside [1, T = %d]", n_tile, T);
  FRPA_REQUIRE(j_coord >= 1 && j_coord <= 3, FRPA_ERR_INVALID, "H1: j_coord %d outside [1, 3]", j_coord);
  FRPA_REQUIRE(j_atom_begin >= 1 && j_atom_begin + n_tile - 1 <= n_atoms, FRPA_ERR_INVALID,
               "H1: atoms %d..%d outside [1, n_atoms = %d]", j_atom_begin, j_atom_begin + n_tile - 1, n_atoms);
  if (P.n_valid == 0)
    return;
  FRPA_REQUIRE(wave && gwave && hwave && avn && vsigma && gradient_rho && pref1 && grho1 && vhg && rho1 && v2rho2 &&
                   v2rhosigma && v2sigma2 && partition_tab && H,
               FRPA_ERR_INVALID, "H1: null array");
  if (!P.adj)
    P.adj = new AdjState();
  AdjState &st = *P.adj;
  const Relabel &R = get_relabel(P);
  const int n = P.n_basis;
  const size_t nn = (size_t)n * n * n_tile;
  st.d_hs.grow(nn);
  fill_f64(st.d_hs.ptr, nn, 0.0, s);
  GgaIn in{gwave, hwave, avn, vsigma, v2rho2, v2rhosigma, v2sigma2, gradient_rho, pref1, grho1, vhg, rho1,
           partition_tab, n_atoms, j_atom_begin, j_coord, P.nmb, P.n_batches};
  st.d_gga_cp.grow((size_t)(2 * P.wave_len));
  if (st.n_gga_items == 0)
    build_gga_items(P, st);
  for (int t = 0; t < n_tile; t++) {
    hipLaunchKernelGGL(h1_gga_prep_kernel, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, R.d_srow.ptr, P.d_map.ptr,
                       wave, in, t, P.n_full, st.d_gga_cp.ptr);
    check_launch("h1_gga_prep_kernel");
    hipLaunchKernelGGL(h1_gga_kernel, dim3(st.n_gga_items), dim3(256), 0, s, P.d_desc.ptr, st.d_gga_items.ptr,
                       R.d_sib.ptr, st.d_gga_cp.ptr, n, st.d_hs.ptr + (size_t)n * n * t);
    check_launch("h1_gga_kernel");
  }
  hipLaunchKernelGGL(h1_unrelabel_kernel, dim3((n + 63) / 64, (n + 63) / 64, n_tile), dim3(256), 0, s, st.d_hs.ptr,
                     R.d_pinv.ptr, n, nullptr, H);
  check_launch("h1_unrelabel_kernel");
}

} // namespace frpa
