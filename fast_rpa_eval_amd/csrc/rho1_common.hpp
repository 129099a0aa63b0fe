// rho1 plan: batch metadata resident in HBM, shared by the forward and adjoint kernels.
#pragma once

#include "common.hpp"

namespace frpa {

// One valid batch (n_compute_c > 0), in processing order.
struct BatchDesc {
  long long wave_off;  // offset of the batch's (nc_pad, np) block in wave_batches_compress
  long long ib_off;    // offset of its basis list in Rho1Plan::d_ibasis0
  int nc;              // n_compute_c
  int np;              // n_point_batches
  int pstart;          // n_point_batches_prefix_sum[b]
  int b;               // batch index in the caller's numbering (bias / gradient layouts)
};
static_assert(sizeof(BatchDesc) == 32, "BatchDesc layout");

// Order 3 of the INT8 forward: a work item is the row tiles [I0, I1) of one point pass q of a batch;
// the items of a batch are claimed back to back on one XCD, so they share its phi digits and P'
// tiles in that XCD's L2.  Each row tile leaves its per-point sums (C components: rho1, and the
// three gradient components for the GGA launch) in a partial buffer at poff + (I C + c) npc + p, and
// a finishing kernel adds them up per pass in row-tile order.
struct TileItem {
  int di;          // BatchDesc index
  int q;           // point pass (points q pch .. + npc - 1)
  int I0, I1;      // row tiles
  long long poff;  // partial-buffer offset of the pass (doubles, per perturbation)
};
struct PassRec {
  int di, q, nT, npc;
  long long poff;
};
struct TileSched {
  int pch = 0;                  // points per pass the lists were built for (0 = not built)
  int n_items = 0, n_recs = 0;
  long long part_doubles = 0;   // partial-buffer doubles per perturbation
  DevBuf<TileItem> d_items;     // in launch-position order (XCD-interleaved)
  DevBuf<PassRec> d_recs;
};

// Basis relabeling: new label pi[mu] = order of first appearance of mu over the batches in the
// caller's order (a batch's functions in ascending index), so that the compute sets of a spatially
// ordered batch list become long runs of consecutive labels.  Per batch, its compute rows sorted by
// new label: srow[k] = local row of the caller's list, sib[k] = pi of that row's basis function
// (ascending), both compact with the same offsets as Rho1Plan::d_ibasis0.
struct Relabel {
  std::vector<int> h_pi, h_pinv, h_srow, h_sib, h_spos;  // spos: inverse of srow per batch
  DevBuf<int> d_pi, d_pinv, d_srow, d_sib, d_spos;
};

// How a rho1 forward kernel walks the batches (chosen by frpa_rho1_autotune, or the default):
//   order 0: descending work (the BatchDesc order); order 1: XCD-aware spatial order -- the batches in
//   the caller's order cut into 8 contiguous ranges, launch position 8 k + x taking the k-th batch of
//   range x, so that workgroup x (mod 8), which the dispatcher places on XCD x, works through one
//   spatial region and its L2 sees neighbouring batches;
//   persistent: a grid of a few workgroups per CU that claim positions 8 k + x from per-XCD
//   counters instead of one workgroup per batch;
//   npg: point groups per wave and pass (0 = the default for the precision).
struct LaunchCfg {
  int order = 0;
  int persistent = 0;
  int npg = 0;
};
constexpr int N_XCD = 8;

struct OzakiState;  // INT8 split data (rho1_ozaki.hip)
struct AdjState;    // basis-tile work lists of the adjoint (rho1_adjoint.hip)
struct AdjOzState;  // digits of the INT8-split adjoint (rho1_adjoint_ozaki.hip)
struct LowpState;   // converted operands of the FP32 / BF16x3 forward (rho1_lowp.hip)

struct Rho1Plan {
  int n_batches = 0, n_basis = 0, n_full = 0, nmb = 0;
  int n_valid = 0, max_np = 0, max_nc = 0;
  long long wave_len = 0;
  std::vector<BatchDesc> h_desc;    // processing order (descending work)
  std::vector<int> h_ibasis0;       // 0-based basis indices, compact
  DevBuf<BatchDesc> d_desc;
  DevBuf<int> d_ibasis0;
  DevBuf<int> d_map;
  OzakiState *ozaki = nullptr;
  AdjState *adj = nullptr;
  AdjOzState *adjoz = nullptr;
  LowpState *lowp = nullptr;
  Relabel *relabel = nullptr;       // built on first use (get_relabel)
  DevBuf<int> d_order_spatial;      // launch position -> BatchDesc index, order 1
  // order 2 (INT8 forward): the spatial order with every point pass of PCH points a work item of
  // its own, (BatchDesc index << 8) | pass; the passes of a batch are adjacent in one XCD's range
  DevBuf<int> d_order_items;
  int n_items = 0, items_pch = 0;
  TileSched tiles[2];               // order 3 (INT8 forward): [0] rho1 only, [1] with the gradient
  DevBuf<double> d_tile_part;       // order 3 partial sums (grow-only)
  DevBuf<int> d_counters;           // N_XCD work counters of the persistent mode
  int n_cu = 0;                     // compute units of the device the plan was made on
  bool wave_off_quads = false;      // every BatchDesc::wave_off a multiple of 4 doubles
  // bias part 2 (rho1_bias2.hip): DM + DM^T on the rows of the tile's atoms, rebuilt per call
  DevBuf<int> d_b2_rowmap;          // n_basis: local row of a tile basis function, -1 otherwise
  DevBuf<int> d_b2_rows;            // 1 + n_basis: [0] = number of rows R, then their basis indices
  DevBuf<double> d_b2_tab;          // (b2_cap, n_basis): column c holds the R rows at c
  int b2_cap = 0;
  LaunchCfg cfg[6];                 // per precision (FP64, INT8 split S = 5, S = 4, FP32, BF16x3, FP64 symmetric)
  int adj_spin_limit = 0;           // polls per slot wait of the INT8 adjoint (0 = its default)
  ~Rho1Plan();
};

struct Rho1Call {
  int T = 1, n_tile = 1;
  const double *wave = nullptr, *gwave = nullptr, *P1 = nullptr, *bias = nullptr, *gbias = nullptr;
  double *rho = nullptr, *grad = nullptr;
};

Rho1Plan *make_rho1_plan(int n_batches, int n_basis, int n_full_points, int n_max_batch_size,
                         const int *n_point_batches, const int *prefix, const int *map, const int *ncc,
                         const long long *ibasis_ptr, const int *ibasis, const long long *wave_offset);
void rho1_forward_fp64(Rho1Plan &P, const Rho1Call &c, hipStream_t s);
void rho1_forward_ozaki(Rho1Plan &P, const Rho1Call &c, int slices, hipStream_t s);
void rho1_forward_lowp(Rho1Plan &P, const Rho1Call &c, int precision, hipStream_t s);
void rho1_adjoint(Rho1Plan &P, int T, int precision, const double *wave, const double *g, double *H, hipStream_t s);
const Relabel &get_relabel(Rho1Plan &P);
// Kernel arguments of a launch configuration: the position -> BatchDesc map (nullptr = identity),
// the zeroed work counters (nullptr = one workgroup per position), and the grid.
struct LaunchArgs {
  const int *order;
  int *counters;
  dim3 grid;
};
// n_pos: launch positions per perturbation (-1: one per valid batch; order 2 passes n_items)
LaunchArgs launch_args(Rho1Plan &P, const LaunchCfg &cfg, int T, int blocks_per_cu, hipStream_t s, int n_pos = -1);
// order 2 position map for point passes of pch points (built on first use, rebuilt if pch changes)
void build_item_order(Rho1Plan &P, int pch);
// order 3 row-tile work items for point passes of pch points (gradient launch: every tile pair)
TileSched &build_tile_items(Rho1Plan &P, int pch, bool grad);
void fill_i32(int *p, size_t n, int v, hipStream_t s);

// Workgroup barrier that first waits until this wave's LDS operations have completed.
// __syncthreads() alone lets the compiler drop that wait on paths where it takes LDS operations as
// ordered across waves; on gfx950 a wave that reached such a barrier with LDS writes still
// outstanding (its partial sums of a block reduction) has had them read stale by another wave right
// after the barrier.  Used wherever one wave reads what other waves wrote to LDS.
__device__ __forceinline__ void block_sync() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt at their maxima (no wait)
  __syncthreads();
}

// Persistent-mode work claiming: item k of XCD group x is launch position 8 (k % len) + x of
// perturbation t = k / len, where len = number of positions congruent to x.
__device__ __forceinline__ bool claim_item(int *counters, int n_pos, int T, int &pos, int &t, int *s_item) {
  const int x = blockIdx.x % N_XCD;
  const int len = n_pos / N_XCD + (x < n_pos % N_XCD ? 1 : 0);
  block_sync();  // every thread has read the previous claim
  if (threadIdx.x == 0)
    *s_item = atomicAdd(&counters[x], 1);
  block_sync();
  const int k = *s_item;
  if (len == 0 || k >= len * T)
    return false;
  pos = N_XCD * (k % len) + x;
  t = k / len;
  return true;
}
void ozaki_release(OzakiState *st);
void adj_release(AdjState *st);
void adjoz_release(AdjOzState *st);
void lowp_release(LowpState *st);
// FP32 / BF16x3 forward: (re)converts the wave's phi for that precision; returns the bytes held
long long lowp_prepare_wave(Rho1Plan &P, const double *wave, int precision, hipStream_t s);
// INT8-split adjoint of one call: int64 fixed-point atomics into the zeroed relabelled upper-triangle
// scratch Hs; returns the device array of the per-perturbation bounds B that fix the scales
const double *h1_ozaki_push(Rho1Plan &P, int T, int S, const double *wave, const double *g, unsigned long long *Hs,
                            hipStream_t s);
// INT8-split adjoint digits: (re)build for a wave, forget them (next call rebuilds), whether built
void adjoz_prepare(Rho1Plan &P, int S, const double *wave, hipStream_t s);
void adjoz_invalidate(Rho1Plan &P);
bool adjoz_built(const Rho1Plan &P);
long long adjoz_bytes(const Rho1Plan &P);  // device bytes of the adjoint's digits, scales and point maxima
// Slot waits of the INT8 adjoint that gave up since the last read (read and cleared; synchronises s)
int adjoz_take_wait_failures(Rho1Plan &P, hipStream_t s);

// Scale 2^E of the adjoint's fixed-point accumulation for the bound B >= |every partial sum|:
// 2 B 2^E <= 2^61 (1 when B = 0)
__host__ __device__ inline double fixed_scale(double B) {
  if (!(B > 0.0))
    return 1.0;
  int e = 0;
  (void)frexp(2.0 * B, &e);  // 2 B < 2^e
  return ldexp(1.0, 61 - e);
}
long long ozaki_prepare_wave(Rho1Plan &P, const double *wave, int slices, hipStream_t s);
void synth_wave(const Rho1Plan &P, int n_atoms, const double *atom_xyz, const int *basis_atom, const int *basis_l,
                const int *basis_m, const double *basis_alpha, int n_basis, const double *coord_points, double *wave,
                hipStream_t s);

// first_order_rho_bias_part2 (nmb, n_batches, T) for the atoms j_atom_begin .. + n_tile - 1 (rho1_bias2.hip);
// avn = atom_valid_n_compute_c_batches (n_atoms + 1, n_batches) on the device.  sides = 1 with
// accumulate: the v6 inline bias (DM one-sided), added to bias.
void rho1_bias_part2(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                     const double *wave, const double *gwave, const double *DM, double *bias, hipStream_t s,
                     int sides = 2, bool accumulate = false);

// Bias part 2 of the first-order H (evaluate_first_order_H_bias_part2.cpp:13-218) for the atoms
// j_atom_begin .. + n_tile - 1, added into H (n_basis, n_basis, T) (rho1_h1_bias2.hip).
void rho1_h1_bias_part2(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                        const double *wave, const double *gwave, const double *hpsi, const double *vrho,
                        const double *vsigma, const double *gradient_rho, const double *partition_tab, double *H,
                        hipStream_t s);

// Full first-order H (LDA + GGA terms of evaluate_first_order_H.cpp:13-222) for the perturbations
// t < n_tile (atoms j_atom_begin + t), accumulated into H (n_basis, n_basis, T) (rho1_adjoint.hip).
void rho1_h1_gga(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                 const double *wave, const double *gwave, const double *hwave, const double *vsigma,
                 const double *gradient_rho, const double *pref1, const double *grho1, const double *vhg,
                 const double *rho1, const double *v2rho2, const double *v2rhosigma, const double *v2sigma2,
                 const double *partition_tab, double *H, hipStream_t s);

} // namespace frpa
