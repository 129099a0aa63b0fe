// Adjoint of rho1 (the LDA first-order H integration) with the per-batch product emulated on INT8
// MFMA, the counterpart of the INT8-split forward:
//   H(i, k) += sum_p g(map(p)) phi_i(p) phi_k(p)   for the batch's compute rows i, k.
// Split (radix-254 signed digits, round to nearest, as in the forward) with one scale per ROW, the
// point index being the contraction index:
//   Y(i, p) = g_p phi_i(p):  Y sy_i = sum_s y_s 254^-s,   sy_i = 127 / max_p |Y(i, p)|   (per call
//                            and pass of at most 320 points, made while staging tile I)
//   phi(k, p):               phi sf_k = sum_u f_u 254^-u,  sf_k = 127 / max_p |phi(k, p)| (per wave)
//   H(i, k) = sum_l 254^-l C_l(i, k) / (sy_i sf_k),  C_l = sum_{s+u=l} sum_p y_s(i, p) f_u(k, p)
// with the levels s + u <= S - 1 (exact int32: |C_l| <= 5 np 127^2).
// Reference: evaluate_first_order_H.cpp:144-215 (contract = 0.5 w v1 phi, dsyr2k, scatter-add).
//
// Work items are groups of consecutive batches (the caller's order, contiguous points, at most one
// 320-point pass, union U of the compute sets with |U|^2 <= 0.8 sum nc^2; build_groups): over U the
// group is one batch whose rows are zero at the points of members lacking them, so its product adds
// the sum of the members' products with one atomic per H element instead of one per member -- the
// atomics, which bound this kernel, drop ~1.2x on a spatially ordered batch list for ~4% more
// multiplications.
// Layout: per group, S digit planes of phi, rows in relabelled order (get_relabel), |U| rounded to 64
// rows (nc64) and the points rounded to 64; a plane is a sequence of 64-point chunks of nc64 rows x
// 64 bytes (row r, point p at (p / 64) nc64 64 + 64 r + p % 64), so that a lane's MFMA fragment (16
// consecutive points of one row) is one 16-byte load and a wave's fragments (16 rows x 64 points)
// one contiguous kilobyte.  The kernel computes the Y digits
// of one 64-row tile I for all points of the pass from the phi digits and g into LDS, and
// streams the phi digits of the tiles J >= I from global memory (two steps of register prefetch, no
// barrier inside the J loop); the upper triangle of each tile pair is added into the relabelled
// scratch H (see rho1_adjoint.hip) as 64-bit fixed point with integer atomics: value * 2^E rounded to
// an int64, 2^E per call and perturbation from B = sum over points |g(p)| max_i phi_i(p)^2 (summed
// in a fixed order), which bounds |H(i, k)| and every partial sum of it, so that 2 B 2^E <= 2^61 (no
// overflow; resolution 2^-E <= 4 B / 2^62).  Integer adds are exact and commute, so H does not
// depend on the order in which the atomics land: the INT8-split adjoint is bit-reproducible (the
// FP64 atomics of the other adjoints are not), at the speed of the FP64 atomics.  Non-finite g or
// phi leave H undefined.
#include <algorithm>
#include <cstdlib>
#include <exception>
#include <iterator>
#include <numeric>
#include <thread>
#include <vector>

#include "rho1_common.hpp"

namespace frpa {

typedef int v4i __attribute__((ext_vector_type(4)));

// One batch of a group: where its wave block is and how its rows map onto the group's rows.
struct GMember {
  long long wave_off;    // its (ld, np) block in wave_batches_compress
  long long rowmap_off;  // its row map: group row -> its row in the caller's list, or -1
  int ld;                // leading dimension of its block, n_compute_c rounded up to 4
  int np;                // its points
  int p0;                // its first point within the group
  int pad;
};
static_assert(sizeof(GMember) == 32, "GMember layout");

struct AzItem {
  int g;       // group (its index in group order), or -1: nothing to do
  int i0, i1;  // its row tiles I in [i0, i1)
  int pad;
};
static_assert(sizeof(AzItem) == 16, "AzItem layout");

struct AdjOzState {
  int S = 0;
  const double *wave = nullptr;  // the wave the phi digits were made from
  double merge = -1.0;           // merge threshold the groups were made with
  long long bytes = 0;           // bytes of the phi digits
  // Groups of consecutive batches (in the caller's order) merged into one work item, in launch
  // order: nc = rows of the union of their compute sets, np = their points (contiguous in the
  // point map), pstart = that of the first, ib_off = offset of the group's rows in d_gsib / d_finv.
  int n_groups = 0;
  DevBuf<BatchDesc> d_gdesc;
  DevBuf<int> d_gsib;            // relabelled basis index of every group row, ascending per group
  DevBuf<int> d_gmem;            // n_groups + 1: group g's members are [d_gmem[g], d_gmem[g + 1])
  DevBuf<GMember> d_members;
  DevBuf<int> d_mdi;             // per member: its BatchDesc index (into Rho1Plan::d_desc)
  DevBuf<int> d_rowmap;          // per member, per group row: its row in the caller's list, or -1
  DevBuf<long long> d_off;       // per group: byte offset of its digit planes
  DevBuf<int2> d_dtiles;         // (group, 64-row tile) work items of h1_digits_kernel
  DevBuf<signed char> d_fdig;    // phi digits (per wave)
  DevBuf<double> d_finv;         // 1 / sf per group row (compact, at ib_off)
  DevBuf<double> d_pmax;         // max_i |phi_i(p)| per valid point (pstart + p)
  // Work items of the kernel: (group, row tiles [i0, i1)), in grid order; a group's items are
  // consecutive in one XCD's share of the grid (see xcd_items), g < 0 pads the grid.
  int n_items = 0;
  DevBuf<AzItem> d_items;
  DevBuf<double> d_bound;        // per call: B per perturbation
  DevBuf<double> d_bpart;        // per call: B's partial sums, one per workgroup of h1_bound_kernel
  DevBuf<int> d_spin_fail;       // waits of the producer / consumer mode that gave up (0 when all is well)
  int *h_gsib = nullptr;         // pinned staging of d_gsib (hipHostMalloc), kept for the next set-up
  size_t h_gsib_bytes = 0;
  AdjOzState() = default;
  AdjOzState(const AdjOzState &) = delete;
  AdjOzState &operator=(const AdjOzState &) = delete;
  ~AdjOzState() {
    if (h_gsib)
      (void)hipHostFree(h_gsib);
  }
};

void adjoz_release(AdjOzState *st) { delete st; }

namespace {

constexpr double RADIX = 254.0;
constexpr int AZ_THREADS = 512;
constexpr int PMAX = 320;  // points held in LDS per pass (larger batches take several passes)
constexpr int NH = (PMAX / 16 + 7) / 8;  // 16-point groups per wave when a tile's Y digits are staged
constexpr int GMAX = 16;   // batches per group

__device__ __forceinline__ int ceil64(int x) { return (x + 63) & ~63; }

// Row maps of the groups (once per wave): for member m of group g, rowmap[rowmap_off(m) + r] = the
// caller's local row of group row r in m's block, or -1 where m lacks that basis function.  One
// workgroup per group: the map is cleared, then each member row finds its group row by binary search
// of its label in the group's ascending labels (every member label is in the union).
__global__ __launch_bounds__(256) void h1_rowmap_kernel(const BatchDesc *__restrict__ gdesc,
                                                        const int *__restrict__ gmem,
                                                        const GMember *__restrict__ members,
                                                        const int *__restrict__ mdi,
                                                        const BatchDesc *__restrict__ descs,
                                                        const int *__restrict__ sib_all,
                                                        const int *__restrict__ srow_all,
                                                        const int *__restrict__ gsib, int *__restrict__ rowmap) {
  const BatchDesc gd = gdesc[blockIdx.x];
  const int *U = gsib + gd.ib_off;
  const int nU = gd.nc;
  const int m0 = gmem[blockIdx.x], m1 = gmem[blockIdx.x + 1];
  for (int m = m0; m < m1; m++) {
    int *rm = rowmap + members[m].rowmap_off;
    for (int r = threadIdx.x; r < nU; r += blockDim.x)
      rm[r] = -1;
  }
  __syncthreads();  // cleared before any member row is placed
  for (int m = m0; m < m1; m++) {
    const BatchDesc d = descs[mdi[m]];
    const int *lab = sib_all + d.ib_off, *srow = srow_all + d.ib_off;
    int *rm = rowmap + members[m].rowmap_off;
    for (int q = threadIdx.x; q < d.nc; q += blockDim.x) {
      const int x = lab[q];
      int lo = 0, hi = nU;  // first r with U[r] >= x
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (U[mid] < x)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < nU)
        rm[lo] = srow[q];
    }
  }
}

// Phi digit planes of the groups (once per wave): rows in the group's order (ascending relabelled
// index), one scale per row over all the group's points; a member's points where the row is not in
// its compute set are zero.  One workgroup per (group, 64-row tile): lane = row of the tile, so every
// wave load reads consecutive rows of one point (a member's rows of the tile are mostly consecutive
// there); the 4 waves take every fourth point for the row maxima, then a 16-point piece each of
// every 64-point chunk for the digits (the tile's wave values again), staged in LDS so that each
// plane's 4 KB of the chunk is written contiguously.  Rows nc .. nc64-1 and points np .. npk-1 are
// zero.
template <int S>
__global__ __launch_bounds__(256) void h1_digits_kernel(const int2 *__restrict__ tiles,
                                                        const BatchDesc *__restrict__ gdesc,
                                                        const long long *__restrict__ off,
                                                        const int *__restrict__ gmem,
                                                        const GMember *__restrict__ members,
                                                        const int *__restrict__ rowmap,
                                                        const double *__restrict__ wave,
                                                        signed char *__restrict__ dig, double *__restrict__ inv) {
  const int2 tile = tiles[blockIdx.x];
  const BatchDesc d = gdesc[tile.x];
  const int nc = d.nc, np = d.np;
  const int nc64 = ceil64(nc), npk = ceil64(np);
  const size_t plane = (size_t)nc64 * npk;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = 64 * tile.y + lane;
  __shared__ GMember sm[GMAX];
  __shared__ int lrs[GMAX][64];  // member m's row of tile row lane, or -1
  __shared__ double red[4][64];
  const int m0 = gmem[tile.x], nm = gmem[tile.x + 1] - m0;
  if (tid < nm)
    sm[tid] = members[m0 + tid];
  __syncthreads();
  for (int e = tid; e < 64 * nm; e += 256) {
    const int m = e >> 6, rr = 64 * tile.y + (e & 63);
    lrs[m][e & 63] = rr < nc ? rowmap[sm[m].rowmap_off + rr] : -1;
  }
  __syncthreads();
  double mx = 0.0;
  for (int m = 0; m < nm; m++) {
    const int lr = lrs[m][lane];
    if (lr < 0)
      continue;
    const double *W = wave + sm[m].wave_off + lr;
    const size_t ld = (size_t)sm[m].ld;
    const int mp = sm[m].np;
    int p = w;
    for (; p + 60 < mp; p += 64) {  // 16 loads in flight per lane
      double x[16];
#pragma unroll
      for (int j = 0; j < 16; j++)
        x[j] = W[ld * (p + 4 * j)];
#pragma unroll
      for (int j = 0; j < 16; j++)
        mx = fmax(mx, fabs(x[j]));
    }
    for (; p < mp; p += 4)
      mx = fmax(mx, fabs(W[ld * p]));
  }
  red[w][lane] = mx;
  __syncthreads();
  const double rmax = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
  const double sc = rmax > 0.0 ? 127.0 / rmax : 1.0;
  if (w == 0 && r < nc)
    inv[d.ib_off + r] = 1.0 / sc;
  // digits of one 64-point chunk at a time: wave w takes the chunk's points 16 w .. + 15 of every row
  // (its 16 loads in flight together), leaves them in LDS as the chunk's layout, and the block writes
  // each plane's 4 KB of the tile contiguously
  __shared__ __attribute__((aligned(16))) signed char stage[S][64 * 64];
  signed char *obase = dig + off[tile.x] + (size_t)64 * 64 * tile.y;
  // member m holds group points [pend - its np, pend); Wr = its column of this row (nullptr: absent)
  int m = 0, pend = nm ? sm[0].np : 0;
  const double *Wr = nullptr;
  if (r < nc && nm) {
    const int lr = lrs[0][lane];
    Wr = lr >= 0 ? wave + sm[0].wave_off + lr : nullptr;
  }
  for (int cq = 0; cq < npk / 64; cq++) {
    const int p0 = 64 * cq + 16 * w;
    double x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int p = p0 + j;
      x[j] = 0.0;
      if (r < nc && p < np) {
        while (p >= pend) {  // p < np = the members' points, so m stays < nm
          m++;
          pend += sm[m].np;
          const int lr = lrs[m][lane];
          Wr = lr >= 0 ? wave + sm[m].wave_off + lr : nullptr;
        }
        if (Wr)
          x[j] = Wr[(size_t)sm[m].ld * (p - (pend - sm[m].np))];
      }
    }
    unsigned int w4[S][4];
#pragma unroll
    for (int sl = 0; sl < S; sl++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        w4[sl][u] = 0u;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      double v = x[j] * sc;
#pragma unroll
      for (int sl = 0; sl < S; sl++) {
        const double dg = rint(v);
        w4[sl][j >> 2] |= (unsigned int)(unsigned char)(signed char)dg << (8 * (j & 3));
        v = (v - dg) * RADIX;
      }
    }
#pragma unroll
    for (int sl = 0; sl < S; sl++)
      *reinterpret_cast<v4i *>(&stage[sl][64 * lane + 16 * w]) =
          v4i{(int)w4[sl][0], (int)w4[sl][1], (int)w4[sl][2], (int)w4[sl][3]};
    __syncthreads();
    for (int e = tid; e < S * 256; e += 256) {
      const int sl = e >> 8, q = e & 255;
      *reinterpret_cast<v4i *>(obase + sl * plane + (size_t)cq * nc64 * 64 + 16 * q) =
          *reinterpret_cast<const v4i *>(&stage[sl][16 * q]);
    }
    __syncthreads();
  }
}

// max_i |phi_i(p)| over the batch's rows for every point (once per wave): one wave per point, lanes
// down the point's column.
__global__ __launch_bounds__(256) void h1_pointmax_kernel(const BatchDesc *__restrict__ descs,
                                                          const double *__restrict__ wave, double *__restrict__ pmax) {
  const BatchDesc d = descs[blockIdx.x];
  const int ld = (d.nc + 3) & ~3;
  const double *W = wave + d.wave_off;
  const int lane = threadIdx.x & 63;
  for (int p = threadIdx.x >> 6; p < d.np; p += 4) {
    const double *Wp = W + (size_t)ld * p;
    double mx = 0.0;
    int r = lane;
    for (; r + 64 * 7 < d.nc; r += 64 * 8) {  // 8 loads in flight per lane
      double x[8];
#pragma unroll
      for (int j = 0; j < 8; j++)
        x[j] = Wp[r + 64 * j];
#pragma unroll
      for (int j = 0; j < 8; j++)
        mx = fmax(mx, fabs(x[j]));
    }
    for (; r < d.nc; r += 64)
      mx = fmax(mx, fabs(Wp[r]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      mx = fmax(mx, __shfl_xor(mx, o));
    if (lane == 0)
      pmax[d.pstart + p] = mx;
  }
}

// part[t][block] = sum over the block's valid points q of |g(map(q), t)| pmax(q)^2
__global__ __launch_bounds__(256) void h1_bound_kernel(const int *__restrict__ map, const double *__restrict__ g,
                                                       int n_full, const double *__restrict__ pmax, int n_points,
                                                       double *__restrict__ part) {
  __shared__ double red[4];
  const int t = blockIdx.y;
  double v = 0.0;
  for (int q = blockIdx.x * 256 + threadIdx.x; q < n_points; q += gridDim.x * 256) {
    const double m = pmax[q];
    v += fabs(g[(size_t)map[q] + (size_t)n_full * t]) * m * m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(size_t)gridDim.x * t + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// bound[t] = sum of the n_part partial sums of perturbation t, in a fixed order (one workgroup per t)
__global__ __launch_bounds__(256) void h1_bound_sum_kernel(const double *__restrict__ part, int n_part,
                                                           double *__restrict__ bound) {
  __shared__ double red[4];
  const int t = blockIdx.x;
  double v = 0.0;
  for (int i = threadIdx.x; i < n_part; i += 256)
    v += part[(size_t)n_part * t + i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    bound[t] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Block: 8 waves; wave w multiplies phi rows 16 (w & 3) .. +15 of a tile J (MFMA rows) against the
// 64 Y columns of tile I, the two halves w >> 2 taking alternate J, so D(k, i) lands with i along
// lane % 16 (consecutive relabelled rows of H for the atomics).
//
// PC (producer / consumer, the default): waves 0 .. PC_CW - 1 multiply -- the units (J, 16-row group)
// of a tile I dealt round-robin -- and leave each 16 x 64 tile-pair block, in fixed point, in an LDS
// slot of their own; the other PC_NA waves issue every atomic from the slots.  A wave's
// vector-memory counter is in order, so a multiplying wave that issued atomics would wait for their
// memory-side completion before using its next phi-digit loads; with the atomics in other waves the
// two overlap.  Slot handover: release / acquire flags in LDS, every wait bounded by spin_limit polls
// (a producer waits for its slot to empty, an atomic wave for the next block; the count restarts at
// every block taken).  A wave that gives up adds 1 to *spin_fail and goes on -- a producer drops its
// block rather than overwrite a full slot -- so the kernel always ends and H is known to be incomplete
// (the host reports it).  Every wave stages the Y digits; an atomic wave drains all blocks of tile I
// before the barrier that starts tile I + 1, and the slot flags are cleared after that barrier.
constexpr int PC_CW = 6;         // multiplying waves of the producer / consumer mode
constexpr int PC_NA = 8 - PC_CW; // atomic waves; atomic wave aw serves the multiplying waves aw + PC_NA j
constexpr int PC_PER = (PC_CW + PC_NA - 1) / PC_NA;
constexpr int PC_SPIN = 1 << 16; // default bound on the polls of one wait (each ~64 clocks: ~2 ms)

template <int S, bool PC>
__global__ __launch_bounds__(AZ_THREADS) void h1_ozaki_kernel(const AzItem *__restrict__ items,
                                                              const BatchDesc *__restrict__ descs,
                                                              const long long *__restrict__ off,
                                                              const int *__restrict__ sib_all,
                                                              const int *__restrict__ map,
                                                              const double *__restrict__ g, int n_full,
                                                              const signed char *__restrict__ fdig,
                                                              const double *__restrict__ finv, int n_basis,
                                                              const double *__restrict__ bound,
                                                              unsigned long long *__restrict__ Hs,
                                                              int *__restrict__ spin_fail, int spin_limit) {
  constexpr int YROW = PMAX + 16;  // LDS bytes per (plane, row) of the staged Y tile
  __shared__ __attribute__((aligned(16))) signed char Ys[S * 64 * YROW];
  __shared__ double gs[PMAX];      // g(map(p), t) of the pass's points
  __shared__ double ysc_s[64];     // 1 / sy of the tile's rows
  // the slots (q(k row kk, i row ii) at kk * 64 + ii) and the row maxima of the Y staging share LDS:
  // the staging ends with a barrier before the first block is multiplied, and the next tile's staging
  // starts after the barrier that follows the last block drained
  constexpr int SLOT_WORDS = PC ? PC_CW * 16 * 64 : 1;
  __shared__ __attribute__((aligned(16))) long long slot_mem[SLOT_WORDS > 8 * 64 ? SLOT_WORDS : 8 * 64];
  long long(*slot)[16 * 64] = reinterpret_cast<long long(*)[16 * 64]>(slot_mem);
  double(*red)[64] = reinterpret_cast<double(*)[64]>(slot_mem);
  __shared__ int slot_full[PC_CW], slot_cols[PC_CW][16];      // cols: H column labels of the block's 16 rows
  const AzItem item = items[blockIdx.x];
  if (item.g < 0)
    return;  // grid padding: the whole block leaves before any barrier
  const int di = item.g, t = blockIdx.y;
  const BatchDesc d = descs[di];
  const int nc = d.nc, np = d.np;
  const int nc64 = ceil64(nc), npk = ceil64(np);
  const size_t plane = (size_t)nc64 * npk;
  const signed char *F = fdig + off[di];
  const double *fi = finv + d.ib_off;
  const int *sib = sib_all + d.ib_off;
  const double *gt = g + (size_t)n_full * t;
  unsigned long long *Ht = Hs + (size_t)n_basis * n_basis * t;
  const double scale = fixed_scale(bound[t]);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int kw = w & 3, iw = w >> 2;
  const int nT = nc64 / 64;

  for (int pp = 0; pp < npk; pp += PMAX) {
    const int npp = min(PMAX, npk - pp);  // points of this pass (multiple of 64)
    const int nck = npp / 64;
    block_sync();
    for (int j = tid; j < npp; j += AZ_THREADS)
      gs[j] = pp + j < np ? gt[map[d.pstart + pp + j]] : 0.0;
    for (int I = item.i0; I < item.i1; I++) {
      // Y = g phi for the tile's 64 sorted rows and the pass's points, as S digit planes in LDS with
      // one scale per row.  phi comes from its own digits (x = phi sf, 16 points of one row per
      // 16-byte load per plane, read once for the row maximum and once more for the digits): z = g x
      // gives the row maximum m; the Y digits are those of 127 z / m and 1 / sy = m / (127 sf).  Row a =
      // tid % 64; the 8 waves take the 16-point groups w, w + 8, w + 16 of the pass.
      const int a = tid & 63;
      const int ia = 64 * I + a;
      block_sync();  // gs written; the previous tile's Y, scales and slots are no longer read
      if (PC && tid < PC_CW)
        slot_full[tid] = 0;  // read again only after the barrier that ends the staging
      v4i fr[S];
      auto load_fr = [&](int c16) {
#pragma unroll
        for (int sl = 0; sl < S; sl++)
          fr[sl] = *reinterpret_cast<const v4i *>(F + sl * plane + (size_t)((pp >> 6) + (c16 >> 2)) * nc64 * 64 +
                                                  (size_t)64 * ia + 16 * (c16 & 3));
      };
      auto zval = [&](int c16, int j) {
        double x = (double)(signed char)((unsigned)fr[S - 1][j >> 2] >> (8 * (j & 3)));
#pragma unroll
        for (int sl = S - 2; sl >= 0; sl--)
          x = x * (1.0 / RADIX) + (double)(signed char)((unsigned)fr[sl][j >> 2] >> (8 * (j & 3)));
        return x * gs[16 * c16 + j];
      };
      double mx = 0.0;
      for (int h = 0; h < NH; h++) {
        const int c16 = w + 8 * h;
        if (c16 < npp / 16) {
          load_fr(c16);
#pragma unroll
          for (int j = 0; j < 16; j++)
            mx = fmax(mx, fabs(zval(c16, j)));
        }
      }
      red[w][a] = mx;
      block_sync();
      if (w == 0) {
        double m = red[0][a];
#pragma unroll
        for (int q = 1; q < 8; q++)
          m = fmax(m, red[q][a]);
        red[0][a] = m > 0.0 ? 127.0 / m : 1.0;
        ysc_s[a] = ia < nc && m > 0.0 ? m / 127.0 * fi[ia] : 0.0;
      }
      block_sync();
      const double sc = red[0][a];
      for (int h = 0; h < NH; h++) {
        const int c16 = w + 8 * h;
        if (c16 < npp / 16) {
          load_fr(c16);  // again (L2): holding them across the barrier would spill
          unsigned int w4[S][4];
#pragma unroll
          for (int sl = 0; sl < S; sl++)
#pragma unroll
            for (int u = 0; u < 4; u++)
              w4[sl][u] = 0u;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            double x = zval(c16, j) * sc;
#pragma unroll
            for (int sl = 0; sl < S; sl++) {
              const double dg = rint(x);
              w4[sl][j >> 2] |= (unsigned int)(unsigned char)(signed char)dg << (8 * (j & 3));
              x = (x - dg) * RADIX;
            }
          }
#pragma unroll
          for (int sl = 0; sl < S; sl++)
            *reinterpret_cast<v4i *>(&Ys[(sl * 64 + a) * YROW + 16 * c16]) =
                v4i{(int)w4[sl][0], (int)w4[sl][1], (int)w4[sl][2], (int)w4[sl][3]};
        }
      }
      block_sync();
      double ysc[4];
#pragma unroll
      for (int gi = 0; gi < 4; gi++)
        ysc[gi] = ysc_s[16 * gi + lr];
      if constexpr (PC) {
        // units (J, kw) = (I + u / 4, u % 4), u < 4 (nT - I): multiplying wave w takes u = w, w + 6, ...
        const int NU = 4 * (nT - I);
        auto units_of = [&](int cw) { return NU > cw ? (NU - cw + PC_CW - 1) / PC_CW : 0; };
        if (w < PC_CW) {
          auto unit_jk = [&](int st, int &J, int &kwu) {
            const int u = w + PC_CW * (st / nck);
            J = I + u / 4;
            kwu = u % 4;
          };
          auto load_a = [&](int st, v4i (&a)[S]) {
            int J, kwu;
            unit_jk(st, J, kwu);
            const int c = st % nck;
            const signed char *src =
                F + (size_t)((pp >> 6) + c) * nc64 * 64 + (size_t)64 * (64 * J + 16 * kwu + lr) + 16 * lk;
#pragma unroll
            for (int sl = 0; sl < S; sl++)
              a[sl] = *reinterpret_cast<const v4i *>(src + sl * plane);
          };
          const int nsteps = units_of(w) * nck;
          v4i a0[S], a1[S], a2[S];
          if (nsteps > 0)
            load_a(0, a0);
          if (nsteps > 1)
            load_a(1, a1);
          v4i acc[S][4];
          for (int st = 0, c = 0; st < nsteps; st++) {
            if (c == 0) {
#pragma unroll
              for (int l = 0; l < S; l++)
#pragma unroll
                for (int gi = 0; gi < 4; gi++)
                  acc[l][gi] = v4i{0, 0, 0, 0};
            }
            if (st + 2 < nsteps)
              load_a(st + 2, a2);
#pragma unroll
            for (int gi = 0; gi < 4; gi++) {
              v4i b[S];
#pragma unroll
              for (int sl = 0; sl < S; sl++)
                b[sl] =
                    *reinterpret_cast<const v4i *>(&Ys[(sl * 64 + 16 * gi + lr) * YROW + 64 * c + 16 * lk]);
#pragma unroll
              for (int sa = 0; sa < S; sa++)
#pragma unroll
                for (int sb = 0; sb < S - sa; sb++)
                  acc[sa + sb][gi] =
                      __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[sa], b[sb], acc[sa + sb][gi], 0, 0, 0);
            }
            if (c == nck - 1) {
              int J, kwu;
              unit_jk(st, J, kwu);
              // wait until this wave's slot is empty, fill it, hand it over; a wait that gives up
              // drops the block (counted) instead of overwriting the one still in the slot
              int spins = 0;
              bool free_slot = true;
              while (__hip_atomic_load(&slot_full[w], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0) {
                if (++spins >= spin_limit) {
                  free_slot = false;
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
              }
              if (!free_slot) {
                if (lane == 0)
                  atomicAdd(spin_fail, 1);
              } else {
                // acc[l][gi][r] = C_l(k = 64 J + 16 kwu + 4 lk + r, i = 64 I + 16 gi + lr)
#pragma unroll
                for (int gi = 0; gi < 4; gi++)
#pragma unroll
                  for (int r = 0; r < 4; r++) {
                    const int k = 64 * J + 16 * kwu + 4 * lk + r;
                    double v = (double)acc[S - 1][gi][r];
#pragma unroll
                    for (int l = S - 2; l >= 0; l--)
                      v = v * (1.0 / RADIX) + (double)acc[l][gi][r];
                    slot[w][(4 * lk + r) * 64 + 16 * gi + lr] =
                        k < nc ? (long long)rint(v * ysc[gi] * fi[k] * scale) : 0ll;
                  }
                if (lane < 16) {
                  const int k = 64 * J + 16 * kwu + lane;
                  slot_cols[w][lane] = k < nc ? sib[k] : -1;
                }
                if (lane == 0)
                  __hip_atomic_store(&slot_full[w], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
              }
            }
#pragma unroll
            for (int sl = 0; sl < S; sl++) {
              a0[sl] = a1[sl];
              a1[sl] = a2[sl];
            }
            if (++c == nck)
              c = 0;
          }
        } else {
          // atomic wave aw serves multiplying waves aw + PC_NA e: exactly units_of(cw) blocks each
          const int aw = w - PC_CW;
          int left[PC_PER], total = 0;
#pragma unroll
          for (int e = 0; e < PC_PER; e++) {
            const int cw = aw + PC_NA * e;
            left[e] = cw < PC_CW ? units_of(cw) : 0;
            total += left[e];
          }
          const int i = 64 * I + lane;
          const int xi = i < nc ? sib[i] : 0;
          const size_t x = (size_t)xi;
          int spins = 0;  // polls since the last block taken
          while (total > 0) {
            bool got = false;
#pragma unroll
            for (int e = 0; e < PC_PER; e++) {
              const int cw = aw + PC_NA * e;
              if (left[e] > 0 &&
                  __hip_atomic_load(&slot_full[cw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0) {
                // the block's column labels come with it (no global load here, which would wait for
                // this wave's earlier atomics); sorted labels ascend, so i <= k is xi <= col
                long long qv[16];
                int col[16];
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                  qv[kk] = slot[cw][kk * 64 + lane];
                  col[kk] = slot_cols[cw][kk];
                }
                if (lane == 0)
                  __hip_atomic_store(&slot_full[cw], 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                for (int kk = 0; kk < 16; kk++)
                  if (i < nc && col[kk] >= 0 && xi <= col[kk])
                    atomicAdd(&Ht[x + (size_t)n_basis * col[kk]], (unsigned long long)qv[kk]);
                left[e]--;
                total--;
                got = true;
              }
            }
            if (got) {
              spins = 0;
            } else {
              if (++spins >= spin_limit)
                break;
              __builtin_amdgcn_s_sleep(1);
            }
          }
          if (total > 0 && lane == 0)
            atomicAdd(spin_fail, 1);
        }
      } else {
        // The two halves of the block (iw) take alternate column tiles J = I + iw, I + iw + 2, ...; a
        // wave multiplies phi rows 16 kw .. + 15 of its J against all 64 Y rows of tile I.
        // A fragments of step s = (J, c) = (I + iw + 2 (s / nck), s % nck): phi digits of row
        // 64 J + 16 kw + lr, points pp + 64 c + 16 lk .. + 15; step s + 2 is in flight while s runs.
        auto load_a = [&](int st, v4i (&a)[S]) {
          const int J = I + iw + 2 * (st / nck), c = st % nck;
          const signed char *src =
              F + (size_t)((pp >> 6) + c) * nc64 * 64 + (size_t)64 * (64 * J + 16 * kw + lr) + 16 * lk;
  #pragma unroll
          for (int sl = 0; sl < S; sl++)
            a[sl] = *reinterpret_cast<const v4i *>(src + sl * plane);
        };
        const int npairs = nT - I > iw ? (nT - I - iw + 1) / 2 : 0;
        const int nsteps = npairs * nck;
        v4i a0[S], a1[S], a2[S];
        if (nsteps > 0)
          load_a(0, a0);
        if (nsteps > 1)
          load_a(1, a1);
        v4i acc[S][4];
        for (int st = 0, J = I + iw, c = 0; st < nsteps; st++) {
          if (c == 0) {
  #pragma unroll
            for (int l = 0; l < S; l++)
  #pragma unroll
              for (int gi = 0; gi < 4; gi++)
                acc[l][gi] = v4i{0, 0, 0, 0};
          }
          if (st + 2 < nsteps)
            load_a(st + 2, a2);
  #pragma unroll
          for (int gi = 0; gi < 4; gi++) {
            v4i b[S];
  #pragma unroll
            for (int sl = 0; sl < S; sl++)
              b[sl] = *reinterpret_cast<const v4i *>(&Ys[(sl * 64 + 16 * gi + lr) * YROW + 64 * c + 16 * lk]);
  #pragma unroll
            for (int sa = 0; sa < S; sa++)
  #pragma unroll
              for (int sb = 0; sb < S - sa; sb++)
                acc[sa + sb][gi] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[sa], b[sb], acc[sa + sb][gi], 0, 0, 0);
          }
          if (c == nck - 1) {
            // acc[l][gi][r] = C_l(k = 64 J + 16 kw + 4 lk + r, i = 64 I + 16 gi + lr)
  #pragma unroll
            for (int gi = 0; gi < 4; gi++) {
              const int i = 64 * I + 16 * gi + lr;
              const size_t x = i < nc ? (size_t)sib[i] : 0;
  #pragma unroll
              for (int r = 0; r < 4; r++) {
                const int k = 64 * J + 16 * kw + 4 * lk + r;
                if (i < nc && k < nc && i <= k) {
                  double v = (double)acc[S - 1][gi][r];
  #pragma unroll
                  for (int l = S - 2; l >= 0; l--)
                    v = v * (1.0 / RADIX) + (double)acc[l][gi][r];
                  const long long q = (long long)rint(v * ysc[gi] * fi[k] * scale);
                  atomicAdd(&Ht[x + (size_t)n_basis * sib[k]], (unsigned long long)q);
                }
              }
            }
          }
  #pragma unroll
          for (int sl = 0; sl < S; sl++) {
            a0[sl] = a1[sl];
            a1[sl] = a2[sl];
          }
          if (++c == nck) {
            c = 0;
            J += 2;
          }
        }
      }
    }
  }
}

// Merge threshold of the groups: consecutive batches of the caller's order whose points are
// contiguous join a group while its points fit one pass (PMAX) and the union U of the compute sets
// keeps |U|^2 <= merge * sum of the members' n_compute^2 (the atomics of a group scale with |U|^2, its
// multiplications with |U|^2 times its points).  FRPA_ADJOINT_MERGE overrides it; 0 keeps every
// batch alone.
double merge_threshold() {
  const char *e = std::getenv("FRPA_ADJOINT_MERGE");
  return e && e[0] ? std::atof(e) : 0.8;
}

// Runs job(i) for i in [0, n) on up to `threads` host threads, the calling one included (job i on
// thread i mod the thread count; a thread that cannot be started has its jobs run here); an exception
// from a job is rethrown here once every thread has joined.
template <class Job>
void host_parallel(int n, int threads, const Job &job) {
  const int nt = std::max(1, std::min(n, threads));
  std::vector<std::exception_ptr> err(nt);
  auto body = [&](int t) {
    try {
      for (int i = t; i < n; i += nt)
        job(i);
    } catch (...) {
      err[t] = std::current_exception();
    }
  };
  std::vector<std::thread> pool;
  int started = 1;
  try {
    for (; started < nt; started++)
      pool.emplace_back(body, started);
  } catch (...) {
  }
  for (int t = started; t < nt; t++)
    body(t);
  body(0);
  for (std::thread &th : pool)
    th.join();
  for (const std::exception_ptr &e : err)
    if (e)
      std::rethrow_exception(e);
}

int host_threads() { return (int)std::max(1u, std::thread::hardware_concurrency()); }

// The groups of a contiguous piece of the caller's batch order (BatchDesc indices order[0 .. n)):
// consecutive batches whose points are contiguous join a group while its points fit one pass (PMAX),
// it has fewer than GMAX members and the union U of the compute sets keeps |U|^2 <= merge * sum nc^2.
struct GroupList {
  std::vector<std::vector<int>> mem;   // members (BatchDesc indices) per group
  std::vector<std::vector<int>> rows;  // union of the members' relabelled indices per group, ascending
};

void group_range(const Rho1Plan &P, const Relabel &R, const int *order, int n, double merge, GroupList &out) {
  std::vector<int> U;
  long long sq = 0;
  int np_end = 0, np_sum = 0;
  for (int k = 0; k < n; k++) {
    const int idx = order[k];
    const BatchDesc &d = P.h_desc[idx];
    const int *lab = R.h_sib.data() + d.ib_off;
    if (!out.mem.empty() && merge > 0.0 && (int)out.mem.back().size() < GMAX && d.pstart == np_end &&
        np_sum + d.np <= PMAX) {
      U.clear();
      std::set_union(out.rows.back().begin(), out.rows.back().end(), lab, lab + d.nc, std::back_inserter(U));
      const long long sq2 = sq + (long long)d.nc * d.nc;
      if ((double)U.size() * (double)U.size() <= merge * (double)sq2) {
        out.mem.back().push_back(idx);
        out.rows.back().swap(U);
        sq = sq2;
        np_end = d.pstart + d.np;
        np_sum += d.np;
        continue;
      }
    }
    out.mem.push_back({idx});
    out.rows.emplace_back(lab, lab + d.nc);
    sq = (long long)d.nc * d.nc;
    np_end = d.pstart + d.np;
    np_sum = d.np;
  }
}

// The caller's order is cut into GROUP_PIECES contiguous pieces of equal batch count, grouped on up
// to GROUP_PIECES host threads (a group never spans two pieces; the cut does not depend on the
// thread count, so the groups, and H, are the same on every machine).
constexpr int GROUP_PIECES = 16;

// Builds the groups into st (their row maps on the device, stream s); returns their descriptors
// (group order).
std::vector<BatchDesc> build_groups(Rho1Plan &P, AdjOzState &st, double merge, hipStream_t s) {
  const Relabel &R = get_relabel(P);
  std::vector<int> order(P.n_valid);
  for (int i = 0; i < P.n_valid; i++)
    order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return P.h_desc[x].b < P.h_desc[y].b; });
  const int npieces = std::max(1, std::min(GROUP_PIECES, P.n_valid));
  std::vector<GroupList> pieces(npieces);
  auto piece = [&](int c) {
    const int a = (int)((long long)P.n_valid * c / npieces), b = (int)((long long)P.n_valid * (c + 1) / npieces);
    group_range(P, R, order.data() + a, b - a, merge, pieces[c]);
  };
  host_parallel(npieces, host_threads(), piece);
  std::vector<std::vector<int>> mem, rows;
  for (GroupList &g : pieces) {
    std::move(g.mem.begin(), g.mem.end(), std::back_inserter(mem));
    std::move(g.rows.begin(), g.rows.end(), std::back_inserter(rows));
  }
  const int ng = (int)mem.size();
  // group order (their index and the layout of their digits): descending work |U|^2 points, ties by
  // the first member's position; the launch order is that of the work items (xcd_items)
  auto work = [&](int g) {
    long long np = 0;
    for (int i : mem[g])
      np += P.h_desc[i].np;
    const long long n64 = ((long long)rows[g].size() + 63) / 64 * 64;
    return n64 * n64 * ((np + 63) / 64 * 64);
  };
  std::vector<long long> wk(ng);
  std::vector<int> gorder(ng);
  for (int g = 0; g < ng; g++) {
    wk[g] = work(g);
    gorder[g] = g;
  }
  std::stable_sort(gorder.begin(), gorder.end(), [&](int x, int y) { return wk[x] > wk[y]; });
  std::vector<BatchDesc> gdesc(ng);
  std::vector<int> gmem(ng + 1, 0), mdi;
  std::vector<GMember> members;
  long long rm_off = 0;  // row-map entries so far (one per member and group row)
  long long n_gsib = 0;  // group rows so far
  for (int k = 0; k < ng; k++) {
    const int g = gorder[k];
    BatchDesc gd{};
    gd.ib_off = n_gsib;
    gd.nc = (int)rows[g].size();
    gd.pstart = P.h_desc[mem[g][0]].pstart;
    gd.b = P.h_desc[mem[g][0]].b;
    int p0 = 0;
    for (int i : mem[g]) {
      const BatchDesc &d = P.h_desc[i];
      GMember M{};
      M.wave_off = d.wave_off;
      M.rowmap_off = rm_off;
      rm_off += gd.nc;
      M.ld = (d.nc + 3) & ~3;
      M.np = d.np;
      M.p0 = p0;
      p0 += d.np;
      members.push_back(M);
      mdi.push_back(i);
    }
    gd.np = p0;
    gdesc[k] = gd;
    n_gsib += gd.nc;
    gmem[k + 1] = (int)members.size();
  }
  st.n_groups = ng;
  st.d_gdesc.upload(gdesc);
  // the groups' rows, in group order, through a pinned host buffer kept by the plan (filled on the
  // host threads; one DMA copy): 4 bytes per group row, ~0.1 GB for 60 000 batches of 100-1200 rows
  const size_t gbytes = (size_t)std::max<long long>(n_gsib, 1) * sizeof(int);
  if (st.h_gsib_bytes < gbytes) {
    if (st.h_gsib)
      FRPA_HIP(hipHostFree(st.h_gsib));
    st.h_gsib = nullptr;
    st.h_gsib_bytes = 0;
    FRPA_HIP(hipHostMalloc((void **)&st.h_gsib, gbytes, hipHostMallocDefault));
    st.h_gsib_bytes = gbytes;
  }
  {
    int *dst = st.h_gsib;
    auto fill = [&](int k0, int k1) {
      for (int k = k0; k < k1; k++)
        std::copy(rows[gorder[k]].begin(), rows[gorder[k]].end(), dst + gdesc[k].ib_off);
    };
    host_parallel(GROUP_PIECES, host_threads(), [&](int c) {
      fill((int)((long long)ng * c / GROUP_PIECES), (int)((long long)ng * (c + 1) / GROUP_PIECES));
    });
  }
  st.d_gsib.alloc((size_t)std::max<long long>(n_gsib, 1));
  if (n_gsib)
    FRPA_HIP(hipMemcpy(st.d_gsib.ptr, st.h_gsib, (size_t)n_gsib * sizeof(int), hipMemcpyHostToDevice));
  st.d_gmem.upload(gmem);
  st.d_members.upload(members.empty() ? std::vector<GMember>(1) : members);
  st.d_mdi.upload(mdi.empty() ? std::vector<int>{0} : mdi);
  st.d_rowmap.alloc((size_t)std::max<long long>(rm_off, 1));
  if (ng) {
    hipLaunchKernelGGL(h1_rowmap_kernel, dim3(ng), dim3(256), 0, s, st.d_gdesc.ptr, st.d_gmem.ptr,
                       st.d_members.ptr, st.d_mdi.ptr, P.d_desc.ptr, R.d_sib.ptr, R.d_srow.ptr, st.d_gsib.ptr,
                       st.d_rowmap.ptr);
    check_launch("h1_rowmap_kernel");
  }
  st.merge = merge;
  return gdesc;
}

// Work items of the kernel: ranges of row tiles I of a group (tile I has the 4 (nT - I) units of tiles
// J >= I; a range closes once it holds 8 or more column tiles), all items of a group consecutive in
// one XCD's share of the grid, so that they run side by side and read the group's phi digits from
// that XCD's L2 (a group alone per block streamed them from memory once per row tile).  The groups
// in the caller's (spatial) order are cut into N_XCD contiguous ranges of equal work, one per XCD,
// so that the H elements one XCD updates at a time are few and neighbouring.  Workgroup w of the
// launch runs on XCD w % N_XCD: item k of XCD x's list is grid position N_XCD k + x, and shorter
// lists are padded with empty items.
std::vector<AzItem> xcd_items(const std::vector<BatchDesc> &gdesc) {
  const int ng = (int)gdesc.size();
  std::vector<int> spatial(ng);
  std::iota(spatial.begin(), spatial.end(), 0);
  std::stable_sort(spatial.begin(), spatial.end(), [&](int a, int b) { return gdesc[a].b < gdesc[b].b; });
  auto work = [&](int g) {
    const double nT = (double)((gdesc[g].nc + 63) / 64);
    return (double)((gdesc[g].np + 63) / 64) * nT * (nT + 1.0) / 2.0;
  };
  double total = 0.0;
  for (int g = 0; g < ng; g++)
    total += work(g);
  std::vector<std::vector<AzItem>> lists(N_XCD);
  double before = 0.0;  // work of the groups ahead of g in the spatial order
  for (int g : spatial) {
    const int x = total > 0.0 ? std::min(N_XCD - 1, (int)(before / total * N_XCD)) : 0;
    const int nT = (gdesc[g].nc + 63) / 64;
    for (int I = 0, I0 = 0, w = 0; I < nT; I++) {
      w += nT - I;
      if (w >= 8 || I == nT - 1) {
        lists[x].push_back(AzItem{g, I0, I + 1, 0});
        I0 = I + 1;
        w = 0;
      }
    }
    before += work(g);
  }
  size_t len = 0;
  for (const auto &l : lists)
    len = std::max(len, l.size());
  std::vector<AzItem> out(N_XCD * len, AzItem{-1, 0, 0, 0});
  for (int x = 0; x < N_XCD; x++)
    for (size_t k = 0; k < lists[x].size(); k++)
      out[N_XCD * k + x] = lists[x][k];
  return out;
}

template <int S>
void prepare_f(Rho1Plan &P, AdjOzState &st, const double *wave, double merge, hipStream_t s) {
  const std::vector<BatchDesc> gdesc = build_groups(P, st, merge, s);
  const std::vector<AzItem> items = xcd_items(gdesc);
  st.n_items = (int)items.size();
  st.d_items.upload(items.empty() ? std::vector<AzItem>(1, AzItem{-1, 0, 0, 0}) : items);
  std::vector<long long> off(st.n_groups + 1, 0);
  size_t rows = 0;
  for (int i = 0; i < st.n_groups; i++) {
    const BatchDesc &d = gdesc[i];
    const int nc64 = (d.nc + 63) / 64 * 64, npk = (d.np + 63) / 64 * 64;
    off[i + 1] = off[i] + (long long)S * nc64 * npk;
    rows += d.nc;
  }
  st.d_off.upload(off);
  st.bytes = std::max<long long>(off[st.n_groups], 16);
  st.d_fdig.alloc(st.bytes);
  st.d_finv.alloc(std::max<size_t>(rows, 1));
  std::vector<int2> tiles;  // (group, 64-row tile) of the digits kernel
  for (int i = 0; i < st.n_groups; i++)
    for (int I = 0; I < (gdesc[i].nc + 63) / 64; I++)
      tiles.push_back(make_int2(i, I));
  if (!tiles.empty()) {
    st.d_dtiles.upload(tiles);
    hipLaunchKernelGGL((h1_digits_kernel<S>), dim3((unsigned)tiles.size()), dim3(256), 0, s, st.d_dtiles.ptr,
                       st.d_gdesc.ptr, st.d_off.ptr, st.d_gmem.ptr, st.d_members.ptr, st.d_rowmap.ptr, wave,
                       st.d_fdig.ptr, st.d_finv.ptr);
    check_launch("h1_digits_kernel");
  }
  int nvp = 0;
  for (const BatchDesc &d : P.h_desc)
    nvp = std::max(nvp, d.pstart + d.np);
  st.d_pmax.alloc(std::max(nvp, 1));
  fill_f64(st.d_pmax.ptr, st.d_pmax.n, 0.0, s);  // points of empty batches stay 0
  if (P.n_valid) {
    hipLaunchKernelGGL(h1_pointmax_kernel, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, wave, st.d_pmax.ptr);
    check_launch("h1_pointmax_kernel");
  }
  st.S = S;
  st.wave = wave;
}

template <int S>
void run(Rho1Plan &P, AdjOzState &st, int T, const double *wave, const double *g, unsigned long long *Hs,
         hipStream_t s) {
  const int nvp = (int)st.d_pmax.n;
  const int bblocks = std::max(1, std::min((nvp + 255) / 256, 4096));
  st.d_bound.grow((size_t)T);
  st.d_bpart.grow((size_t)bblocks * T);
  hipLaunchKernelGGL(h1_bound_kernel, dim3(bblocks, T), dim3(256), 0, s, P.d_map.ptr, g, P.n_full, st.d_pmax.ptr,
                     nvp, st.d_bpart.ptr);
  check_launch("h1_bound_kernel");
  hipLaunchKernelGGL(h1_bound_sum_kernel, dim3(T), dim3(256), 0, s, st.d_bpart.ptr, bblocks, st.d_bound.ptr);
  check_launch("h1_bound_sum_kernel");
  const char *pc = std::getenv("FRPA_ADJOINT_PC");
  auto kern = (pc && pc[0] == '0') ? h1_ozaki_kernel<S, false> : h1_ozaki_kernel<S, true>;
  const int spin_limit = P.adj_spin_limit > 0 ? P.adj_spin_limit : PC_SPIN;
  if (st.n_items) {
    hipLaunchKernelGGL(kern, dim3(st.n_items, T), dim3(AZ_THREADS), 0, s, st.d_items.ptr, st.d_gdesc.ptr, st.d_off.ptr,
                       st.d_gsib.ptr, P.d_map.ptr, g, P.n_full, st.d_fdig.ptr, st.d_finv.ptr, P.n_basis,
                       st.d_bound.ptr, Hs, st.d_spin_fail.ptr, spin_limit);
    check_launch("h1_ozaki_kernel");
  }
}

} // namespace

// (Re)builds the adjoint's groups and phi digits for this wave (once per wave; frpa_rho1_prepare_wave
// calls it again after the wave changed in place).
void adjoz_prepare(Rho1Plan &P, int S, const double *wave, hipStream_t s) {
  FRPA_REQUIRE(S == 4 || S == 5, FRPA_ERR_INVALID, "INT8 split supports 4 or 5 slices, not %d", S);
  if (!P.adjoz)
    P.adjoz = new AdjOzState();
  const double merge = merge_threshold();
  if (S == 5)
    prepare_f<5>(P, *P.adjoz, wave, merge, s);
  else
    prepare_f<4>(P, *P.adjoz, wave, merge, s);
}

// Waits of the INT8 adjoint that gave up since the last read, read and cleared (synchronises s).
int adjoz_take_wait_failures(Rho1Plan &P, hipStream_t s) {
  if (!P.adjoz || !P.adjoz->d_spin_fail.ptr)
    return 0;
  int fails = 0;
  FRPA_HIP(hipMemcpyAsync(&fails, P.adjoz->d_spin_fail.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  FRPA_HIP(hipStreamSynchronize(s));
  if (fails) {
    fill_i32(P.adjoz->d_spin_fail.ptr, 1, 0, s);
    FRPA_HIP(hipStreamSynchronize(s));
  }
  return fails;
}

void adjoz_invalidate(Rho1Plan &P) {
  if (P.adjoz)
    P.adjoz->wave = nullptr;
}

bool adjoz_built(const Rho1Plan &P) { return P.adjoz && P.adjoz->wave; }

long long adjoz_bytes(const Rho1Plan &P) {
  if (!adjoz_built(P))
    return 0;
  const AdjOzState &st = *P.adjoz;
  return st.bytes + (long long)(st.d_finv.n + st.d_pmax.n) * 8;
}

const double *h1_ozaki_push(Rho1Plan &P, int T, int S, const double *wave, const double *g, unsigned long long *Hs,
                            hipStream_t s) {
  FRPA_REQUIRE(S == 4 || S == 5, FRPA_ERR_INVALID, "INT8 split supports 4 or 5 slices, not %d", S);
  if (!P.adjoz)
    P.adjoz = new AdjOzState();
  AdjOzState &st = *P.adjoz;
  if (!st.d_spin_fail.ptr) {  // zeroed once; cleared again only when read (adjoz_take_wait_failures)
    st.d_spin_fail.alloc(1);
    fill_i32(st.d_spin_fail.ptr, 1, 0, s);
  }
  const double merge = merge_threshold();
  if (st.wave != wave || st.S != S || st.merge != merge)
    adjoz_prepare(P, S, wave, s);
  if (S == 5)
    run<5>(P, st, T, wave, g, Hs, s);
  else
    run<4>(P, st, T, wave, g, Hs, s);
  return st.d_bound.ptr;
}

} // namespace frpa
