// rho1 forward with the GEMM W = P1loc * phi emulated on INT8 MFMA (Ozaki-style split).
//
// Both operands are written as sums of signed 8-bit digits in radix 254:
//   P(r, c) * sa_r = sum_{s<S} a_s(r, c) 254^-s,   sa_r = 127 / bound_r             (per P row)
//   phi(i, p) * sb_p = sum_{s<S} b_s(i, p) 254^-s,  sb_p = 127 / max_i |phi(i, p)|  (per point)
// with round-to-nearest digits: |x| <= 127 and |x - rint(x)| <= 1/2, so every digit fits int8 and
// carries log2(254) ~ 8 bits.  The digit products a_s * b_u are exact INT8 MFMA GEMMs with exact
// int32 accumulation (|sum| <= 5 * nc * 127^2 < 2^31 for nc < 26000); the levels s + u <= S - 1
// are kept, S(S+1)/2 products, and recombined in FP64:
//   W(i, p) = sum_l 254^-l C_l(i, p) / (sa_i sb_p),   C_l = sum_{s+u=l} a_s b_u   (exact int32)
// S = 5 (15 INT8 GEMMs) keeps max|d rho1| / max|rho1| ~ 1e-12; S = 4 (10 GEMMs) ~ 5e-10 on the
// synthetic workloads.
//
// Symmetry: rho(p) = phi^T P phi only sees the symmetric part Ps = (P1 + P1^T) / 2, so with the
// batch's compute rows in ascending order
//   rho(p) = 2 sum_i phi_i(p) sum_{k >= i} P'(i, k) phi_k(p),   P'(i, k) = Ps(i, k) (k > i), Ps(i, i) / 2
// and only the 64 x 64 tile pairs J >= I of P' are multiplied (half the MFMAs and half the P1
// gathers of the full product).  The ddot, the bias and the scatter stay FP64.
//
// GGA gradient (GRAD): grad rho1(j, p) = 2 sum_i dphi(i, j, p) W(i, p) with W = P1loc phi for P1 as
// given (evaluate_first_order_rho.cpp:161-200), so that launch multiplies every tile pair (I, J) of
// the digits of P1 itself -- twice the MFMAs and gathers of the symmetric form -- and takes
// rho1 = sum_i phi_i W_i and the gradient from the same W; dphi is read in FP64 in the epilogue.
//
// Data: P1 is relabeled by first appearance over the batches (get_relabel), so a batch's compute set
// is a few long runs of consecutive labels and the per-tile gathers touch few cache lines.  Per call
// the P' digits are written as one int64 per element (its S digits) in the new labels; the phi
// digits are prepared once per wave (frpa_rho1_prepare_wave) as S int8 planes per batch, rows in the
// batch's sorted order, ncK = nc rounded up to 64 (zero rows): a plane is a sequence of 64-row chunks,
// each np points x 64 bytes (row k, point p at (k / 64) np 64 + 64 p + k % 64), so a lane's B fragment
// (16 consecutive k of one point) is one 16-byte load and a chunk's staged points one contiguous run.
//
// Launch shapes: 8-wave blocks over 128-point passes or 4-wave blocks over 64-point passes with the
// phi chunks staged through registers, or (ASYNC) 4-wave blocks whose phi chunks go straight from
// global memory into LDS and whose P' gathers for the next K step are in flight during the current
// step's MFMAs; work items are whole batches, point passes, or ranges of row tiles of a point pass.
#include <algorithm>
#include <utility>
#include <vector>

#include "rho1_common.hpp"

namespace frpa {

typedef int v4i __attribute__((ext_vector_type(4)));

struct OzakiState {
  int S = 0;
  const double *wave = nullptr;     // the wave the digits were made from
  long long bytes = 0;
  DevBuf<long long> d_boff;         // per BatchDesc: byte offset of its S digit planes
  DevBuf<signed char> d_bdig;       // phi digits
  DevBuf<double> d_binv;            // 1 / sb_p per valid point (indexed by pstart + p)
  // per call
  DevBuf<unsigned long long> d_adig;  // P' digits (byte s = digit s) in the new labels, (n, n, T)
  DevBuf<double> d_ainv;            // 1 / sa in the new labels, (n_basis, T)
  DevBuf<unsigned long long> d_rowbits;  // bound_r = max(max_c |P1(r, c)|, max_c |P1(c, r)|), IEEE bits
  int n_valid_points = 0;
};

void ozaki_release(OzakiState *st) { delete st; }

namespace {

constexpr double RADIX = 254.0;

__device__ __forceinline__ int ceil64(int x) { return (x + 63) & ~63; }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = fmax(v, __shfl_xor(v, o));
  return v;
}

// ---- phi digits (one-off) -------------------------------------------------------------------
// One wave per point: lanes run over the batch's rows in sorted order (row k is the caller's local
// row srow[k]); a wave store writes the point's 64-byte run of one chunk of a plane.  Batches of up
// to PHI_ROWS rows (the usual case) stage srow in LDS and hold the point's column in registers, so
// all its loads are in flight at once and the column is read once; larger ones loop over it twice.
constexpr int PHI_ROWS = 1280;
template <int S>
__global__ __launch_bounds__(256) void phi_digits_kernel(const BatchDesc *__restrict__ descs,
                                                         const long long *__restrict__ boff,
                                                         const int *__restrict__ srow_all,
                                                         const double *__restrict__ wave,
                                                         signed char *__restrict__ bdig, double *__restrict__ binv) {
  const BatchDesc d = descs[blockIdx.x];
  const int ld = (d.nc + 3) & ~3, ncK = ceil64(d.nc);
  const double *W = wave + d.wave_off;
  const int *srow = srow_all + d.ib_off;
  signed char *out = bdig + boff[blockIdx.x];
  const size_t plane = (size_t)ncK * d.np, chunk = (size_t)64 * d.np;
  const int lane = threadIdx.x & 63;
  __shared__ int srow_s[PHI_ROWS];
  const bool fits = d.nc <= PHI_ROWS;
  if (fits)
    for (int k = threadIdx.x; k < d.nc; k += 256)
      srow_s[k] = srow[k];
  __syncthreads();
  auto put = [&](int k, int p, double r) {  // the S digits of r at row k, point p
#pragma unroll
    for (int sl = 0; sl < S; sl++) {
      const double dg = rint(r);
      out[sl * plane + chunk * (k >> 6) + (size_t)64 * p + (k & 63)] = (signed char)dg;
      r = (r - dg) * RADIX;
    }
  };
  for (int p = threadIdx.x >> 6; p < d.np; p += 4) {
    const double *Wp = W + (size_t)ld * p;
    if (fits) {
      double v[PHI_ROWS / 64];
      double mx = 0.0;
#pragma unroll
      for (int i = 0; i < PHI_ROWS / 64; i++) {
        const int k = lane + 64 * i;
        v[i] = k < d.nc ? Wp[srow_s[k]] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < PHI_ROWS / 64; i++)
        mx = fmax(mx, fabs(v[i]));
      mx = wave_max(mx);
      const double sb = mx > 0.0 ? 127.0 / mx : 1.0;
      if (lane == 0)
        binv[d.pstart + p] = 1.0 / sb;
#pragma unroll
      for (int i = 0; i < PHI_ROWS / 64; i++) {
        const int k = lane + 64 * i;
        if (k < ncK)
          put(k, p, v[i] * sb);
      }
    } else {
      double mx = 0.0;
      for (int k = lane; k < d.nc; k += 64)
        mx = fmax(mx, fabs(Wp[srow[k]]));
      mx = wave_max(mx);
      const double sb = mx > 0.0 ? 127.0 / mx : 1.0;
      if (lane == 0)
        binv[d.pstart + p] = 1.0 / sb;
      for (int k = lane; k < ncK; k += 64)
        put(k, p, k < d.nc ? Wp[srow[k]] * sb : 0.0);
    }
  }
}

// ---- P' digits (per call) -------------------------------------------------------------------
// bound_r: row maxima (column chunks per block) and column maxima (one block per column), combined
// with atomicMax on the bit pattern (non-negative doubles order like their IEEE bits); rowbits
// starts at 0.  For a symmetric P1 both give max_c |P1(r, c)|.
__global__ void p1_rowmax_kernel(const double *__restrict__ P1, int n_basis, int cols_per_block,
                                 unsigned long long *__restrict__ rowbits) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.z;
  if (r >= n_basis)
    return;
  const double *Pt = P1 + (size_t)n_basis * n_basis * t;
  const int c0 = blockIdx.y * cols_per_block, c1 = min(n_basis, c0 + cols_per_block);
  double mx = 0.0;
  for (int c = c0; c < c1; c++)
    mx = fmax(mx, fabs(Pt[r + (size_t)n_basis * c]));
  atomicMax(&rowbits[r + (size_t)n_basis * t], (unsigned long long)__double_as_longlong(mx));
}

__global__ __launch_bounds__(256) void p1_colmax_kernel(const double *__restrict__ P1, int n_basis,
                                                        unsigned long long *__restrict__ rowbits) {
  __shared__ double red[4];
  const int c = blockIdx.x, t = blockIdx.y;
  const double *col = P1 + (size_t)n_basis * n_basis * t + (size_t)n_basis * c;
  double mx = 0.0;
  for (int r = threadIdx.x; r < n_basis; r += 256)
    mx = fmax(mx, fabs(col[r]));
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(&rowbits[c + (size_t)n_basis * t], (unsigned long long)__double_as_longlong(mx));
  }
}

// 1 / sa in the new labels: ainv[x + n t] = bound[pinv[x] + n t] / 127
__global__ void p1_rowscale_kernel(const unsigned long long *__restrict__ rowbits, const int *__restrict__ pinv,
                                   int n_basis, size_t n, double *__restrict__ ainv) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const size_t x = i % n_basis, t = i / n_basis;
  const double mx = __longlong_as_double((long long)rowbits[pinv[x] + (size_t)n_basis * t]);
  ainv[i] = mx > 0.0 ? mx / 127.0 : 1.0;
}

// Digits of P'(pinv[x], pinv[y]) / ainv[x] (FULL: of P1(pinv[x], pinv[y]) / ainv[x]) at x + n y,
// digit s in byte s of one 8-byte word (one gather load per element); one 64 x 64 tile of the new
// labels per block; the transposed tile P1(pinv[y], pinv[x]) goes through LDS so both reads run down
// columns.  |P1(r, c)| and |P'(r, c)| are both at most bound_r, the row scale.  The symmetric
// forward reads P' only at x <= y (a batch's sorted rows ascend in the new labels, and it multiplies
// the column tiles J >= I), so only the tiles on and above the diagonal are made; a padded column
// may read a stale word below it, which meets zero phi digits (exact integer products).
template <int S, bool FULL>
__global__ __launch_bounds__(256) void p1_digits_kernel(const double *__restrict__ P1, const int *__restrict__ pinv,
                                                        const double *__restrict__ ainv, int n_basis,
                                                        unsigned long long *__restrict__ adig) {
  __shared__ double tr[64][65];
  if (!FULL && blockIdx.x > blockIdx.y)
    return;  // below the diagonal: never read by the symmetric forward
  const int x0 = 64 * blockIdx.x, y0 = 64 * blockIdx.y, t = blockIdx.z;
  const size_t nn = (size_t)n_basis * n_basis;
  const double *Pt = P1 + nn * t;
  const int a = threadIdx.x & 63, b0 = threadIdx.x >> 6;
  // transposed read: thread (j = a, i = b0 + 4u) loads P1(pinv[y0 + j], pinv[x0 + i])
  if (!FULL) {
    const int y = y0 + a;
    const int cy = y < n_basis ? pinv[y] : 0;
#pragma unroll 4
    for (int u = 0; u < 16; u++) {
      const int i = b0 + 4 * u, x = x0 + i;
      tr[i][a] = (x < n_basis && y < n_basis) ? Pt[cy + (size_t)n_basis * pinv[x]] : 0.0;
    }
  }
  __syncthreads();
  const int x = x0 + a;
  if (x >= n_basis)
    return;
  const int rx = pinv[x];
  const double inv = 1.0 / ainv[x + (size_t)n_basis * t];
#pragma unroll 4
  for (int u = 0; u < 16; u++) {
    const int j = b0 + 4 * u, y = y0 + j;
    if (y >= n_basis)
      break;
    double v;
    if (FULL) {
      v = Pt[rx + (size_t)n_basis * pinv[y]];
    } else {
      v = 0.5 * (Pt[rx + (size_t)n_basis * pinv[y]] + tr[a][j]);
      if (x == y)
        v *= 0.5;
    }
    double z = v * inv;
    unsigned long long packed = 0;
#pragma unroll
    for (int sl = 0; sl < S; sl++) {
      const double dg = rint(z);
      packed |= (unsigned long long)(unsigned char)(signed char)dg << (8 * sl);
      z = (z - dg) * RADIX;
    }
    adig[x + (size_t)n_basis * y + nn * t] = packed;
  }
}

// byte s of each of four packed elements -> one dword (element q in byte q)
__device__ __forceinline__ int gather_bytes(unsigned long long e0, unsigned long long e1, unsigned long long e2,
                                            unsigned long long e3, int s) {
  const unsigned b0 = (unsigned)(e0 >> (8 * s)) & 0xffu, b1 = (unsigned)(e1 >> (8 * s)) & 0xffu;
  const unsigned b2 = (unsigned)(e2 >> (8 * s)) & 0xffu, b3 = (unsigned)(e3 >> (8 * s)) & 0xffu;
  return (int)(b0 | (b1 << 8) | (b2 << 16) | (b3 << 24));
}

// ---- fused forward ------------------------------------------------------------------------
// Block: 4 H waves; wave w owns rows 16 (w & 3) .. +15 of each 64-row tile I and, for H = 2, one half
// of the block's point chunk (w >> 2): up to H * 16 * NPG points per pass, so the P' digits of a
// batch are gathered once for batches of up to 16 H NPG points.  H = 2 (8 waves, 128-point passes
// at NPG = 4) fills the LDS with one block per CU; H = 1 (4 waves, 64-point passes) fits two blocks
// per CU, whose barriers and gathers overlap each other's MFMAs.  For row tile I the K loop runs over the
// column tiles J = I .. nT - 1 (chunks of 64 sorted rows); the phi digits of chunk J + 1 are staged
// through registers into the other LDS buffer and the P' digits of chunk J + 1 are gathered into
// registers while chunk J's MFMAs run (one barrier per chunk).
constexpr int BROW = 80;  // bytes per (digit, point) row of a staged B chunk: 64 k + 16 pad
// ASYNC: the B chunks are staged with direct global -> LDS loads (global_load_lds_dwordx4: lane l of
// a wave writes its 16 bytes at the wave's LDS base + 16 l), so a chunk is 1 KB blocks of 16
// unpadded 64-byte rows and no registers hold it in flight; the P' gathers of chunk J + 1 are
// issued before chunk J's MFMAs and both are waited for once, before the barrier that ends the step.
constexpr int BROW_ASYNC = 64;
typedef __attribute__((address_space(3))) void lds_void;

template <int S>
__device__ __forceinline__ void repack_a(const unsigned long long (&e)[16], v4i (&a)[S]) {
#pragma unroll
  for (int sl = 0; sl < S; sl++)
    a[sl] = v4i{gather_bytes(e[0], e[1], e[2], e[3], sl), gather_bytes(e[4], e[5], e[6], e[7], sl),
                gather_bytes(e[8], e[9], e[10], e[11], sl), gather_bytes(e[12], e[13], e[14], e[15], sl)};
}

// ASYNC gather: P' digits of 16 elements as (lo, hi) dword pairs, loaded through a buffer resource
// with 32-bit byte offsets (one add per element; every offset lies inside the perturbation's
// n_basis^2 elements), and the digit planes rebuilt with byte permutes: for elements 4 h .. 4 h + 3,
// digits 0-3 are a 4 x 4 byte transpose of the low dwords (8 v_perm) and digit 4 collects byte 0 of
// the high dwords (3 v_perm).  v_perm_b32(x, y, sel): selector byte 0-3 takes byte of y, 4-7 of x.
typedef unsigned u2 __attribute__((ext_vector_type(2)));

template <int S>
__device__ __forceinline__ void repack_a_perm(const u2 (&el)[16], v4i (&a)[S]) {
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const unsigned L0 = el[4 * h].x, L1 = el[4 * h + 1].x, L2 = el[4 * h + 2].x, L3 = el[4 * h + 3].x;
    const unsigned t01lo = __builtin_amdgcn_perm(L1, L0, 0x05010400u), t01hi = __builtin_amdgcn_perm(L1, L0, 0x07030602u);
    const unsigned t23lo = __builtin_amdgcn_perm(L3, L2, 0x05010400u), t23hi = __builtin_amdgcn_perm(L3, L2, 0x07030602u);
    a[0][h] = (int)__builtin_amdgcn_perm(t23lo, t01lo, 0x05040100u);
    a[1][h] = (int)__builtin_amdgcn_perm(t23lo, t01lo, 0x07060302u);
    a[2][h] = (int)__builtin_amdgcn_perm(t23hi, t01hi, 0x05040100u);
    a[3][h] = (int)__builtin_amdgcn_perm(t23hi, t01hi, 0x07060302u);
    if constexpr (S == 5) {
      const unsigned u = __builtin_amdgcn_perm(el[4 * h + 1].y, el[4 * h].y, 0x05010400u);
      const unsigned v = __builtin_amdgcn_perm(el[4 * h + 3].y, el[4 * h + 2].y, 0x05010400u);
      a[4][h] = (int)__builtin_amdgcn_perm(v, u, 0x05040100u);
    }
  }
}

// One point chunk; this wave multiplies NG (exact, compile-time) groups of 16 points starting at
// group gp0 of the chunk.  Every wave of the block runs the same barriers whatever its NG.
// LDS: three staged phi-digit chunks -- buffer 2 holds the diagonal chunk J = I for the whole row
// tile (its digits also give phi(i, p) for the ddot, so the FP64 wave is not read), buffers 0 and 1
// alternate for J > I.
// GRAD: the gradient launch (every column tile J, starting at J = I so that buffer 2 still holds
// the diagonal chunk; P1's own digits; dphi from g.gwave in the epilogue).
struct GradArgs {
  const double *gwave;  // gradient_wave_batches_compress: batch block at 3 wave_off, (nc_pad, 3, np)
  const double *gbias;  // first_order_gradient_rho_bias_batches_atoms (3, nmb, n_batches, T), nullable
  double *grad;         // first_order_gradient_rho (3, nmb, n_batches, T)
  const int *srow_all;  // caller's local row of every sorted row (Relabel::d_srow)
};

// Row tiles [I0, I1) of the chunk.  Each row tile's per-point sums over its 64 rows (per component
// c: rho1, and with GRAD the three gradient components) are combined across the four row groups in
// LDS, ((g0 + g1) + g2) + g3, and then either added up over the row tiles in order by this block
// (part == nullptr: the whole batch, final values written) or stored at part[(I C + c) npc + p] for
// the finishing kernel, which adds them up in the same order -- so every launch order gives the
// same bits.
// Final value of component c at point p of batch d from its sum over the row tiles
// (evaluate_first_order_rho.cpp:187-200): rho1 = 2 sum (symmetric form; GRAD: 1 sum) + bias,
// grad rho1 = 2 sum - gradient bias.
template <bool GRAD>
__device__ __forceinline__ void ozaki_final(int c, double tot, const BatchDesc &d, int t, int p,
                                            const int *__restrict__ map, const double *__restrict__ bias,
                                            double *__restrict__ rho, int n_full, int nmb, int n_batches_all,
                                            int n_tile, const GradArgs &ga) {
  const size_t bi = (size_t)p + (size_t)nmb * (d.b + (size_t)n_batches_all * t);
  if (c == 0) {
    double sum = (GRAD ? 1.0 : 2.0) * tot;
    if (bias)
      sum += bias[bi];
    if (t < n_tile)
      rho[(size_t)map[d.pstart + p] + (size_t)n_full * t] = sum;
  } else {
    double gv = 2.0 * tot;
    if (ga.gbias)
      gv -= ga.gbias[(c - 1) + 3 * bi];
    ga.grad[(c - 1) + 3 * bi] = gv;
  }
}

template <int S, int NPG, int NG, int H, bool GRAD, bool ASYNC>
__device__ __forceinline__ void ozaki_chunk(signed char *Bs, int *jbs, double *red, const BatchDesc &d, int t,
                                            int pc, int npc, int gp0, int I0, int I1, double *__restrict__ part,
                                            const int *__restrict__ sib,
                                            const int *__restrict__ map, const unsigned long long *__restrict__ At,
                                            const double *__restrict__ ai,
                                            const signed char *__restrict__ Bg, const double *__restrict__ binv,
                                            int n_basis, const double *__restrict__ bias, double *__restrict__ rho,
                                            int n_full, int nmb, int n_batches_all, int n_tile, const GradArgs &ga) {
  constexpr int THREADS = 256 * H;
  constexpr int PCH = 16 * H * NPG;             // point slots per chunk
  constexpr int RB = ASYNC ? BROW_ASYNC : BROW;  // bytes per staged (digit, point) row
  constexpr int BUF = S * PCH * RB;             // bytes per LDS buffer
  constexpr int NQ = S * PCH * 4;               // 16-byte pieces of one staged chunk
  constexpr int QPT = (NQ + THREADS - 1) / THREADS;
  const int nc = d.nc, ncK = ceil64(nc);
  const size_t plane = (size_t)ncK * d.np;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rg = w & 3;
  const int lr = lane & 15, lk = lane >> 4;
  const int nT = ncK / 64;
  constexpr int C = GRAD ? 4 : 1;
  // per-tile sums of row group rg, component c at redI[(4 c + rg) PCH + p]; the register-staged
  // GRAD launch keeps them in staging buffer 1, which no wave touches from the last K step of a tile
  // until the second K step of the next one (the ASYNC launch gives them an array of their own)
  double *redI = (GRAD && !ASYNC) ? reinterpret_cast<double *>(Bs + BUF) : red;
  // reducing threads: element e = tid + THREADS u is component e / PCH of point e % PCH
  constexpr int RPT = (C * PCH + THREADS - 1) / THREADS;
  double tot[RPT];
#pragma unroll
  for (int u = 0; u < RPT; u++)
    tot[u] = 0.0;
  auto reduce_tile = [&](int I) {
#pragma unroll
    for (int u = 0; u < RPT; u++) {
      const int e = tid + THREADS * u, rc = e / PCH, rp = e % PCH;
      if (rc < C && rp < npc) {
        const double *r = redI + 4 * rc * PCH + rp;
        const double v = ((r[0] + r[PCH]) + r[2 * PCH]) + r[3 * PCH];
        if (part)
          part[((size_t)I * C + rc) * npc + rp] = v;
        else
          tot[u] += v;
      }
    }
  };
  const int ncp = (nc + 3) & ~3;
  const double *GW = GRAD ? ga.gwave + 3 * d.wave_off : nullptr;
  const int *srow = GRAD ? ga.srow_all + d.ib_off : nullptr;

  v4i bq[QPT];
  auto load_b = [&](int c) {
    const int j0 = 64 * c;
#pragma unroll
    for (int u = 0; u < QPT; u++) {
      const int e = tid + THREADS * u;
      const int q = e & 3, p = (e >> 2) % PCH, sl = (e >> 2) / PCH;
      bq[u] = (e < NQ && p < npc)
                  ? *reinterpret_cast<const v4i *>(Bg + sl * plane + (size_t)64 * d.np * (j0 >> 6) +
                                                   (size_t)64 * (pc + p) + 16 * q)
                  : v4i{0, 0, 0, 0};
    }
  };
  auto store_b = [&](int c, int buf) {
    signed char *B = Bs + buf * BUF;
#pragma unroll
    for (int u = 0; u < QPT; u++) {
      const int e = tid + THREADS * u;
      const int q = e & 3, p = (e >> 2) % PCH, sl = (e >> 2) / PCH;
      if (e < NQ)
        *reinterpret_cast<v4i *>(&B[(sl * PCH + p) * BROW + 16 * q]) = bq[u];
    }
    if (tid < 64) {
      const int j = 64 * c + tid;
      jbs[64 * buf + tid] = j < nc ? sib[j] : -1;
    }
  };
  // ASYNC: chunk c's digits straight into buffer buf, 1 KB block k of 64 pieces by wave k mod (4 H);
  // points past npc read the last point's digits (their sums are never used)
  auto stage_b = [&](int c, signed char *B) {
    constexpr int NBLK = S * PCH / 16, NW = 4 * H;
#pragma unroll
    for (int u = 0; u < (NBLK + NW - 1) / NW; u++) {
      const int blk = NW * u + w;
      if (blk < NBLK) {
        const int e = 64 * blk + lane;
        const int q = e & 3, p = min((e >> 2) % PCH, npc - 1), sl = (e >> 2) / PCH;
        __builtin_amdgcn_global_load_lds(
            static_cast<const void *>(Bg + sl * plane + (size_t)64 * d.np * c + (size_t)64 * (pc + p) + 16 * q),
            (lds_void *)(B + 1024 * blk), 16, 0, 0);
      }
    }
  };

  for (int I = I0; I < I1; I++) {
    const int irow = 64 * I + 16 * rg + lr;
    const bool iok = irow < nc;
    const size_t arow = iok ? (size_t)sib[irow] : 0;
    v4i acc[S][NG > 0 ? NG : 1];
#pragma unroll
    for (int l = 0; l < S; l++)
#pragma unroll
      for (int g = 0; g < NG; g++)
        acc[l][g] = v4i{0, 0, 0, 0};

    unsigned long long e[16];
    // A fragments: row irow, k = 64 c + 16 lk + q, q < 16; P'(i, k) = 0 below the diagonal (k < i);
    // GRAD: the full P1
    auto load_a = [&](int c, int buf) {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int jb = jbs[64 * buf + 16 * lk + q];
        const bool ok = iok && jb >= 0 && (GRAD || c > I || 64 * c + 16 * lk + q >= irow);
        e[q] = ok ? At[arow + (size_t)n_basis * jb] : 0ull;
      }
    };
    block_sync();  // the previous tile is done with all buffers, its row-group sums are in LDS
    if (I > I0)
      reduce_tile(I - 1);
    v4i a[S];
    if constexpr (ASYNC) {
      // symmetric launch: column tiles c = I .. nT - 1; GRAD: every tile, c = (I + st) mod nT; the
      // chunk of step st + 1 goes to buffer st & 1.  jbs holds byte
      // offsets of the chunk's columns in At (0 past nc: those k meet zero phi digits); rows past nc
      // read row 0 (their sums are multiplied by zero scales); only the symmetric launch's diagonal
      // chunk is masked.
      unsigned *jco = reinterpret_cast<unsigned *>(jbs);
      const unsigned nb8 = 8u * (unsigned)n_basis;
      const __amdgpu_buffer_rsrc_t rsrc =
          __builtin_amdgcn_make_buffer_rsrc((void *)At, 0, (int)(nb8 * (unsigned)n_basis), 0x00020000);
      const unsigned arowB = 8u * (unsigned)arow;
      u2 el[16];
      auto gather = [&](int buf) {
        const uint4 *cp = reinterpret_cast<const uint4 *>(jco + 64 * buf + 16 * lk);
#pragma unroll
        for (int h = 0; h < 4; h++) {
          const uint4 c4 = cp[h];
          el[4 * h] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.x, 0, 0);
          el[4 * h + 1] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.y, 0, 0);
          el[4 * h + 2] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.z, 0, 0);
          el[4 * h + 3] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.w, 0, 0);
        }
      };
      auto col_off = [&](int j) { return j < nc ? nb8 * (unsigned)sib[j] : 0u; };
      const int nsteps = GRAD ? nT : nT - I;
      auto tile_of = [&](int st) { return GRAD ? (I + st) % nT : I + st; };
      // The staging buffers reach the K steps as distinct restrict pointers, each step's buffers
      // fixed at compile time (step 0 reads the diagonal chunk from B2, then B0 and B1 alternate,
      // two steps per iteration), so the compiler can tell that a step's B reads do not touch the
      // buffer whose direct loads are still in flight and leaves the P' gathers in flight across
      // the MFMAs (with one dynamically indexed buffer it waits for every load before the first
      // B read).
      auto ksteps = [&](signed char *__restrict__ B0, signed char *__restrict__ B1, signed char *__restrict__ B2) {
        stage_b(I, B2);
        if (nsteps > 1)
          stage_b(tile_of(1), B0);
        if (tid < 64) {
          jco[128 + tid] = col_off(64 * I + tid);
          if (nsteps > 1)
            jco[tid] = col_off(64 * tile_of(1) + tid);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        gather(2);
        if (!GRAD) {
#pragma unroll
          for (int q = 0; q < 16; q++)  // P'(i, k) = 0 below the diagonal (k < i)
            if (64 * I + 16 * lk + q < irow)
              el[q] = u2{0u, 0u};
        }
        repack_a_perm<S>(el, a);
        // step st: phi chunk in B, P' column offsets of step st + 1 in jco slot st & 1; chunk st + 2
        // is staged into Bnext once every wave has read B
        auto step = [&](int st, const signed char *B, signed char *Bnext) {
          const bool hn = st + 1 < nsteps, hnn = st + 2 < nsteps;
          const int bn = st & 1, bnn = (st + 1) & 1;
          // labels of chunk st + 2 (index clamped into the batch; used only if hnn), issued before
          // the gathers and consumed after the MFMAs: waiting for them leaves the gathers in flight
          const int jn = 64 * tile_of(st + 2) + lane;
          const int sj = sib[min(jn, nc - 1)];
          __builtin_amdgcn_sched_barrier(0);
          if (hn)
            gather(bn);  // in flight during this step's MFMAs
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int g = 0; g < NG; g++) {
            v4i b[S];
#pragma unroll
            for (int sl = 0; sl < S; sl++)
              b[sl] = *reinterpret_cast<const v4i *>(&B[(sl * PCH + 16 * (gp0 + g) + lr) * RB + 16 * lk]);
#pragma unroll
            for (int sa = 0; sa < S; sa++)
#pragma unroll
              for (int sb = 0; sb < S - sa; sb++)
                acc[sa + sb][g] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[sa], b[sb], acc[sa + sb][g], 0, 0, 0);
          }
          if (hnn && tid < 64)
            jco[64 * bnn + tid] = jn < nc ? nb8 * (unsigned)sj : 0u;
          __builtin_amdgcn_s_waitcnt(0);  // chunk st + 1 staged, its P' gathers arrived
          __syncthreads();
          if (hnn)
            stage_b(tile_of(st + 2), Bnext);
          if (hn)
            repack_a_perm<S>(el, a);
        };
        step(0, B2, B1);
        for (int st = 1; st < nsteps; st += 2) {
          step(st, B0, B0);
          if (st + 1 < nsteps)
            step(st + 1, B1, B1);
        }
      };
      ksteps(Bs, Bs + BUF, Bs + 2 * BUF);
    } else {
      load_b(I);
      store_b(I, 2);
      __syncthreads();
      load_a(I, 2);
      repack_a<S>(e, a);
      // step st visits column tile c = I + st (GRAD: (I + st) mod nT, every tile)
      const int nsteps = GRAD ? nT : nT - I;
      for (int st = 0; st < nsteps; st++) {
        const int cn = GRAD ? (I + st + 1) % nT : I + st + 1;  // the next step's tile
        const int cur = st == 0 ? 2 : (st - 1) & 1, nxt = st & 1;
        const bool more = st + 1 < nsteps;
        if (more)
          load_b(cn);
        const signed char *B = Bs + cur * BUF;
  #pragma unroll
        for (int g = 0; g < NG; g++) {
          v4i b[S];
  #pragma unroll
          for (int sl = 0; sl < S; sl++)
            b[sl] = *reinterpret_cast<const v4i *>(&B[(sl * PCH + 16 * (gp0 + g) + lr) * RB + 16 * lk]);
  #pragma unroll
          for (int sa = 0; sa < S; sa++)
  #pragma unroll
            for (int sb = 0; sb < S - sa; sb++)
              acc[sa + sb][g] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[sa], b[sb], acc[sa + sb][g], 0, 0, 0);
        }
        if (more)
          store_b(cn, nxt);
        __syncthreads();
        if (more) {
          load_a(cn, nxt);
          repack_a<S>(e, a);
        }
      }

    }

    // epilogue: v_mfma_i32_16x16x64_i8 leaves D(4 (lane / 16) + r, lane % 16) in acc[r], so
    // acc[l][g][r] = C_l(i = 64 I + 16 rg + 4 lk + r, p = pc + 16 (gp0 + g) + lr); phi(i, p) sb_p is
    // rebuilt from the diagonal chunk's digits (rows 16 rg + 4 lk .. + 3 are one dword per plane):
    //   rho(p) += W(i, p) phi(i, p) = (sum_l 254^-l C_l / sa_i) (sum_s b_s 254^-s) / sb_p^2
    double rinv[4];
    int grow[4];  // GRAD: caller's local row of sorted row i (-1 past nc)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 64 * I + 16 * rg + 4 * lk + r;
      rinv[r] = i < nc ? ai[sib[i]] : 0.0;
      if (GRAD)
        grow[r] = i < nc ? srow[i] : -1;
    }
    const signed char *BD = Bs + 2 * BUF;
#pragma unroll
    for (int g = 0; g < NG; g++) {
      const int p = 16 * (gp0 + g) + lr;
      double part = 0.0, gp[3] = {0.0, 0.0, 0.0};
      if (p < npc) {
        unsigned dw[S];
#pragma unroll
        for (int sl = 0; sl < S; sl++)
          dw[sl] = *reinterpret_cast<const unsigned *>(&BD[(sl * PCH + p) * RB + 16 * rg + 4 * lk]);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          double wv = (double)acc[S - 1][g][r];
#pragma unroll
          for (int l = S - 2; l >= 0; l--)
            wv = wv * (1.0 / RADIX) + (double)acc[l][g][r];
          double ph = (double)(signed char)(dw[S - 1] >> (8 * r));
#pragma unroll
          for (int sl = S - 2; sl >= 0; sl--)
            ph = ph * (1.0 / RADIX) + (double)(signed char)(dw[sl] >> (8 * r));
          part += wv * rinv[r] * ph;
          if (GRAD && grow[r] >= 0) {
            // W(i, p) sb_p = wv / sa_i; dphi(i, j, p) at grow + ncp (j + 3 p)
            const double *gq = GW + grow[r] + (size_t)ncp * 3 * (pc + p);
#pragma unroll
            for (int j = 0; j < 3; j++)
              gp[j] += wv * rinv[r] * gq[(size_t)ncp * j];
          }
        }
        const double bi = binv[d.pstart + pc + p];
        part *= bi * bi;
#pragma unroll
        for (int j = 0; j < 3; j++)
          gp[j] *= bi;
      }
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      if (GRAD) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          gp[j] += __shfl_xor(gp[j], 16);
          gp[j] += __shfl_xor(gp[j], 32);
        }
      }
      if (lk == 0) {
        redI[rg * PCH + p] = part;
        if (GRAD) {
#pragma unroll
          for (int j = 0; j < 3; j++)
            redI[(4 * (j + 1) + rg) * PCH + p] = gp[j];
        }
      }
    }
  }

  block_sync();
  if (I1 > I0)
    reduce_tile(I1 - 1);
  if (!part) {
#pragma unroll
    for (int u = 0; u < RPT; u++) {
      const int e = tid + THREADS * u, rc = e / PCH, rp = e % PCH;
      if (rc < C && rp < npc)
        ozaki_final<GRAD>(rc, tot[u], d, t, pc + rp, map, bias, rho, n_full, nmb, n_batches_all, n_tile, ga);
    }
  }
}

template <int S, int NPG, int H, bool GRAD, bool ASYNC, int... Gs>
__device__ __forceinline__ void ozaki_dispatch(std::integer_sequence<int, Gs...>, int ng, signed char *Bs, int *jbs,
                                               double *red, const BatchDesc &d, int t, int pc, int npc, int gp0,
                                               int I0, int I1, double *part, const int *sib, const int *map, const unsigned long long *At,
                                               const double *ai, const signed char *Bg,
                                               const double *binv, int n_basis, const double *bias, double *rho,
                                               int n_full, int nmb, int n_batches_all, int n_tile,
                                               const GradArgs &ga) {
  ((ng == Gs ? ozaki_chunk<S, NPG, Gs, H, GRAD, ASYNC>(Bs, jbs, red, d, t, pc, npc, gp0, I0, I1, part, sib, map, At, ai, Bg, binv, n_basis,
                                             bias, rho, n_full, nmb, n_batches_all, n_tile, ga)
             : (void)0),
   ...);
}

template <int S, int NPG, int H, bool GRAD, bool ASYNC>
__global__ __launch_bounds__(256 * H, 2) void rho1_ozaki_kernel(
    const BatchDesc *__restrict__ descs, const int *__restrict__ order, int *__restrict__ counters, int n_pos, int T,
    const int *__restrict__ sib_all, const int *__restrict__ map, const unsigned long long *__restrict__ adig,
    const double *__restrict__ ainv, const signed char *__restrict__ bdig,
    const long long *__restrict__ boff, const double *__restrict__ binv, int n_basis,
    const double *__restrict__ bias, double *__restrict__ rho, int n_full, int nmb, int n_batches_all, int n_tile,
    int items, const TileItem *__restrict__ titems, double *__restrict__ part, long long part_stride, GradArgs ga) {
  constexpr int PCH = 16 * H * NPG;
  __shared__ __attribute__((aligned(16))) signed char Bs[3 * S * PCH * (ASYNC ? BROW_ASYNC : BROW)];
  __shared__ int jbs[192];
  __shared__ double red[(GRAD && ASYNC ? 16 : 4) * PCH];
  __shared__ int s_item;
  const int half = H == 2 ? (int)(threadIdx.x >> 8) : 0;
  int pos = blockIdx.x, t = blockIdx.y;
  for (bool first = true;; first = false) {
    if (counters) {
      if (!claim_item(counters, n_pos, T, pos, t, &s_item))
        break;
    } else if (!first) {
      break;
    }
    const int it = order ? order[pos] : pos;
    // items: (BatchDesc index << 8) | point pass; titems: row tiles of a point pass
    const TileItem ti = titems ? titems[pos] : TileItem{0, 0, 0, 0, 0};
    const int di = titems ? ti.di : items ? it >> 8 : it;
    const BatchDesc d = descs[di];
    const int nT = (d.nc + 63) / 64;
    const int I0 = titems ? ti.I0 : 0, I1 = titems ? ti.I1 : nT;
    double *tpart = titems ? part + part_stride * t + ti.poff : nullptr;
    const int *sib = sib_all + d.ib_off;
    const size_t nn = (size_t)n_basis * n_basis * t;
    const double *ai = ainv + (size_t)n_basis * t;
    const signed char *Bg = bdig + boff[di];
    const int pbeg = titems ? PCH * ti.q : items ? PCH * (it & 255) : 0;
    const int pend = (titems || items) ? min(d.np, pbeg + PCH) : d.np;
    for (int pc = pbeg; pc < pend; pc += PCH) {
      const int npc = min(PCH, d.np - pc);
      const int ngt = (npc + 15) >> 4;       // groups of 16 points in this chunk
      const int n0 = H == 2 ? (ngt + 1) >> 1 : ngt;  // the first half takes the extra group
      const int ng = half ? ngt - n0 : n0, gp0 = half ? n0 : 0;
      ozaki_dispatch<S, NPG, H, GRAD, ASYNC>(std::make_integer_sequence<int, NPG + 1>{}, ng, Bs, jbs, red, d, t, pc, npc,
                                      gp0, I0, I1, tpart, sib, map, adig + nn, ai, Bg, binv, n_basis, bias, rho, n_full, nmb,
                                      n_batches_all, n_tile, ga);
    }
  }
}

// Order 3: per pass record, the row tiles' partial sums added up in row-tile order (as the block
// adds them in the other orders) and the final values written.
template <bool GRAD>
__global__ __launch_bounds__(256) void rho1_ozaki_finish_kernel(const PassRec *__restrict__ recs,
                                                                const BatchDesc *__restrict__ descs, int pch,
                                                                const double *__restrict__ part,
                                                                long long part_stride, const int *__restrict__ map,
                                                                const double *__restrict__ bias,
                                                                double *__restrict__ rho, int n_full, int nmb,
                                                                int n_batches_all, int n_tile, GradArgs ga) {
  constexpr int C = GRAD ? 4 : 1;
  const PassRec r = recs[blockIdx.x];
  const int t = blockIdx.y;
  const BatchDesc d = descs[r.di];
  const double *pp = part + part_stride * t + r.poff;
  for (int e = threadIdx.x; e < C * r.npc; e += blockDim.x) {
    const int c = e / r.npc, p = e % r.npc;
    double tot = 0.0;
    for (int I = 0; I < r.nT; I++)
      tot += pp[((size_t)I * C + c) * r.npc + p];
    ozaki_final<GRAD>(c, tot, d, t, pch * r.q + p, map, bias, rho, n_full, nmb, n_batches_all, n_tile, ga);
  }
}

template <int S>
void prepare(Rho1Plan &P, OzakiState &st, const double *wave, hipStream_t s) {
  std::vector<long long> off(P.n_valid + 1, 0);
  for (int i = 0; i < P.n_valid; i++) {
    const BatchDesc &d = P.h_desc[i];
    off[i + 1] = off[i] + (long long)S * ((d.nc + 63) / 64 * 64) * d.np;
  }
  st.d_boff.upload(off);
  st.d_bdig.alloc(std::max<long long>(off[P.n_valid], 16));
  int nvp = 0;
  for (const auto &d : P.h_desc)
    nvp = std::max(nvp, d.pstart + d.np);
  st.n_valid_points = nvp;
  st.d_binv.alloc(std::max(nvp, 1));
  const Relabel &R = get_relabel(P);
  hipLaunchKernelGGL(phi_digits_kernel<S>, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, st.d_boff.ptr,
                     R.d_srow.ptr, wave, st.d_bdig.ptr, st.d_binv.ptr);
  check_launch("phi_digits_kernel");
  st.S = S;
  st.wave = wave;
  st.bytes = off[P.n_valid] + (long long)nvp * 8;
}

template <int S, int NPG, int H, bool GRAD, bool ASYNC = false>
void launch(Rho1Plan &P, const Rho1Call &c, const LaunchCfg &cfg, hipStream_t s) {
  OzakiState &st = *P.ozaki;
  const Relabel &R = get_relabel(P);
  static int per_cu = -1;
  if (per_cu < 0) {
    FRPA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rho1_ozaki_kernel<S, NPG, H, GRAD, ASYNC>,
                                                          256 * H, 0));
    per_cu = std::max(per_cu, 1);
  }
  constexpr int PCH = 16 * H * NPG;
  const bool items = cfg.order == 2;
  TileSched *ts = cfg.order == 3 ? &build_tile_items(P, PCH, GRAD) : nullptr;
  if (items)
    build_item_order(P, PCH);
  const int n_pos = ts ? ts->n_items : items ? P.n_items : P.n_valid;
  if (ts)
    P.d_tile_part.grow((size_t)ts->part_doubles * c.T);
  LaunchCfg lc = cfg;
  if (ts)
    lc.order = 0;  // the item list is already in launch-position order
  const LaunchArgs la = launch_args(P, lc, c.T, per_cu, s, n_pos);
  const GradArgs ga{c.gwave, c.gbias, c.grad, R.d_srow.ptr};
  hipLaunchKernelGGL((rho1_ozaki_kernel<S, NPG, H, GRAD, ASYNC>), la.grid, dim3(256 * H), 0, s, P.d_desc.ptr, la.order,
                     la.counters, n_pos, c.T, R.d_sib.ptr, P.d_map.ptr, st.d_adig.ptr, st.d_ainv.ptr,
                     st.d_bdig.ptr, st.d_boff.ptr, st.d_binv.ptr, P.n_basis, c.bias, c.rho, P.n_full, P.nmb,
                     P.n_batches, c.n_tile, (int)items, ts ? ts->d_items.ptr : nullptr,
                     ts ? P.d_tile_part.ptr : nullptr, ts ? ts->part_doubles : 0LL, ga);
  check_launch("rho1_ozaki_kernel");
  if (ts) {
    hipLaunchKernelGGL((rho1_ozaki_finish_kernel<GRAD>), dim3(ts->n_recs, c.T), dim3(256), 0, s, ts->d_recs.ptr,
                       P.d_desc.ptr, PCH, P.d_tile_part.ptr, ts->part_doubles, P.d_map.ptr, c.bias, c.rho, P.n_full,
                       P.nmb, P.n_batches, c.n_tile, ga);
    check_launch("rho1_ozaki_finish_kernel");
  }
}

template <int S, bool GRAD>
void forward(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  OzakiState &st = *P.ozaki;
  const Relabel &R = get_relabel(P);
  const int n = P.n_basis;
  const size_t nbb = (size_t)n * n * c.T;
  const size_t nr = (size_t)n * c.T;
  st.d_adig.grow(nbb);
  st.d_ainv.grow(nr);
  st.d_rowbits.grow(nr);
  fill_f64(reinterpret_cast<double *>(st.d_rowbits.ptr), nr, 0.0, s);  // +0.0 has all-zero bits
  const int chunk = 256;
  hipLaunchKernelGGL(p1_rowmax_kernel, dim3((n + 255) / 256, (n + chunk - 1) / chunk, c.T), dim3(256), 0, s, c.P1,
                     n, chunk, st.d_rowbits.ptr);
  check_launch("p1_rowmax_kernel");
  hipLaunchKernelGGL(p1_colmax_kernel, dim3(n, c.T), dim3(256), 0, s, c.P1, n, st.d_rowbits.ptr);
  check_launch("p1_colmax_kernel");
  hipLaunchKernelGGL(p1_rowscale_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, st.d_rowbits.ptr,
                     R.d_pinv.ptr, n, nr, st.d_ainv.ptr);
  check_launch("p1_rowscale_kernel");
  hipLaunchKernelGGL((p1_digits_kernel<S, GRAD>), dim3((n + 63) / 64, (n + 63) / 64, c.T), dim3(256), 0, s, c.P1,
                     R.d_pinv.ptr, st.d_ainv.ptr, n, st.d_adig.ptr);
  check_launch("p1_digits_kernel");
  // point-chunk width 16 H NPG: NPG at most the widest whose accumulators, prefetched digits and
  // staged B fit 256 registers per lane (two waves per SIMD) without spilling: 4 groups for 5 digits,
  // 5 for 4 (the default when some batch has more than 128 points).  cfg.npg = 2 selects the
  // four-wave block (H = 1, 64-point passes, two blocks per CU); cfg.npg = 1 the same block with its
  // phi chunks staged by direct global -> LDS loads and the P' gathers one step ahead.
  const LaunchCfg &cfg = P.cfg[S == 5 ? FRPA_PREC_OZAKI : FRPA_PREC_OZAKI4];
  // (the buffer-resource gather needs the n_basis^2 elements of one perturbation under 4 GiB)
  if (cfg.npg == 1 && (unsigned long long)n * n * 8 < (1ull << 32)) {
    launch<S, 4, 1, GRAD, true>(P, c, cfg, s);
  } else if (cfg.npg == 2 || cfg.npg == 1) {
    launch<S, 4, 1, GRAD>(P, c, cfg, s);
  } else if constexpr (S == 5) {
    launch<S, 4, 2, GRAD>(P, c, cfg, s);
  } else {
    const int npg = cfg.npg ? cfg.npg : (P.max_np <= 128 ? 4 : 5);
    if (npg == 4)
      launch<S, 4, 2, GRAD>(P, c, cfg, s);
    else
      launch<S, 5, 2, GRAD>(P, c, cfg, s);
  }
}

} // namespace

long long ozaki_prepare_wave(Rho1Plan &P, const double *wave, int S, hipStream_t s) {
  FRPA_REQUIRE(S == 4 || S == 5, FRPA_ERR_INVALID, "INT8 split supports 4 or 5 slices, not %d", S);
  if (!P.ozaki)
    P.ozaki = new OzakiState();
  if (P.n_valid == 0)
    return 0;
  if (S == 4)
    prepare<4>(P, *P.ozaki, wave, s);
  else
    prepare<5>(P, *P.ozaki, wave, s);
  return P.ozaki->bytes;
}

void rho1_forward_ozaki(Rho1Plan &P, const Rho1Call &c, int S, hipStream_t s) {
  if (P.n_valid == 0)
    return;
  if (!P.ozaki || P.ozaki->wave != c.wave || P.ozaki->S != S)
    ozaki_prepare_wave(P, c.wave, S, s);
  const bool grad = c.grad != nullptr;
  FRPA_REQUIRE(!grad || c.gwave, FRPA_ERR_INVALID, "gradient output without gradient_wave");
  if (P.ozaki->S == 4)
    grad ? forward<4, true>(P, c, s) : forward<4, false>(P, c, s);
  else
    grad ? forward<5, true>(P, c, s) : forward<5, false>(P, c, s);
}

} // namespace frpa
