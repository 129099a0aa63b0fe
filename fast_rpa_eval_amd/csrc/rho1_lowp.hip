// rho1 forward on the symmetric half of P1 in FP64 and in reduced precision -- the same-algorithm
// FP64 comparison of the INT8 split and its reduced-precision comparison paths:
//   FRPA_PREC_FP64_SYM: FP64 operands as given, v_mfma_f64_16x16x4_f64 (FP64 accumulation);
//   FRPA_PREC_FP32    : operands rounded to FP32, v_mfma_f32_16x16x4_f32;
//   FRPA_PREC_BF16X3  : each operand split x = hi + lo with hi = bf16(x), lo = bf16(x - hi), three
//                       products hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16.
// The FP32 MFMA accumulators hold one K step (64 compute rows) and are added into FP64 registers
// after every step (compensation at K-step granularity); the ddot rho(p) = sum_i phi(i, p) W(i, p),
// the bias and the scatter are FP64, with phi(i, p) the converted operand (FP32, or hi + lo; FP64:
// the wave itself).
// Reference: evaluate_first_order_rho.cpp:149-207 (the same per-batch GEMM + ddot).
//
// Data path of the INT8 split's direct-to-LDS launch (rho1_ozaki.hip), with its symmetric form
// rho(p) = 2 sum_i phi_i(p) sum_{k >= i} P'(i, k) phi_k(p), P' the symmetric part of P1 with the
// diagonal halved, tile pairs J >= I of the batch's label-sorted compute rows:
//   * phi is converted once per wave into 4 bytes per element -- FP32, or a bf16 hi plane and a bf16
//     lo plane -- per batch (nc rounded up to 64) rows in sorted order x np points, rows contiguous;
//   * P' is converted per call into one 4-byte word per element (FP32, or bf16 hi | lo << 16) in
//     the relabelled (n_basis, n_basis, T) layout and gathered as the MFMA A operand through a
//     buffer resource, the next K step's gathers in flight during the current step's MFMAs;
//   * FP64: phi is staged from wave_batches_compress itself (the caller's row order, K steps of 32
//     rows: four sub-rows of 8 doubles per point; rows at or past n_compute are replaced by zero
//     when read, padding rows past nc_pad by row 0 when staged), P' (FP64) per call in the caller's
//     basis numbering, gathered at (i_basis(i), i_basis(k));
//   * a K step's phi chunk (64 rows x 64 points, 16 KB) goes straight from global memory into LDS
//     (global_load_lds_dwordx4) as four 64-byte sub-rows per point: FP32 rows 16 q .. 16 q + 15
//     (sub-row q), BF16x3 hi rows 32 h .. + 31 (sub-row h) and lo rows (sub-row 2 + h); a lane
//     (m = lane % 16, kq = lane / 16) of MFMA q (FP32) holds k = 16 kq + q, of MFMA h (BF16x3) the
//     eight k = 16 kq + 8 h .. + 7 -- the same K order for both operands, so the sums are those of
//     the GEMM.
#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

#include "rho1_common.hpp"

namespace frpa {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

// The converted wave of one precision (FP32 or BF16x3), each kept while the same wave pointer is passed.
struct LowpConv {
  const double *wave = nullptr;  // the wave the phi planes were made from
  long long bytes = 0;
  DevBuf<long long> d_boff;      // per BatchDesc: byte offset of its phi planes
  DevBuf<unsigned char> d_bdat;  // phi planes (4 bytes per element)
};
struct LowpState {
  LowpConv conv[2];              // [0] FP32, [1] BF16x3
  DevBuf<unsigned> d_a;          // per call: P' words in the new labels (FP64: caller's numbering), (n, n, T)
};

void lowp_release(LowpState *st) { delete st; }

namespace {

constexpr int LP_PCH = 64;              // points per pass
constexpr int LP_BUF = 4 * LP_PCH * 64;  // bytes of one staged K step: 4 sub-rows of 64 bytes per point

__device__ __forceinline__ int ceil64(int x) { return (x + 63) & ~63; }

__device__ __forceinline__ unsigned short bf16_bits(__bf16 h) { return __builtin_bit_cast(unsigned short, h); }

// x = hi + lo, hi = bf16(x) and lo = bf16(x - hi) (round to nearest), as (hi bits, lo bits)
__device__ __forceinline__ void bf16_split(double x, unsigned short &hi, unsigned short &lo) {
  const __bf16 h = (__bf16)(float)x;
  const __bf16 l = (__bf16)(float)(x - (double)(float)h);
  hi = bf16_bits(h);
  lo = bf16_bits(l);
}

// ---- phi planes (once per wave) ---------------------------------------------------------------
// One wave per point, lanes down the batch's sorted rows (row k is the caller's local row srow[k]).
template <int PREC>
__global__ __launch_bounds__(256) void phi_lowp_kernel(const BatchDesc *__restrict__ descs,
                                                       const long long *__restrict__ boff,
                                                       const int *__restrict__ srow_all,
                                                       const double *__restrict__ wave,
                                                       unsigned char *__restrict__ bdat) {
  const BatchDesc d = descs[blockIdx.x];
  const int ld = (d.nc + 3) & ~3, ncK = ceil64(d.nc);
  const double *W = wave + d.wave_off;
  const int *srow = srow_all + d.ib_off;
  unsigned char *out = bdat + boff[blockIdx.x];
  const size_t plane = (size_t)ncK * d.np;
  const int lane = threadIdx.x & 63;
  for (int p = threadIdx.x >> 6; p < d.np; p += 4) {
    const double *Wp = W + (size_t)ld * p;
    for (int k = lane; k < ncK; k += 64) {
      const double x = k < d.nc ? Wp[srow[k]] : 0.0;
      if constexpr (PREC == FRPA_PREC_FP32) {
        reinterpret_cast<float *>(out)[(size_t)ncK * p + k] = (float)x;
      } else {
        unsigned short hi, lo;
        bf16_split(x, hi, lo);
        unsigned short *o = reinterpret_cast<unsigned short *>(out);
        o[(size_t)ncK * p + k] = hi;
        o[plane + (size_t)ncK * p + k] = lo;
      }
    }
  }
}

// ---- P' words (per call) ------------------------------------------------------------------------
// P'(pinv[x], pinv[y]) at x + n y: one 64 x 64 tile of the new labels per block, the transposed
// tile P1(pinv[y], pinv[x]) through LDS so both reads run down columns.
// FP64_SYM: the FP64 words (two per element) in the caller's numbering (pinv = nullptr: identity).
template <int PREC>
__global__ __launch_bounds__(256) void p1_lowp_kernel(const double *__restrict__ P1, const int *__restrict__ pinv,
                                                      int n_basis, unsigned *__restrict__ out) {
  __shared__ double tr[64][65];
  const int x0 = 64 * blockIdx.x, y0 = 64 * blockIdx.y, t = blockIdx.z;
  const size_t nn = (size_t)n_basis * n_basis;
  const double *Pt = P1 + nn * t;
  const int a = threadIdx.x & 63, b0 = threadIdx.x >> 6;
  auto lab = [&](int z) { return pinv ? pinv[z] : z; };
  {
    const int y = y0 + a;
    const int cy = y < n_basis ? lab(y) : 0;
#pragma unroll 4
    for (int u = 0; u < 16; u++) {
      const int i = b0 + 4 * u, x = x0 + i;
      tr[i][a] = (x < n_basis && y < n_basis) ? Pt[cy + (size_t)n_basis * lab(x)] : 0.0;
    }
  }
  __syncthreads();
  const int x = x0 + a;
  if (x >= n_basis)
    return;
  const int rx = lab(x);
#pragma unroll 4
  for (int u = 0; u < 16; u++) {
    const int j = b0 + 4 * u, y = y0 + j;
    if (y >= n_basis)
      break;
    double v = 0.5 * (Pt[rx + (size_t)n_basis * lab(y)] + tr[a][j]);
    if (x == y)
      v *= 0.5;
    if constexpr (PREC == FRPA_PREC_FP64_SYM) {
      reinterpret_cast<double *>(out)[x + (size_t)n_basis * y + nn * t] = v;
    } else {
      unsigned wv;
      if constexpr (PREC == FRPA_PREC_FP32) {
        wv = __float_as_uint((float)v);
      } else {
        unsigned short hi, lo;
        bf16_split(v, hi, lo);
        wv = (unsigned)hi | ((unsigned)lo << 16);
      }
      out[x + (size_t)n_basis * y + nn * t] = wv;
    }
  }
}

// ---- fused forward ------------------------------------------------------------------------------
// Block: 4 waves, wave rg owns rows 16 rg .. + 15 of each 64-row tile I, all 64 points of the pass
// (4 groups of 16).  Three LDS buffers as in the INT8 split: buffer 2 holds the first chunk of the
// row tile (FP32 / BF16x3: the diagonal chunk J = I, whose values also give phi(i, p) for the ddot),
// buffers 0 and 1 alternate for the later chunks.  Row tiles [I0, I1) of one pass; per-tile sums
// combined across the row groups in LDS, ((g0 + g1) + g2) + g3, then added up over the row tiles in
// order here (part == nullptr) or by the finishing kernel (order 3), so every launch order gives the
// same bits.
template <int PREC>
__device__ __forceinline__ void lowp_chunk(unsigned char *Bs, unsigned *jco, double *red, const BatchDesc &d, int t,
                                           int pc, int npc, int I0, int I1, double *__restrict__ part,
                                           const int *__restrict__ sib, const int *__restrict__ map,
                                           const unsigned *__restrict__ At, const unsigned char *__restrict__ Bg,
                                           const double *__restrict__ W, int n_basis, const double *__restrict__ bias,
                                           double *__restrict__ rho, int n_full, int nmb, int n_batches_all,
                                           int n_tile) {
  constexpr bool FP32 = PREC == FRPA_PREC_FP32;
  constexpr bool F64 = PREC == FRPA_PREC_FP64_SYM;
  constexpr int KC = F64 ? 32 : 64;  // compute rows per K step
  constexpr int ESZ = F64 ? 8 : 4;   // bytes per gathered P' element
  constexpr int NQ = F64 ? 8 : 16;   // P' elements a lane gathers per K step
  const int nc = d.nc, ncK = ceil64(nc), ld = (nc + 3) & ~3;
  const size_t plane2 = (size_t)2 * ncK * d.np;  // bytes of the BF16x3 hi plane
  const int tid = threadIdx.x, lane = tid & 63, rg = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nC = (nc + KC - 1) / KC;  // K steps over the batch's rows
  double tot = 0.0;  // the point of reducing thread tid (tid < npc)
  auto reduce_tile = [&](int I) {
    if (tid < npc) {
      const double *r = red + tid;
      const double v = ((r[0] + r[LP_PCH]) + r[2 * LP_PCH]) + r[3 * LP_PCH];
      if (part)
        part[(size_t)I * npc + tid] = v;
      else
        tot += v;
    }
  };
  // chunk c straight into buffer buf: 1 KB block k = 64 pieces of 16 bytes by wave k mod 4; piece e:
  // 16-byte part q = e & 3 of sub-row sl = e / (4 PCH) of point p; points past npc read the last
  // point's values (their sums are never used); FP64: sub-row sl = rows KC c + 8 sl .. + 7 of the
  // wave block, a piece past nc_pad reads row 0 of the point
  auto stage_b = [&](int c, unsigned char *B) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int blk = 4 * u + rg;
      const int e = 64 * blk + lane;
      const int q = e & 3, p = min((e >> 2) % LP_PCH, npc - 1), sl = (e >> 2) / LP_PCH;
      const void *src;
      if constexpr (F64) {
        int row = KC * c + 8 * sl + 2 * q;
        row = row < ld ? row : 0;
        src = W + (size_t)ld * (pc + p) + row;
      } else {
        const size_t col = (size_t)ncK * (pc + p) + 64 * c;
        src = FP32 ? Bg + 4 * (col + 16 * sl) + 16 * q
                   : Bg + (sl >> 1) * plane2 + 2 * (col + 32 * (sl & 1)) + 16 * q;
      }
      __builtin_amdgcn_global_load_lds(src, (lds_void *)(B + 1024 * blk), 16, 0, 0);
    }
  };
  const unsigned nbe = (unsigned)ESZ * (unsigned)n_basis;
  auto col_off = [&](int j) { return j < nc ? nbe * (unsigned)sib[j] : 0u; };
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)At, 0, (int)(nbe * (unsigned)n_basis), 0x00020000);

  for (int I = I0; I < I1; I++) {
    const int irow = 64 * I + 16 * rg + lr;
    const bool iok = irow < nc;
    const unsigned arowB = (unsigned)ESZ * (unsigned)(iok ? sib[irow] : 0);
    double accd[4][4];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        accd[g][r] = 0.0;
    unsigned el[16];
    v2u el64[8];
    // P' words of row irow, columns k = KC c + (KC / 4) lk + q (labels from jco, buffer buf); columns
    // past nc read column 0 and meet zero phi rows; rows past nc read row 0 and are never summed
    auto gather = [&](int buf) {
      const uint4 *cp = reinterpret_cast<const uint4 *>(jco + 64 * buf + NQ * lk);
#pragma unroll
      for (int h = 0; h < NQ / 4; h++) {
        const uint4 c4 = cp[h];
        if constexpr (F64) {
          el64[4 * h] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.x, 0, 0);
          el64[4 * h + 1] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.y, 0, 0);
          el64[4 * h + 2] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.z, 0, 0);
          el64[4 * h + 3] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, arowB + c4.w, 0, 0);
        } else {
          el[4 * h] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, arowB + c4.x, 0, 0);
          el[4 * h + 1] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, arowB + c4.y, 0, 0);
          el[4 * h + 2] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, arowB + c4.z, 0, 0);
          el[4 * h + 3] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, arowB + c4.w, 0, 0);
        }
      }
    };
    // A operands of chunk c, P'(i, k) = 0 below the diagonal (k < i): FP64 a64[q] / FP32 a32[q] =
    // element q; BF16x3 MFMA h takes elements 8 h .. 8 h + 7 (hi and lo)
    double a64[8];
    float a32[16];
    v8bf ahi[2], alo[2];
    auto repack = [&](int c) {
      if constexpr (F64) {
#pragma unroll
        for (int q = 0; q < 8; q++)
          a64[q] = KC * c + 8 * lk + q < irow ? 0.0 : __builtin_bit_cast(double, el64[q]);
      } else {
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (64 * c + 16 * lk + q < irow)
            el[q] = 0u;
        if constexpr (FP32) {
#pragma unroll
          for (int q = 0; q < 16; q++)
            a32[q] = __uint_as_float(el[q]);
        } else {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            v4u hw, lw;
#pragma unroll
            for (int m = 0; m < 4; m++) {
              const unsigned x0 = el[8 * h + 2 * m], x1 = el[8 * h + 2 * m + 1];
              hw[m] = __builtin_amdgcn_perm(x1, x0, 0x05040100u);  // low halves: hi of both
              lw[m] = __builtin_amdgcn_perm(x1, x0, 0x07060302u);  // high halves: lo of both
            }
            ahi[h] = __builtin_bit_cast(v8bf, hw);
            alo[h] = __builtin_bit_cast(v8bf, lw);
          }
        }
      }
    };
    block_sync();  // the previous tile is done with all buffers, its row-group sums are in LDS
    if (I > I0)
      reduce_tile(I - 1);
    // the row tile's K steps: chunks c0 .. nC - 1, c0 the one holding row 64 I
    const int c0 = 64 * I / KC;
    const int nsteps = nC - c0;
    // The staging buffers reach the K steps as distinct restrict pointers, each step's buffers
    // fixed at compile time (step 0 reads the diagonal chunk from B2, then B0 and B1 alternate, two
    // steps per iteration), so the compiler can tell that a step's B reads do not touch the buffer
    // whose direct loads are still in flight and leaves the P' gathers in flight across the MFMAs.
    auto ksteps = [&](unsigned char *__restrict__ B0, unsigned char *__restrict__ B1, unsigned char *__restrict__ B2) {
    stage_b(c0, B2);
    if (nsteps > 1)
      stage_b(c0 + 1, B0);
    if (tid < 64) {
      jco[128 + tid] = col_off(KC * c0 + tid);
      if (nsteps > 1)
        jco[tid] = col_off(KC * (c0 + 1) + tid);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    gather(2);
    repack(c0);
    // step st: chunk c0 + st in B, P' column offsets of step st + 1 in jco slot st & 1; chunk
    // c0 + st + 2 is staged into Bnext once every wave has read B
    auto step = [&](int st, const unsigned char *B, unsigned char *Bnext) {
      const int c = c0 + st;
      const bool hn = st + 1 < nsteps, hnn = st + 2 < nsteps;
      const int bn = st & 1, bnn = (st + 1) & 1;
      // labels of chunk c + 2 (index clamped into the batch; used only if hnn), loaded before the
      // gathers and consumed after the MFMAs
      const int jn = KC * (c + 2) + lane;
      const int sj = sib[min(jn, nc - 1)];
      __builtin_amdgcn_sched_barrier(0);
      if (hn)
        gather(bn);  // in flight during this step's MFMAs
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int p = 16 * g + lr;
        if constexpr (F64) {
          // sub-row lk of point p: rows KC c + 8 lk .. + 7; rows at or past nc read as zero
          const double *row = reinterpret_cast<const double *>(B + (lk * LP_PCH + p) * 64);
          v4d acc = v4d{accd[g][0], accd[g][1], accd[g][2], accd[g][3]};
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const double b = KC * c + 8 * lk + q < nc ? row[q] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a64[q], b, acc, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 4; r++)
            accd[g][r] = acc[r];
        } else {
          v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
          if constexpr (FP32) {
            // sub-row lk of point p: rows 16 lk .. + 15
            const v4u *row = reinterpret_cast<const v4u *>(B + (lk * LP_PCH + p) * 64);
#pragma unroll
            for (int h = 0; h < 4; h++) {
              const v4u b4 = row[h];
#pragma unroll
              for (int u = 0; u < 4; u++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a32[4 * h + u], __uint_as_float(b4[u]), acc, 0, 0, 0);
            }
          } else {
            // rows 16 lk + 8 h .. + 7: hi in sub-row lk / 2, lo in sub-row 2 + lk / 2, at byte
            // 32 (lk % 2) + 16 h of the point's 64-byte row
            const unsigned char *rh = B + ((lk >> 1) * LP_PCH + p) * 64 + 32 * (lk & 1);
            const unsigned char *rl = B + ((2 + (lk >> 1)) * LP_PCH + p) * 64 + 32 * (lk & 1);
#pragma unroll
            for (int h = 0; h < 2; h++) {
              const v8bf bh = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u *>(rh + 16 * h));
              const v8bf bl = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u *>(rl + 16 * h));
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[h], bh, acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[h], bl, acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[h], bh, acc, 0, 0, 0);
            }
          }
#pragma unroll
          for (int r = 0; r < 4; r++)
            accd[g][r] += (double)acc[r];
        }
      }
      if (hnn && tid < 64)
        jco[64 * bnn + tid] = jn < nc ? nbe * (unsigned)sj : 0u;
      __builtin_amdgcn_s_waitcnt(0);  // chunk c + 1 staged, its P' gathers arrived
      __syncthreads();
      if (hnn)
        stage_b(c + 2, Bnext);
      if (hn)
        repack(c + 1);
    };
    step(0, B2, B1);
    for (int st = 1; st < nsteps; st += 2) {
      step(st, B0, B0);
      if (st + 1 < nsteps)
        step(st + 1, B1, B1);
    }
    };
    ksteps(Bs, Bs + LP_BUF, Bs + 2 * LP_BUF);

    // epilogue: the FP32 / BF16 MFMAs leave D(4 (lane / 16) + r, lane % 16) in acc[r], the FP64 one
    // D(4 r + lane / 16, lane % 16), so accd[g][r] = W(i, p = pc + 16 g + lr) with i = 64 I + 16 rg +
    // (4 lk + r, FP64: 4 r + lk); phi(i, p) from the diagonal chunk (local row 16 rg + 4 lk + r), FP64:
    // from the wave
    const unsigned char *BD = Bs + 2 * LP_BUF;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int p = 16 * g + lr;
      double part_p = 0.0;
      if (p < npc) {
        double ph[4];
        if constexpr (F64) {
          const double *Wp = W + (size_t)ld * (pc + p) + 64 * I + 16 * rg + lk;
#pragma unroll
          for (int r = 0; r < 4; r++)
            ph[r] = 64 * I + 16 * rg + 4 * r + lk < nc ? Wp[4 * r] : 0.0;
        } else if constexpr (FP32) {
          const v4u f = *reinterpret_cast<const v4u *>(BD + (rg * LP_PCH + p) * 64 + 16 * lk);
#pragma unroll
          for (int r = 0; r < 4; r++)
            ph[r] = (double)__uint_as_float(f[r]);
        } else {
          const int off = 32 * (rg & 1) + 8 * lk;
          const v2u hv = *reinterpret_cast<const v2u *>(BD + ((rg >> 1) * LP_PCH + p) * 64 + off);
          const v2u lv = *reinterpret_cast<const v2u *>(BD + ((2 + (rg >> 1)) * LP_PCH + p) * 64 + off);
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const unsigned sh = 16 * (r & 1);
            const float hi = __uint_as_float(((hv[r >> 1] >> sh) & 0xffffu) << 16);
            const float lo = __uint_as_float(((lv[r >> 1] >> sh) & 0xffffu) << 16);
            ph[r] = (double)hi + (double)lo;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (64 * I + 16 * rg + (F64 ? 4 * r + lk : 4 * lk + r) < nc)
            part_p += accd[g][r] * ph[r];
      }
      part_p += __shfl_xor(part_p, 16);
      part_p += __shfl_xor(part_p, 32);
      if (lk == 0)
        red[rg * LP_PCH + p] = part_p;
    }
  }
  block_sync();
  if (I1 > I0)
    reduce_tile(I1 - 1);
  if (!part && tid < npc) {
    // rho1 = 2 sum (symmetric form) + bias (evaluate_first_order_rho.cpp:187-207)
    const int p = pc + tid;
    double sum = 2.0 * tot;
    if (bias)
      sum += bias[(size_t)p + (size_t)nmb * (d.b + (size_t)n_batches_all * t)];
    if (t < n_tile)
      rho[(size_t)map[d.pstart + p] + (size_t)n_full * t] = sum;
  }
}

// sib_all: per batch the P' index of each compute row -- FP32 / BF16x3: its new label (sorted rows),
// FP64: its 0-based basis index (the caller's rows)
template <int PREC>
__global__ __launch_bounds__(256, 3) void rho1_lowp_kernel(
    const BatchDesc *__restrict__ descs, const int *__restrict__ order, int *__restrict__ counters, int n_pos, int T,
    const int *__restrict__ sib_all, const int *__restrict__ map, const unsigned *__restrict__ adat,
    const unsigned char *__restrict__ bdat, const long long *__restrict__ boff, const double *__restrict__ wave,
    int n_basis, const double *__restrict__ bias, double *__restrict__ rho, int n_full, int nmb, int n_batches_all,
    int n_tile, int items, const TileItem *__restrict__ titems, double *__restrict__ part, long long part_stride) {
  constexpr size_t AW = PREC == FRPA_PREC_FP64_SYM ? 2 : 1;  // 4-byte words per P' element
  __shared__ __attribute__((aligned(16))) unsigned char Bs[3 * LP_BUF];
  __shared__ __attribute__((aligned(16))) unsigned jco[192];
  __shared__ double red[4 * LP_PCH];
  __shared__ int s_item;
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
    const int pbeg = titems ? LP_PCH * ti.q : items ? LP_PCH * (it & 255) : 0;
    const int pend = (titems || items) ? min(d.np, pbeg + LP_PCH) : d.np;
    for (int pc = pbeg; pc < pend; pc += LP_PCH)
      lowp_chunk<PREC>(Bs, jco, red, d, t, pc, min(LP_PCH, d.np - pc), I0, I1, tpart, sib_all + d.ib_off, map,
                       adat + AW * n_basis * n_basis * t, bdat ? bdat + boff[di] : nullptr,
                       wave + d.wave_off, n_basis, bias, rho, n_full, nmb, n_batches_all, n_tile);
  }
}

// Order 3: per pass record, the row tiles' partial sums added up in row-tile order and the final
// values written (rho1 = 2 sum + bias).
__global__ __launch_bounds__(256) void rho1_lowp_finish_kernel(const PassRec *__restrict__ recs,
                                                               const BatchDesc *__restrict__ descs,
                                                               const double *__restrict__ part, long long part_stride,
                                                               const int *__restrict__ map,
                                                               const double *__restrict__ bias,
                                                               double *__restrict__ rho, int n_full, int nmb,
                                                               int n_batches_all, int n_tile) {
  const PassRec r = recs[blockIdx.x];
  const int t = blockIdx.y;
  const BatchDesc d = descs[r.di];
  const double *pp = part + part_stride * t + r.poff;
  for (int e = threadIdx.x; e < r.npc; e += blockDim.x) {
    double tot = 0.0;
    for (int I = 0; I < r.nT; I++)
      tot += pp[(size_t)I * r.npc + e];
    const int p = LP_PCH * r.q + e;
    double sum = 2.0 * tot;
    if (bias)
      sum += bias[(size_t)p + (size_t)nmb * (d.b + (size_t)n_batches_all * t)];
    if (t < n_tile)
      rho[(size_t)map[d.pstart + p] + (size_t)n_full * t] = sum;
  }
}

constexpr int conv_index(int prec) { return prec == FRPA_PREC_FP32 ? 0 : 1; }

template <int PREC>
void prepare(Rho1Plan &P, LowpConv &st, const double *wave, hipStream_t s) {
  std::vector<long long> off(P.n_valid + 1, 0);
  for (int i = 0; i < P.n_valid; i++) {
    const BatchDesc &d = P.h_desc[i];
    off[i + 1] = off[i] + 4LL * ((d.nc + 63) / 64 * 64) * d.np;
  }
  st.d_boff.upload(off);
  st.d_bdat.alloc(std::max<long long>(off[P.n_valid], 16));
  const Relabel &R = get_relabel(P);
  hipLaunchKernelGGL(phi_lowp_kernel<PREC>, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, st.d_boff.ptr,
                     R.d_srow.ptr, wave, st.d_bdat.ptr);
  check_launch("phi_lowp_kernel");
  st.wave = wave;
  st.bytes = off[P.n_valid];
}

template <int PREC>
void forward(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  constexpr bool F64 = PREC == FRPA_PREC_FP64_SYM;
  LowpState &st = *P.lowp;
  const Relabel &R = get_relabel(P);
  const int n = P.n_basis;
  st.d_a.grow((size_t)n * n * c.T * (F64 ? 2 : 1));
  // FP64: P' in the caller's numbering, gathered at the batches' own basis indices
  hipLaunchKernelGGL(p1_lowp_kernel<PREC>, dim3((n + 63) / 64, (n + 63) / 64, c.T), dim3(256), 0, s, c.P1,
                     F64 ? nullptr : R.d_pinv.ptr, n, st.d_a.ptr);
  check_launch("p1_lowp_kernel");
  static int per_cu = -1;
  if (per_cu < 0) {
    FRPA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rho1_lowp_kernel<PREC>, 256, 0));
    per_cu = std::max(per_cu, 1);
  }
  const LaunchCfg &cfg = P.cfg[PREC];
  const bool items = cfg.order == 2;
  TileSched *ts = cfg.order == 3 ? &build_tile_items(P, LP_PCH, false) : nullptr;
  if (items)
    build_item_order(P, LP_PCH);
  const int n_pos = ts ? ts->n_items : items ? P.n_items : P.n_valid;
  if (ts)
    P.d_tile_part.grow((size_t)ts->part_doubles * c.T);
  LaunchCfg lc = cfg;
  if (ts)
    lc.order = 0;  // the item list is already in launch-position order
  const LaunchArgs la = launch_args(P, lc, c.T, per_cu, s, n_pos);
  const LowpConv &cv = st.conv[conv_index(PREC)];
  hipLaunchKernelGGL(rho1_lowp_kernel<PREC>, la.grid, dim3(256), 0, s, P.d_desc.ptr, la.order, la.counters, n_pos,
                     c.T, F64 ? P.d_ibasis0.ptr : R.d_sib.ptr, P.d_map.ptr, st.d_a.ptr,
                     F64 ? nullptr : cv.d_bdat.ptr, cv.d_boff.ptr, c.wave, n, c.bias, c.rho, P.n_full, P.nmb,
                     P.n_batches, c.n_tile, (int)items, ts ? ts->d_items.ptr : nullptr,
                     ts ? P.d_tile_part.ptr : nullptr, ts ? ts->part_doubles : 0LL);
  check_launch("rho1_lowp_kernel");
  if (ts) {
    hipLaunchKernelGGL(rho1_lowp_finish_kernel, dim3(ts->n_recs, c.T), dim3(256), 0, s, ts->d_recs.ptr,
                       P.d_desc.ptr, P.d_tile_part.ptr, ts->part_doubles, P.d_map.ptr, c.bias, c.rho, P.n_full,
                       P.nmb, P.n_batches, c.n_tile);
    check_launch("rho1_lowp_finish_kernel");
  }
}

} // namespace

long long lowp_prepare_wave(Rho1Plan &P, const double *wave, int precision, hipStream_t s) {
  FRPA_REQUIRE(precision == FRPA_PREC_FP32 || precision == FRPA_PREC_BF16X3, FRPA_ERR_INVALID,
               "not a reduced-precision rho1 path: %d", precision);
  if (!P.lowp)
    P.lowp = new LowpState();
  if (P.n_valid == 0)
    return 0;
  LowpConv &cv = P.lowp->conv[conv_index(precision)];
  if (precision == FRPA_PREC_FP32)
    prepare<FRPA_PREC_FP32>(P, cv, wave, s);
  else
    prepare<FRPA_PREC_BF16X3>(P, cv, wave, s);
  return cv.bytes;
}

void rho1_forward_lowp(Rho1Plan &P, const Rho1Call &c, int precision, hipStream_t s) {
  FRPA_REQUIRE(!c.grad, FRPA_ERR_UNSUPPORTED,
               "the FP64 symmetric-half / FP32 / BF16x3 paths compute rho1 only (the GGA gradient needs the "
               "full product: FRPA_PREC_FP64 or FRPA_PREC_OZAKI)");
  if (P.n_valid == 0)
    return;
  const bool f64 = precision == FRPA_PREC_FP64_SYM;
  // the P' gathers address one perturbation's n_basis^2 elements through a buffer resource (32-bit offsets)
  FRPA_REQUIRE((unsigned long long)P.n_basis * P.n_basis * (f64 ? 8 : 4) < (1ull << 32), FRPA_ERR_UNSUPPORTED,
               "rho1 precision %d: n_basis %d too large (n_basis^2 elements of P' must stay under 4 GiB)", precision,
               P.n_basis);
  if (f64) {
    // phi is staged from the wave in 16-byte pieces: the wave and every batch block 16-byte aligned
    FRPA_REQUIRE(P.wave_off_quads && (reinterpret_cast<uintptr_t>(c.wave) & 15) == 0, FRPA_ERR_UNSUPPORTED,
                 "FP64 symmetric-half rho1 needs a 16-byte aligned wave with batch offsets that are multiples of 4");
    if (!P.lowp)
      P.lowp = new LowpState();
    forward<FRPA_PREC_FP64_SYM>(P, c, s);
    return;
  }
  if (!P.lowp || P.lowp->conv[conv_index(precision)].wave != c.wave)
    lowp_prepare_wave(P, c.wave, precision, s);
  if (precision == FRPA_PREC_FP32)
    forward<FRPA_PREC_FP32>(P, c, s);
  else
    forward<FRPA_PREC_BF16X3>(P, c, s);
}

} // namespace frpa
