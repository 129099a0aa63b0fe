// rho1 forward with the GEMM W = P1loc * phi emulated on INT8 MFMA (Ozaki-style split).
//
// Both operands are written as sums of signed 8-bit digits in radix 254:
//   P1(r, c) * sa_r = sum_{s<S} a_s(r, c) 254^-s,   sa_r = 127 / max_c |P1(r, c)|   (per P1 row)
//   phi(i, p) * sb_p = sum_{s<S} b_s(i, p) 254^-s,  sb_p = 127 / max_i |phi(i, p)|  (per point)
// with round-to-nearest digits: |x| <= 127 and |x - rint(x)| <= 1/2, so every digit fits int8 and
// carries log2(254) ~ 8 bits.  The digit products a_s * b_u are exact INT8 MFMA GEMMs with exact
// int32 accumulation (|sum| <= 5 * 1200 * 127^2 < 2^31); the levels s + u <= S - 1 are kept,
// S(S+1)/2 products, and recombined in FP64:
//   W(i, p) = sum_l 254^-l C_l(i, p) / (sa_i sb_p),   C_l = sum_{s+u=l} a_s b_u   (exact int32)
// The ddot rho(p) = sum_i phi(i, p) W(i, p), the bias and the scatter stay FP64, fused as in the
// FP64 kernel.  S = 5 (15 INT8 GEMMs) keeps max|d rho1| / max|rho1| ~ 1e-12; S = 4 (10 GEMMs)
// ~ 5e-10 on the synthetic workloads.
//
// Data: the phi digits are prepared once per wave (frpa_rho1_prepare_wave) as S int8 planes per
// batch, (ncK, np) column-major with ncK = nc rounded up to 64 (zero rows), so a lane's B
// fragment (16 consecutive k of one point) is one 16-byte load.  The P1 digits are rebuilt per call
// (P1 changes every call): one int64 per element holding its S digits, gathered exactly like the
// FP64 path gathers P1 (one 8-byte load per element), then byte-transposed into the S A fragments.
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
  DevBuf<long long> d_adig;         // P1 digits, (n_basis, n_basis, T)
  DevBuf<double> d_ainv;            // 1 / sa_r, (n_basis, T)
  DevBuf<unsigned long long> d_rowbits;  // row maxima of |P1| as IEEE bits
  int n_valid_points = 0;
};

void ozaki_release(OzakiState *st) { delete st; }

namespace {

constexpr double RADIX = 254.0;

__device__ __forceinline__ int ceil64(int x) { return (x + 63) & ~63; }

// ---- phi digits (one-off) -------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void phi_digits_kernel(const BatchDesc *__restrict__ descs,
                                                         const long long *__restrict__ boff,
                                                         const double *__restrict__ wave,
                                                         signed char *__restrict__ bdig, double *__restrict__ binv) {
  const BatchDesc d = descs[blockIdx.x];
  const int ld = (d.nc + 3) & ~3, ncK = ceil64(d.nc);
  const double *W = wave + d.wave_off;
  signed char *out = bdig + boff[blockIdx.x];
  const size_t plane = (size_t)ncK * d.np;
  for (int p = threadIdx.x; p < d.np; p += blockDim.x) {
    double mx = 0.0;
    for (int i = 0; i < d.nc; i++)
      mx = fmax(mx, fabs(W[i + (size_t)ld * p]));
    const double sb = mx > 0.0 ? 127.0 / mx : 1.0;
    binv[d.pstart + p] = 1.0 / sb;
    for (int i = 0; i < ncK; i++) {
      double r = i < d.nc ? W[i + (size_t)ld * p] * sb : 0.0;
#pragma unroll
      for (int s = 0; s < S; s++) {
        const double dg = rint(r);
        out[s * plane + (size_t)ncK * p + i] = (signed char)dg;
        r = (r - dg) * RADIX;
      }
    }
  }
}

// ---- P1 digits (per call) -------------------------------------------------------------------
// max_c |P1(r, c, t)|: column chunks per block, combined with atomicMax on the bit pattern
// (non-negative doubles order like their IEEE bits); rowbits must start at 0.
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

__global__ void p1_rowscale_kernel(const unsigned long long *__restrict__ rowbits, size_t n,
                                   double *__restrict__ ainv) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const double mx = __longlong_as_double((long long)rowbits[i]);
  ainv[i] = mx > 0.0 ? mx / 127.0 : 1.0;
}

template <int S>
__global__ void p1_digits_kernel(const double *__restrict__ P1, const double *__restrict__ ainv, int n_basis,
                                 size_t n_total, long long *__restrict__ adig) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n_total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t r = e % n_basis, t = e / ((size_t)n_basis * n_basis);
    double x = P1[e] / ainv[r + (size_t)n_basis * t];
    unsigned long long packed = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const double dg = rint(x);
      packed |= (unsigned long long)(unsigned char)(signed char)dg << (8 * s);
      x = (x - dg) * RADIX;
    }
    adig[e] = (long long)packed;
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
constexpr int BROW = 80;  // bytes per (digit, point) row of a staged B chunk: 64 k + 16 pad

// One point chunk with exactly NPG groups of 16 points (compile-time loops, no per-group branches),
// K-chunks of 64 double-buffered in LDS with the next chunk's phi digits prefetched into registers.
template <int S, int NPG, int PCH>
__device__ __forceinline__ void ozaki_chunk(signed char *Bs0, signed char *Bs1, int *jbs0, int *jbs1, double *red,
                                            const BatchDesc &d, int t, int pc, int npc, const int *__restrict__ ib,
                                            const int *__restrict__ map, const double *__restrict__ W,
                                            const long long *__restrict__ At, const double *__restrict__ ai,
                                            const signed char *__restrict__ Bg, const double *__restrict__ binv,
                                            int n_basis, const double *__restrict__ bias, double *__restrict__ rho,
                                            int n_full, int nmb, int n_batches_all, int n_tile) {
  constexpr int NQ = S * NPG * 16 * 4;          // 16-byte pieces of one staged chunk
  constexpr int QPT = (NQ + 255) / 256;         // per thread
  const int nc = d.nc, ld = (nc + 3) & ~3, ncK = ceil64(nc);
  const size_t plane = (size_t)ncK * d.np;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nchunk = ncK / 64;
  double racc[NPG];
#pragma unroll
  for (int g = 0; g < NPG; g++)
    racc[g] = 0.0;

  for (int i0 = 0; i0 < nc; i0 += 64) {
    const int irow = i0 + 16 * w + lr;
    const bool iok = irow < nc;
    const size_t arow = iok ? (size_t)ib[irow] : 0;
    v4i acc[S][NPG];
#pragma unroll
    for (int l = 0; l < S; l++)
#pragma unroll
      for (int g = 0; g < NPG; g++)
        acc[l][g] = v4i{0, 0, 0, 0};

    v4i bq[QPT];
    auto load_b = [&](int c) {
      const int j0 = 64 * c;
#pragma unroll
      for (int u = 0; u < QPT; u++) {
        const int e = tid + 256 * u;
        const int q = e & 3, p = (e >> 2) % (NPG * 16), sl = (e >> 2) / (NPG * 16);
        bq[u] = (e < NQ && p < npc)
                    ? *reinterpret_cast<const v4i *>(Bg + sl * plane + (size_t)ncK * (pc + p) + j0 + 16 * q)
                    : v4i{0, 0, 0, 0};
      }
    };
    auto store_b = [&](int c, signed char *Bs, int *jbs) {
#pragma unroll
      for (int u = 0; u < QPT; u++) {
        const int e = tid + 256 * u;
        const int q = e & 3, p = (e >> 2) % (NPG * 16), sl = (e >> 2) / (NPG * 16);
        if (e < NQ)
          *reinterpret_cast<v4i *>(&Bs[(sl * PCH + p) * BROW + 16 * q]) = bq[u];
      }
      if (tid < 64) {
        const int j = 64 * c + tid;
        jbs[tid] = j < nc ? ib[j] : -1;
      }
    };
    __syncthreads();
    load_b(0);
    store_b(0, Bs0, jbs0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
      const signed char *Bs = (c & 1) ? Bs1 : Bs0;
      const int *jbs = (c & 1) ? jbs1 : jbs0;
      // A fragments: row irow, k = 64 c + 16 lk + q, q < 16 (P1 digits gathered like P1 itself)
      unsigned long long e[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int jb = jbs[16 * lk + q];
        e[q] = (iok && jb >= 0) ? (unsigned long long)At[arow + (size_t)n_basis * jb] : 0ull;
      }
      if (c + 1 < nchunk)
        load_b(c + 1);
      v4i a[S];
#pragma unroll
      for (int sl = 0; sl < S; sl++)
        a[sl] = v4i{gather_bytes(e[0], e[1], e[2], e[3], sl), gather_bytes(e[4], e[5], e[6], e[7], sl),
                    gather_bytes(e[8], e[9], e[10], e[11], sl), gather_bytes(e[12], e[13], e[14], e[15], sl)};
#pragma unroll
      for (int g = 0; g < NPG; g++) {
        v4i b[S];
#pragma unroll
        for (int sl = 0; sl < S; sl++)
          b[sl] = *reinterpret_cast<const v4i *>(&Bs[(sl * PCH + 16 * g + lr) * BROW + 16 * lk]);
#pragma unroll
        for (int sa = 0; sa < S; sa++)
#pragma unroll
          for (int sb = 0; sb < S - sa; sb++)
            acc[sa + sb][g] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[sa], b[sb], acc[sa + sb][g], 0, 0, 0);
      }
      if (c + 1 < nchunk)
        store_b(c + 1, (c & 1) ? Bs0 : Bs1, (c & 1) ? jbs0 : jbs1);
      __syncthreads();
    }

    // epilogue: v_mfma_i32_16x16x64_i8 leaves D(4 (lane / 16) + r, lane % 16) in acc[r], so
    // acc[l][g][r] = C_l(i0 + 16w + 4lk + r, pc + 16g + lr)
    double rinv[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = i0 + 16 * w + 4 * lk + r;
      rinv[r] = i < nc ? ai[ib[i]] : 0.0;
    }
#pragma unroll
    for (int g = 0; g < NPG; g++) {
      const int p = 16 * g + lr;
      double part = 0.0;
      if (p < npc) {
        const double bi = binv[d.pstart + pc + p];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = i0 + 16 * w + 4 * lk + r;
          if (i < nc) {
            double wv = (double)acc[S - 1][g][r];
#pragma unroll
            for (int l = S - 2; l >= 0; l--)
              wv = wv * (1.0 / RADIX) + (double)acc[l][g][r];
            part += wv * rinv[r] * bi * W[i + (size_t)ld * (pc + p)];
          }
        }
      }
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      racc[g] += part;
    }
  }

  double(*rd)[PCH] = reinterpret_cast<double(*)[PCH]>(red);
  __syncthreads();
  if (lk == 0) {
#pragma unroll
    for (int g = 0; g < NPG; g++)
      rd[w][16 * g + lr] = racc[g];
  }
  __syncthreads();
  for (int p = tid; p < npc; p += 256) {
    const size_t bi = (size_t)(pc + p) + (size_t)nmb * (d.b + (size_t)n_batches_all * t);
    double sum = rd[0][p] + rd[1][p] + rd[2][p] + rd[3][p];
    if (bias)
      sum += bias[bi];
    if (t < n_tile)
      rho[(size_t)map[d.pstart + pc + p] + (size_t)n_full * t] = sum;
  }
}

template <int S, int MAXPG, int... Gs>
__device__ __forceinline__ void ozaki_dispatch(std::integer_sequence<int, Gs...>, int npg, signed char *Bs0,
                                               signed char *Bs1, int *jbs0, int *jbs1, double *red,
                                               const BatchDesc &d, int t, int pc, int npc, const int *ib,
                                               const int *map, const double *W, const long long *At,
                                               const double *ai, const signed char *Bg, const double *binv,
                                               int n_basis, const double *bias, double *rho, int n_full, int nmb,
                                               int n_batches_all, int n_tile) {
  ((npg == Gs + 1 ? ozaki_chunk<S, Gs + 1, 16 * MAXPG>(Bs0, Bs1, jbs0, jbs1, red, d, t, pc, npc, ib, map, W, At, ai,
                                                         Bg, binv, n_basis, bias, rho, n_full, nmb, n_batches_all,
                                                         n_tile)
                  : (void)0),
   ...);
}

template <int S, int MAXPG>
__global__ __launch_bounds__(256) void rho1_ozaki_kernel(
    const BatchDesc *__restrict__ descs, const int *__restrict__ ibasis0, const int *__restrict__ map,
    const double *__restrict__ wave, const long long *__restrict__ adig, const double *__restrict__ ainv,
    const signed char *__restrict__ bdig, const long long *__restrict__ boff, const double *__restrict__ binv,
    int n_basis, const double *__restrict__ bias, double *__restrict__ rho, int n_full, int nmb, int n_batches_all,
    int n_tile) {
  constexpr int PCH = 16 * MAXPG;
  __shared__ __attribute__((aligned(16))) signed char Bs0[S * PCH * BROW];
  __shared__ __attribute__((aligned(16))) signed char Bs1[S * PCH * BROW];
  __shared__ int jbs0[64], jbs1[64];
  __shared__ double red[4 * PCH];
  const BatchDesc d = descs[blockIdx.x];
  const int t = blockIdx.y;
  const double *W = wave + d.wave_off;
  const int *ib = ibasis0 + d.ib_off;
  const long long *At = adig + (size_t)n_basis * n_basis * t;
  const double *ai = ainv + (size_t)n_basis * t;
  const signed char *Bg = bdig + boff[blockIdx.x];
  for (int pc = 0; pc < d.np; pc += PCH) {
    const int npc = min(PCH, d.np - pc);
    const int npg = (npc + 15) >> 4;
    ozaki_dispatch<S, MAXPG>(std::make_integer_sequence<int, MAXPG>{}, npg, Bs0, Bs1, jbs0, jbs1, red, d, t, pc, npc,
                             ib, map, W, At, ai, Bg, binv, n_basis, bias, rho, n_full, nmb, n_batches_all, n_tile);
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
  hipLaunchKernelGGL(phi_digits_kernel<S>, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, st.d_boff.ptr, wave,
                     st.d_bdig.ptr, st.d_binv.ptr);
  check_launch("phi_digits_kernel");
  st.S = S;
  st.wave = wave;
  st.bytes = off[P.n_valid] + (long long)nvp * 8;
}

template <int S, int NPG>
void launch(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  OzakiState &st = *P.ozaki;
  hipLaunchKernelGGL((rho1_ozaki_kernel<S, NPG>), dim3(P.n_valid, c.T), dim3(256), 0, s, P.d_desc.ptr,
                     P.d_ibasis0.ptr, P.d_map.ptr, c.wave, st.d_adig.ptr, st.d_ainv.ptr, st.d_bdig.ptr,
                     st.d_boff.ptr, st.d_binv.ptr, P.n_basis, c.bias, c.rho, P.n_full, P.nmb, P.n_batches, c.n_tile);
  check_launch("rho1_ozaki_kernel");
}

template <int S>
void forward(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  OzakiState &st = *P.ozaki;
  const size_t nbb = (size_t)P.n_basis * P.n_basis * c.T;
  const size_t nr = (size_t)P.n_basis * c.T;
  st.d_adig.alloc(nbb);
  st.d_ainv.alloc(nr);
  st.d_rowbits.alloc(nr);
  fill_f64(reinterpret_cast<double *>(st.d_rowbits.ptr), nr, 0.0, s);  // +0.0 has all-zero bits
  const int chunk = 256;
  hipLaunchKernelGGL(p1_rowmax_kernel, dim3((P.n_basis + 255) / 256, (P.n_basis + chunk - 1) / chunk, c.T),
                     dim3(256), 0, s, c.P1, P.n_basis, chunk, st.d_rowbits.ptr);
  check_launch("p1_rowmax_kernel");
  hipLaunchKernelGGL(p1_rowscale_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, st.d_rowbits.ptr, nr,
                     st.d_ainv.ptr);
  check_launch("p1_rowscale_kernel");
  const unsigned blocks = (unsigned)std::min<size_t>((nbb + 255) / 256, 65536);
  hipLaunchKernelGGL(p1_digits_kernel<S>, dim3(blocks), dim3(256), 0, s, c.P1, st.d_ainv.ptr, P.n_basis, nbb,
                     st.d_adig.ptr);
  check_launch("p1_digits_kernel");
  // point-chunk width: the largest that keeps the accumulators and staged digits in registers
  // (96 points for 5 digits, 112 for 4); wider batches take several passes
  if (P.max_np <= 64)
    launch<S, 4>(P, c, s);
  else if (S == 5 || P.max_np <= 96)
    launch<S, 6>(P, c, s);
  else
    launch<S, 7>(P, c, s);
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
  FRPA_REQUIRE(!c.grad, FRPA_ERR_UNSUPPORTED, "the INT8-split path computes rho1 only (no GGA gradient)");
  if (P.n_valid == 0)
    return;
  if (!P.ozaki || P.ozaki->wave != c.wave || P.ozaki->S != S)
    ozaki_prepare_wave(P, c.wave, S, s);
  if (P.ozaki->S == 4)
    forward<4>(P, c, s);
  else
    forward<5>(P, c, s);
}

} // namespace frpa
