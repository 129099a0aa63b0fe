// rho1 forward with the GEMM W = P1loc * phi in reduced precision, for comparison with the FP64 and
// INT8-split paths (their error against FP64 is what the comparison states):
//   FRPA_PREC_FP32  : operands rounded to FP32, v_mfma_f32_16x16x4f32;
//   FRPA_PREC_BF16X3: each operand split x = hi + lo with hi = bf16(x), lo = bf16(x - hi), three
//                     products hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16.
// Both accumulate each K-chunk of 32 compute rows in FP32 MFMA accumulators and add the chunk into
// FP64 registers (compensation at chunk granularity); the ddot rho(p) = sum_i phi(i, p) W(i, p), the
// bias and the scatter stay FP64.  Data path as the FP64 kernel: P1 gathered straight from the
// caller's matrix, phi staged through LDS (4 waves of 16 rows x 64 points per pass).
// Reference: evaluate_first_order_rho.cpp:149-207 (the same per-batch GEMM + ddot).
#include <utility>

#include "rho1_common.hpp"

namespace frpa {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int LKC = 32;   // K chunk
constexpr int LNP = 64;   // points per pass
constexpr int LPLD = LNP + 1;

template <int MODE>
__global__ __launch_bounds__(256) void rho1_lowp_kernel(const BatchDesc *__restrict__ descs,
                                                        const int *__restrict__ ibasis0, const int *__restrict__ map,
                                                        const double *__restrict__ wave, const double *__restrict__ P1,
                                                        int n_basis, const double *__restrict__ bias,
                                                        double *__restrict__ rho, int n_full, int nmb,
                                                        int n_batches_all, int n_tile) {
  __shared__ double Bs[LKC * LPLD];
  __shared__ int jbs[LKC];
  __shared__ double red[4][LNP];
  const BatchDesc d = descs[blockIdx.x];
  const int t = blockIdx.y;
  const int nc = d.nc, ld = (nc + 3) & ~3;
  const double *W = wave + d.wave_off;
  const int *ib = ibasis0 + d.ib_off;
  const double *Pt = P1 + (size_t)n_basis * n_basis * t;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (int pc = 0; pc < d.np; pc += LNP) {
    const int np = min(LNP, d.np - pc);
    const double *Wp = W + (size_t)ld * pc;
    double racc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < nc; i0 += 64) {
      const int irow = i0 + 16 * w + lr;
      const bool iok = irow < nc;
      const size_t arow = iok ? (size_t)ib[irow] : 0;
      v4d acc[4];
#pragma unroll
      for (int g = 0; g < 4; g++)
        acc[g] = v4d{0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < nc; k0 += LKC) {
        __syncthreads();
        for (int e = tid; e < LKC * LNP; e += 256) {
          const int k = e % LKC, p = e / LKC, j = k0 + k;
          Bs[k * LPLD + p] = (j < nc && p < np) ? Wp[j + (size_t)ld * p] : 0.0;
        }
        if (tid < LKC)
          jbs[tid] = k0 + tid < nc ? ib[k0 + tid] : -1;
        __syncthreads();
        v4f cacc[4];
#pragma unroll
        for (int g = 0; g < 4; g++)
          cacc[g] = v4f{0.f, 0.f, 0.f, 0.f};
        if constexpr (MODE == FRPA_PREC_FP32) {
          // 16x16x4: A(m = lr, k = lk), B(k = lk, n = lr); 8 steps of 4 k
#pragma unroll
          for (int s = 0; s < LKC / 4; s++) {
            const int jb = jbs[4 * s + lk];
            const float a = (iok && jb >= 0) ? (float)Pt[arow + (size_t)n_basis * jb] : 0.f;
#pragma unroll
            for (int g = 0; g < 4; g++) {
              const float b = (float)Bs[(4 * s + lk) * LPLD + 16 * g + lr];
              cacc[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, cacc[g], 0, 0, 0);
            }
          }
        } else {
          // 16x16x32 bf16: lane holds k = 8 lk .. 8 lk + 7 of row m = lr (A) / column n = lr (B)
          v8bf ah, al;
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const int jb = jbs[8 * lk + q];
            const float x = (iok && jb >= 0) ? (float)Pt[arow + (size_t)n_basis * jb] : 0.f;
            const __bf16 h = (__bf16)x;
            ah[q] = h;
            al[q] = (__bf16)(x - (float)h);
          }
#pragma unroll
          for (int g = 0; g < 4; g++) {
            v8bf bh, bl;
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const float x = (float)Bs[(8 * lk + q) * LPLD + 16 * g + lr];
              const __bf16 h = (__bf16)x;
              bh[q] = h;
              bl[q] = (__bf16)(x - (float)h);
            }
            cacc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, cacc[g], 0, 0, 0);
            cacc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, cacc[g], 0, 0, 0);
            cacc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, cacc[g], 0, 0, 0);
          }
        }
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
          for (int r = 0; r < 4; r++)
            acc[g][r] += (double)cacc[g][r];
      }
      // both MFMA shapes leave D(4 (lane / 16) + r, lane % 16) in acc[r]:
      // acc[g][r] = W(i0 + 16 w + 4 lk + r, pc + 16 g + lr)
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int p = 16 * g + lr;
        double part = 0.0;
        if (p < np) {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int i = i0 + 16 * w + 4 * lk + r;
            if (i < nc)
              part += acc[g][r] * Wp[i + (size_t)ld * p];
          }
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        racc[g] += part;
      }
    }
    __syncthreads();
    if (lk == 0) {
#pragma unroll
      for (int g = 0; g < 4; g++)
        red[w][16 * g + lr] = racc[g];
    }
    __syncthreads();
    for (int p = tid; p < np; p += 256) {
      const size_t bi = (size_t)(pc + p) + (size_t)nmb * (d.b + (size_t)n_batches_all * t);
      double sum = red[0][p] + red[1][p] + red[2][p] + red[3][p];
      if (bias)
        sum += bias[bi];
      if (t < n_tile)
        rho[(size_t)map[d.pstart + pc + p] + (size_t)n_full * t] = sum;
    }
  }
}

} // namespace

void rho1_forward_lowp(Rho1Plan &P, const Rho1Call &c, int precision, hipStream_t s) {
  FRPA_REQUIRE(!c.grad, FRPA_ERR_UNSUPPORTED, "the FP32 / BF16x3 comparison paths compute rho1 only");
  if (P.n_valid == 0)
    return;
  dim3 grid(P.n_valid, c.T);
  if (precision == FRPA_PREC_FP32)
    hipLaunchKernelGGL(rho1_lowp_kernel<FRPA_PREC_FP32>, grid, dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr,
                       P.d_map.ptr, c.wave, c.P1, P.n_basis, c.bias, c.rho, P.n_full, P.nmb, P.n_batches, c.n_tile);
  else
    hipLaunchKernelGGL(rho1_lowp_kernel<FRPA_PREC_BF16X3>, grid, dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr,
                       P.d_map.ptr, c.wave, c.P1, P.n_basis, c.bias, c.rho, P.n_full, P.nmb, P.n_batches, c.n_tile);
  check_launch("rho1_lowp_kernel");
}

} // namespace frpa
