// rho1 forward in FP64 on gfx950: one fused kernel per call.
//
// Reference: fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.cu:20-371 runs three launches per
// tile of 256 batches (gather of P1 into dense per-batch slots, MAGMA vbatched DGEMM writing
// W = P1loc * phi to HBM, and a ddot kernel re-reading W with one thread per point).  Here one
// workgroup owns one batch (and one perturbation t):
//   * the P1 gather is folded into the A-operand load of the MFMA (rows i_basis(i), columns
//     i_basis(j) read straight from P1; consecutive lanes read consecutive rows);
//   * phi (the B operand) is staged through LDS in K-chunks, shared by the four waves;
//   * W = P1loc * phi lives only in MFMA accumulators (v_mfma_f64_16x16x4f64);
//   * the ddot rho(p) = sum_i phi(i, p) W(i, p) (and the GGA gradient sum) is the epilogue,
//     reduced across the 16-row MFMA tile with cross-lane shuffles and across waves in LDS;
//   * bias add and the scatter to first_order_rho(map(p), t) are fused.
#include <utility>

#include "rho1_common.hpp"

namespace frpa {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int KC = 16;  // K chunk staged in LDS (double-buffered)

struct Fp64Smem {
  double *Bs0, *Bs1;  // two K-chunk buffers of KC x (16 MAXPG + 1) doubles
  int *jbs0, *jbs1;
  double *red;        // (NRED, 4, PMAX)
};

// One pass over a point chunk with exactly NPG groups of 16 points: every loop below has a
// compile-time trip count, so the MFMA stream carries no per-group branches and the LDS reads and
// the prefetch of the next chunk can be scheduled around the MFMAs.
template <int NPG, int PMAX, bool GRAD>
__device__ __forceinline__ void rho1_fp64_chunk(const Fp64Smem &sm, const BatchDesc &d, int t, int pc, int np,
                                                const int *__restrict__ ib, const int *__restrict__ map,
                                                const double *__restrict__ W, const double *__restrict__ G,
                                                const double *__restrict__ Pt, int n_basis,
                                                const double *__restrict__ bias, const double *__restrict__ gbias,
                                                double *__restrict__ rho, double *__restrict__ grad, int n_full,
                                                int nmb, int n_batches_all, int n_tile) {
  constexpr int PLD = PMAX + 1;
  constexpr int NP16 = 16 * NPG;
  constexpr int BPT = (KC * NP16 + 255) / 256;  // staged B elements per thread
  const int nc = d.nc, ld = (nc + 3) & ~3;
  const double *Wp = W + (size_t)ld * pc;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nchunk = (nc + KC - 1) / KC;

  double racc[NPG];
  double gacc[3][NPG];
#pragma unroll
  for (int g = 0; g < NPG; g++) {
    racc[g] = 0.0;
    if (GRAD)
      gacc[0][g] = gacc[1][g] = gacc[2][g] = 0.0;
  }

  for (int i0 = 0; i0 < nc; i0 += 64) {
    const int irow = i0 + 16 * w + lr;
    const bool iok = irow < nc;
    const size_t arow = iok ? (size_t)ib[irow] : 0;
    v4d acc[NPG];
#pragma unroll
    for (int g = 0; g < NPG; g++)
      acc[g] = v4d{0.0, 0.0, 0.0, 0.0};

    double breg[BPT];
    auto load_b = [&](int c) {
      const int j0 = c * KC;
#pragma unroll
      for (int q = 0; q < BPT; q++) {
        const int e = tid + 256 * q;
        const int k = e % KC, p = e / KC;
        const int j = j0 + k;
        breg[q] = (e < KC * NP16 && j < nc && p < np) ? Wp[j + (size_t)ld * p] : 0.0;
      }
    };
    auto store_b = [&](int c, double *Bs, int *jbs) {
#pragma unroll
      for (int q = 0; q < BPT; q++) {
        const int e = tid + 256 * q;
        const int k = e % KC, p = e / KC;
        if (e < KC * NP16)
          Bs[k * PLD + p] = breg[q];
      }
      if (tid < KC) {
        const int j = c * KC + tid;
        jbs[tid] = j < nc ? ib[j] : -1;
      }
    };
    __syncthreads();  // previous users of the buffers are done
    load_b(0);
    store_b(0, sm.Bs0, sm.jbs0);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
      const double *Bs = (c & 1) ? sm.Bs1 : sm.Bs0;
      const int *jbs = (c & 1) ? sm.jbs1 : sm.jbs0;
      double a[KC / 4];
#pragma unroll
      for (int s = 0; s < KC / 4; s++) {
        const int jb = jbs[4 * s + lk];
        a[s] = (iok && jb >= 0) ? Pt[arow + (size_t)n_basis * jb] : 0.0;
      }
      if (c + 1 < nchunk)
        load_b(c + 1);  // in flight while the MFMAs below run
#pragma unroll
      for (int s = 0; s < KC / 4; s++) {
        const double *brow = &Bs[(4 * s + lk) * PLD + lr];
#pragma unroll
        for (int g = 0; g < NPG; g++)
          acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], brow[16 * g], acc[g], 0, 0, 0);
      }
      if (c + 1 < nchunk)
        store_b(c + 1, (c & 1) ? sm.Bs0 : sm.Bs1, (c & 1) ? sm.jbs0 : sm.jbs1);
      __syncthreads();
    }

    // epilogue: v_mfma_f64_16x16x4f64 leaves D(4 r + lane / 16, lane % 16) in acc[g][r], so
    // acc[g][r] = W(i0 + 16w + 4 r + lk, pc + 16 g + lr)
#pragma unroll
    for (int g = 0; g < NPG; g++) {
      const int p = 16 * g + lr;
      double part = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
      if (p < np) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = i0 + 16 * w + 4 * r + lk;
          if (i < nc) {
            part += acc[g][r] * Wp[i + (size_t)ld * p];
            if (GRAD) {
              const double *gp = G + (size_t)3 * ld * (pc + p) + i;
              p0 += acc[g][r] * gp[0];
              p1 += acc[g][r] * gp[ld];
              p2 += acc[g][r] * gp[2 * ld];
            }
          }
        }
      }
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      racc[g] += part;
      if (GRAD) {
        p0 += __shfl_xor(p0, 16);
        p0 += __shfl_xor(p0, 32);
        p1 += __shfl_xor(p1, 16);
        p1 += __shfl_xor(p1, 32);
        p2 += __shfl_xor(p2, 16);
        p2 += __shfl_xor(p2, 32);
        gacc[0][g] += p0;
        gacc[1][g] += p1;
        gacc[2][g] += p2;
      }
    }
  }

  double(*red)[4][PMAX] = reinterpret_cast<double(*)[4][PMAX]>(sm.red);
  __syncthreads();
  if (lk == 0) {
#pragma unroll
    for (int g = 0; g < NPG; g++) {
      red[0][w][16 * g + lr] = racc[g];
      if (GRAD) {
        red[1][w][16 * g + lr] = gacc[0][g];
        red[2][w][16 * g + lr] = gacc[1][g];
        red[3][w][16 * g + lr] = gacc[2][g];
      }
    }
  }
  __syncthreads();
  for (int p = tid; p < np; p += 256) {
    const size_t bi = (size_t)(pc + p) + (size_t)nmb * (d.b + (size_t)n_batches_all * t);
    double sum = red[0][0][p] + red[0][1][p] + red[0][2][p] + red[0][3][p];
    if (bias)
      sum += bias[bi];
    if (t < n_tile)
      rho[(size_t)map[d.pstart + pc + p] + (size_t)n_full * t] = sum;
    if (GRAD) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double v = red[1 + c][0][p] + red[1 + c][1][p] + red[1 + c][2][p] + red[1 + c][3][p];
        grad[c + 3 * bi] = v * 2 - (gbias ? gbias[c + 3 * bi] : 0.0);
      }
    }
  }
}

template <int MAXPG, bool GRAD, int... Gs>
__device__ __forceinline__ void rho1_fp64_dispatch(std::integer_sequence<int, Gs...>, int npg, const Fp64Smem &sm,
                                                   const BatchDesc &d, int t, int pc, int np, const int *ib,
                                                   const int *map, const double *W, const double *G,
                                                   const double *Pt, int n_basis, const double *bias,
                                                   const double *gbias, double *rho, double *grad, int n_full,
                                                   int nmb, int n_batches_all, int n_tile) {
  ((npg == Gs + 1 ? rho1_fp64_chunk<Gs + 1, 16 * MAXPG, GRAD>(sm, d, t, pc, np, ib, map, W, G, Pt, n_basis, bias,
                                                               gbias, rho, grad, n_full, nmb, n_batches_all, n_tile)
                  : (void)0),
   ...);
}

template <int MAXPG, bool GRAD>
__global__ __launch_bounds__(256) void rho1_fp64_kernel(const BatchDesc *__restrict__ descs,
                                                        const int *__restrict__ order, int *__restrict__ counters,
                                                        int n_pos, int T, const int *__restrict__ ibasis0,
                                                        const int *__restrict__ map, const double *__restrict__ wave,
                                                        const double *__restrict__ gwave,
                                                        const double *__restrict__ P1, int n_basis,
                                                        const double *__restrict__ bias,
                                                        const double *__restrict__ gbias, double *__restrict__ rho,
                                                        double *__restrict__ grad, int n_full, int nmb,
                                                        int n_batches_all, int n_tile) {
  constexpr int PMAX = 16 * MAXPG;
  constexpr int PLD = PMAX + 1;
  constexpr int NRED = GRAD ? 4 : 1;
  __shared__ double Bs0[KC * PLD], Bs1[KC * PLD];
  __shared__ int jbs0[KC], jbs1[KC];
  __shared__ double red[NRED * 4 * PMAX];
  __shared__ int s_item;
  const Fp64Smem sm{Bs0, Bs1, jbs0, jbs1, red};
  int pos = blockIdx.x, t = blockIdx.y;
  for (bool first = true;; first = false) {
    if (counters) {
      if (!claim_item(counters, n_pos, T, pos, t, &s_item))
        break;
    } else if (!first) {
      break;
    }
    const BatchDesc d = descs[order ? order[pos] : pos];
    const double *W = wave + d.wave_off;
    const double *G = GRAD ? gwave + 3 * d.wave_off : nullptr;
    const int *ib = ibasis0 + d.ib_off;
    const double *Pt = P1 + (size_t)n_basis * n_basis * t;
    for (int pc = 0; pc < d.np; pc += PMAX) {
      const int np = min(PMAX, d.np - pc);
      const int npg = (np + 15) >> 4;
      rho1_fp64_dispatch<MAXPG, GRAD>(std::make_integer_sequence<int, MAXPG>{}, npg, sm, d, t, pc, np, ib, map, W,
                                      G, Pt, n_basis, bias, gbias, rho, grad, n_full, nmb, n_batches_all, n_tile);
    }
  }
}

template <int MAXPG, bool GRAD>
void launch_fp64(Rho1Plan &P, const Rho1Call &c, const LaunchCfg &cfg, hipStream_t s) {
  static int per_cu = -1;
  if (per_cu < 0) {
    FRPA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rho1_fp64_kernel<MAXPG, GRAD>, 256, 0));
    per_cu = std::max(per_cu, 1);
  }
  const LaunchArgs la = launch_args(P, cfg, c.T, per_cu, s);
  hipLaunchKernelGGL((rho1_fp64_kernel<MAXPG, GRAD>), la.grid, dim3(256), 0, s, P.d_desc.ptr, la.order, la.counters,
                     P.n_valid, c.T, P.d_ibasis0.ptr, P.d_map.ptr, c.wave, c.gwave, c.P1, P.n_basis, c.bias,
                     c.gbias, c.rho, c.grad, P.n_full, P.nmb, P.n_batches, c.n_tile);
  check_launch("rho1_fp64_kernel");
}

template <bool GRAD>
void dispatch_fp64(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  const LaunchCfg &cfg = P.cfg[FRPA_PREC_FP64];
  // points per pass 16 MAXPG: by default the smallest of 64 / 128 / 208 that holds the largest
  // batch (208: larger batches loop over chunks)
  const int mp = cfg.npg ? cfg.npg : (P.max_np <= 64 ? 4 : P.max_np <= 128 ? 8 : 13);
  if (mp == 4)
    launch_fp64<4, GRAD>(P, c, cfg, s);
  else if (mp == 8)
    launch_fp64<8, GRAD>(P, c, cfg, s);
  else
    launch_fp64<13, GRAD>(P, c, cfg, s);
}

} // namespace

void rho1_forward_fp64(Rho1Plan &P, const Rho1Call &c, hipStream_t s) {
  if (P.n_valid == 0)
    return;
  if (c.grad)
    dispatch_fp64<true>(P, c, s);
  else
    dispatch_fp64<false>(P, c, s);
}

} // namespace frpa
