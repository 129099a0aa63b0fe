// Bias input of rho1: first_order_rho_bias_part2, the derivative of the unperturbed density
// rho(p) = sum_ik phi_i DM_ik phi_k with respect to displacing the tile's atoms.
// Reference: generate_first_order_rho_bias_part2_ (fast_rap_src_dfpt_gpu_eval/
// generate_first_order_rho_bias_part2.cpp:9-167), a CPU function with two dgemms per (batch, atom).
//
// For batch b and tile atom j (t = j - j_atom_begin), V_t = its compute rows [avn(j-1, b), avn(j, b)):
//   bias(p, b, t) = sum_{k in V_t} phi'(k, p) sum_i (DM(i, k) + DM(k, i)) phi(i, p),
//   phi' = -d phi / d x_{j_coord},  DM(i, k) = density_matrix(i_basis(i), i_basis(k))
// (the reference's two products, sum_i phi_i sum_{k in V} DM_ik phi'_k and
// sum_{i in V} phi'_i sum_k DM_ik phi_k, regrouped per row k of V); 0 where V_t is empty.
//
// One workgroup per batch covers the whole atom tile: the rows of all V_t (disjoint, at most nc) are
// stacked into 64-row tiles, so phi is read once per point pass for every atom of the tile.  The
// A operand (DM + DM^T)(i_basis(V), i_basis(:)) is gathered straight from the density matrix, phi is
// staged through LDS in K-chunks shared by the four waves, W = A phi lives in MFMA accumulators
// (v_mfma_f64_16x16x4f64) and the product with phi' is the epilogue, reduced per atom of the tile
// across lanes with shuffles and across waves in LDS (deterministic).  FP64 throughout.
#include "rho1_common.hpp"

namespace frpa {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int B2_KC = 16;    // K chunk
constexpr int B2_NP = 64;    // points per pass (4 MFMA column groups)
constexpr int B2_PLD = B2_NP + 1;
constexpr int B2_TMAX = 8;   // atoms per tile

__global__ __launch_bounds__(256) void bias2_kernel(const BatchDesc *__restrict__ descs,
                                                   const int *__restrict__ ibasis0, const int *__restrict__ avn,
                                                   int n_atoms, int j_atom_begin, int n_tile, int j_coord,
                                                   const double *__restrict__ wave, const double *__restrict__ gwave,
                                                   const double *__restrict__ DM, int n_basis,
                                                   double *__restrict__ bias, int nmb, int n_batches_all) {
  __shared__ double Bs[2][B2_KC * B2_PLD];
  __shared__ int jbs[2][B2_KC];
  __shared__ double red[4][B2_TMAX][B2_NP];
  __shared__ int s_vs[B2_TMAX], s_off[B2_TMAX + 1];
  const BatchDesc d = descs[blockIdx.x];
  const int nc = d.nc, ld = (nc + 3) & ~3;
  const int *__restrict__ ib = ibasis0 + d.ib_off;
  const double *__restrict__ W = wave + d.wave_off;
  const double *__restrict__ G = gwave + 3 * d.wave_off + (size_t)ld * (j_coord - 1);  // component j_coord
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid == 0) {
    const int *av = avn + (size_t)(n_atoms + 1) * d.b;
    int o = 0;
    for (int t = 0; t < n_tile; t++) {
      // clamped to the batch's rows: a malformed prefix cannot index outside its phi block
      const int vs = min(max(av[j_atom_begin - 1 + t], 0), nc);
      const int ve = min(max(av[j_atom_begin + t], vs), nc);
      s_vs[t] = vs;
      s_off[t] = o;
      o += ve - vs;
    }
    s_off[n_tile] = o;
  }
  __syncthreads();
  const int nR = s_off[n_tile];
  // stacked row r -> (tile atom t, compute row k)
  auto row_of = [&](int r, int &t, int &k) {
    t = 0;
    while (t + 1 < n_tile && s_off[t + 1] <= r)
      t++;
    k = s_vs[t] + r - s_off[t];
  };
  const int nchunk = (nc + B2_KC - 1) / B2_KC;

  for (int pc = 0; pc < d.np; pc += B2_NP) {
    const int npp = min(B2_NP, d.np - pc);
    const double *__restrict__ Wp = W + (size_t)ld * pc;
    for (int e = tid; e < 4 * B2_TMAX * B2_NP; e += 256)
      (&red[0][0][0])[e] = 0.0;
    for (int r0 = 0; r0 < nR; r0 += 64) {
      const int r = r0 + 16 * w + lr;
      const bool rok = r < nR;
      size_t arow = 0;
      if (rok) {
        int t, k;
        row_of(r, t, k);
        arow = (size_t)ib[k];
      }
      v4d acc[4];
#pragma unroll
      for (int g = 0; g < 4; g++)
        acc[g] = v4d{0.0, 0.0, 0.0, 0.0};
      double breg[4];
      auto load_b = [&](int c) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int e = tid + 256 * q;
          const int k = e % B2_KC, p = e / B2_KC, j = c * B2_KC + k;
          breg[q] = (j < nc && p < npp) ? Wp[j + (size_t)ld * p] : 0.0;
        }
      };
      auto store_b = [&](int c, int buf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int e = tid + 256 * q;
          Bs[buf][(e % B2_KC) * B2_PLD + e / B2_KC] = breg[q];
        }
        if (tid < B2_KC) {
          const int j = c * B2_KC + tid;
          jbs[buf][tid] = j < nc ? ib[j] : -1;
        }
      };
      __syncthreads();  // previous users of Bs / jbs are done
      load_b(0);
      store_b(0, 0);
      __syncthreads();
      for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        double a[B2_KC / 4];
#pragma unroll
        for (int s = 0; s < B2_KC / 4; s++) {
          const int jb = jbs[buf][4 * s + lk];
          a[s] = (rok && jb >= 0) ? DM[arow + (size_t)n_basis * jb] + DM[(size_t)jb + (size_t)n_basis * arow] : 0.0;
        }
        if (c + 1 < nchunk)
          load_b(c + 1);
#pragma unroll
        for (int s = 0; s < B2_KC / 4; s++) {
          const double *brow = &Bs[buf][(4 * s + lk) * B2_PLD + lr];
#pragma unroll
          for (int g = 0; g < 4; g++)
            acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], brow[16 * g], acc[g], 0, 0, 0);
        }
        if (c + 1 < nchunk)
          store_b(c + 1, buf ^ 1);
        __syncthreads();
      }
      // acc[g][i] = W(row r0 + 16 w + 4 i + lk, point pc + 16 g + lr); times phi'(k_row, p), summed per t
      int rt[4], rk[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rr = r0 + 16 * w + 4 * i + lk;
        rt[i] = -1;
        rk[i] = 0;
        if (rr < nR)
          row_of(rr, rt[i], rk[i]);
      }
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int p = 16 * g + lr;
        double v[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
          v[i] = (rt[i] >= 0 && p < npp) ? -acc[g][i] * G[rk[i] + (size_t)3 * ld * (pc + p)] : 0.0;
        for (int t = 0; t < n_tile; t++) {
          double part = 0.0;
#pragma unroll
          for (int i = 0; i < 4; i++)
            part += rt[i] == t ? v[i] : 0.0;
          part += __shfl_xor(part, 16);
          part += __shfl_xor(part, 32);
          if (lk == 0)
            red[w][t][p] += part;
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < n_tile * npp; e += 256) {
      const int t = e / npp, p = e % npp;
      bias[(size_t)(pc + p) + (size_t)nmb * (d.b + (size_t)n_batches_all * t)] =
          red[0][t][p] + red[1][t][p] + red[2][t][p] + red[3][t][p];
    }
    __syncthreads();
  }
}

} // namespace

void rho1_bias_part2(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                     const double *wave, const double *gwave, const double *DM, double *bias, hipStream_t s) {
  FRPA_REQUIRE(n_tile >= 1 && n_tile <= T && n_tile <= B2_TMAX, FRPA_ERR_INVALID,
               "bias part 2: n_tile %d outside [1, min(T = %d, %d)]", n_tile, T, B2_TMAX);
  FRPA_REQUIRE(j_coord >= 1 && j_coord <= 3, FRPA_ERR_INVALID, "bias part 2: j_coord %d outside [1, 3]", j_coord);
  FRPA_REQUIRE(j_atom_begin >= 1 && j_atom_begin + n_tile - 1 <= n_atoms, FRPA_ERR_INVALID,
               "bias part 2: atoms %d..%d outside [1, n_atoms = %d]", j_atom_begin, j_atom_begin + n_tile - 1,
               n_atoms);
  if (P.n_valid == 0)
    return;
  FRPA_REQUIRE(avn && wave && gwave && DM && bias, FRPA_ERR_INVALID, "bias part 2: null array");
  hipLaunchKernelGGL(bias2_kernel, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr, avn, n_atoms,
                     j_atom_begin, n_tile, j_coord, wave, gwave, DM, P.n_basis, bias, P.nmb, P.n_batches);
  check_launch("bias2_kernel");
}

} // namespace frpa
