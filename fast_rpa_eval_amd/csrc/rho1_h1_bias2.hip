// Bias part 2 of the first-order H: evaluate_first_order_h_bias_part2_
// (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H_bias_part2.cpp:13-218), a CPU function with one
// dgemm per (batch, atom) there.
//
// For batch b and tile atom j (t = j - j_atom_begin), V_t = its compute rows [avn(j-1, b), avn(j, b)):
//   C(i, p) = w H_times_psi(i, p) + w vrho(p) phi(i, p) + sum_c 2 w vsigma(0, p) grad_rho(c, q) d_c phi(i, p)
//   L(i, k) = sum_p C(i, p) (-d_{j_coord} phi(k, p)),   k in V_t, every compute row i
//   H(ib_i, ib_k, t) += L(i, k),  H(ib_k, ib_i, t) += L(i, k)
// with w = partition_tab(q), q the point's full index.  C does not depend on the atom, so one
// workgroup per batch covers the whole tile: the rows of all V_t are stacked into the columns of one
// product L = C G^T (G = -d_{j_coord} phi of the stacked rows), in 16 x 16 FP64 MFMA tiles over
// points in steps of 4; each lane builds its C and G operands from the wave, gradient wave and
// H_times_psi blocks and the point's prefactors as it goes (every input element is read once per 64
// stacked rows).  Both scatters are FP64 atomics into the caller's H (not reproducible bit for
// bit, like the reference's other GPU reductions; within a few ulps).
#include <algorithm>

#include "rho1_common.hpp"

namespace frpa {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int HB_TMAX = 8;  // atoms per tile
constexpr int HB_RT = 4;    // 16-column tiles of stacked rows per sweep over the points

struct HBiasIn {
  const int *avn;                    // (n_atoms + 1, n_batches)
  const int *map;                    // valid point -> full point
  const double *wave, *gwave, *hpsi; // batch blocks at wave_off (gwave at 3 wave_off)
  const double *vrho, *vsigma;       // (nmb, n_batches), (3, nmb, n_batches)
  const double *grho, *wpart;        // (3, n_full), (n_full)
  int n_atoms, j_atom_begin, n_tile, j_coord, nmb, n_basis;
};

__global__ __launch_bounds__(256) void h1_bias2_kernel(const BatchDesc *__restrict__ descs,
                                                       const int *__restrict__ ibasis0, const HBiasIn in,
                                                       double *__restrict__ H) {
  __shared__ int s_vs[HB_TMAX], s_off[HB_TMAX + 1];
  const BatchDesc d = descs[blockIdx.x];
  const int nc = d.nc, np = d.np, ld = (nc + 3) & ~3;
  const int *__restrict__ ib = ibasis0 + d.ib_off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid == 0) {
    const int *av = in.avn + (size_t)(in.n_atoms + 1) * d.b;
    int o = 0;
    for (int t = 0; t < in.n_tile; t++) {
      const int vs = min(max(av[in.j_atom_begin - 1 + t], 0), nc);
      const int ve = min(max(av[in.j_atom_begin + t], vs), nc);
      s_vs[t] = vs;
      s_off[t] = o;
      o += ve - vs;
    }
    s_off[in.n_tile] = o;
  }
  __syncthreads();
  const int R = s_off[in.n_tile];
  if (R == 0 || np == 0)
    return;
  auto row_of = [&](int r, int &t, int &k) {  // stacked row r -> (tile atom t, compute row k)
    t = 0;
    while (t + 1 < in.n_tile && s_off[t + 1] <= r)
      t++;
    k = s_vs[t] + r - s_off[t];
  };
  const double *__restrict__ W = in.wave + d.wave_off;
  const double *__restrict__ G3 = in.gwave + 3 * d.wave_off;
  const double *__restrict__ Hp = in.hpsi + d.wave_off;
  const double *__restrict__ vr = in.vrho + (size_t)in.nmb * d.b;
  const double *__restrict__ vs = in.vsigma + (size_t)3 * in.nmb * d.b;
  const int jc = in.j_coord - 1;
  const size_t nn = (size_t)in.n_basis * in.n_basis;
  const int nstrip = (nc + 15) >> 4;
  for (int r0 = 0; r0 < R; r0 += 16 * HB_RT) {
    const int nrt = min(HB_RT, (R - r0 + 15) >> 4);
    // this lane's stacked rows (B operand columns): compute row of r0 + 16 rt + lr, or -1
    int kcol[HB_RT];
#pragma unroll
    for (int rt = 0; rt < HB_RT; rt++) {
      const int r = r0 + 16 * rt + lr;
      kcol[rt] = -1;
      if (rt < nrt && r < R) {
        int t, k;
        row_of(r, t, k);
        kcol[rt] = k;
      }
    }
    for (int s = w; s < nstrip; s += 4) {
      const int i = 16 * s + lr;  // A operand row of this lane
      v4d acc[HB_RT];
#pragma unroll
      for (int rt = 0; rt < HB_RT; rt++)
        acc[rt] = v4d{0.0, 0.0, 0.0, 0.0};
      for (int p0 = 0; p0 < np; p0 += 4) {
        const int p = p0 + lk;
        double a = 0.0;
        double b[HB_RT];
        if (p < np) {
          const int q = in.map[d.pstart + p];
          const double wq = in.wpart[q];
          const double f2 = wq * vr[p], f3 = 2.0 * wq * vs[3 * p];
          const double *gq = in.grho + 3 * (size_t)q;
          const double *gp = G3 + (size_t)3 * ld * p;
          if (i < nc)
            a = wq * Hp[i + (size_t)ld * p] + f2 * W[i + (size_t)ld * p] + f3 * gq[0] * gp[i] +
                f3 * gq[1] * gp[ld + i] + f3 * gq[2] * gp[2 * ld + i];
#pragma unroll
          for (int rt = 0; rt < HB_RT; rt++)
            b[rt] = kcol[rt] >= 0 ? -gp[(size_t)ld * jc + kcol[rt]] : 0.0;
        } else {
#pragma unroll
          for (int rt = 0; rt < HB_RT; rt++)
            b[rt] = 0.0;
        }
#pragma unroll
        for (int rt = 0; rt < HB_RT; rt++)
          if (rt < nrt)
            acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[rt], acc[rt], 0, 0, 0);
      }
      // v_mfma_f64_16x16x4f64 leaves D(4 e + lane / 16, lane % 16) in acc[e]:
      // acc[rt][e] = L(i = 16 s + 4 e + lk, stacked row r0 + 16 rt + lr)
#pragma unroll
      for (int rt = 0; rt < HB_RT; rt++) {
        const int r = r0 + 16 * rt + lr;
        if (rt >= nrt || r >= R)
          continue;
        int t, k;
        row_of(r, t, k);
        double *Ht = H + nn * t;
        const size_t bk = (size_t)ib[k];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int ii = 16 * s + 4 * e + lk;
          if (ii < nc) {
            const size_t bi = (size_t)ib[ii];
            atomicAdd(&Ht[bi + (size_t)in.n_basis * bk], acc[rt][e]);
            atomicAdd(&Ht[bk + (size_t)in.n_basis * bi], acc[rt][e]);
          }
        }
      }
    }
  }
}

} // namespace

void rho1_h1_bias_part2(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                        const double *wave, const double *gwave, const double *hpsi, const double *vrho,
                        const double *vsigma, const double *gradient_rho, const double *partition_tab, double *H,
                        hipStream_t s) {
  FRPA_REQUIRE(n_tile >= 1 && n_tile <= T && n_tile <= HB_TMAX, FRPA_ERR_INVALID,
               "H bias part 2: n_tile %d outside [1, min(T = %d, %d)]", n_tile, T, HB_TMAX);
  FRPA_REQUIRE(j_coord >= 1 && j_coord <= 3, FRPA_ERR_INVALID, "H bias part 2: j_coord %d outside [1, 3]", j_coord);
  FRPA_REQUIRE(j_atom_begin >= 1 && j_atom_begin + n_tile - 1 <= n_atoms, FRPA_ERR_INVALID,
               "H bias part 2: atoms %d..%d outside [1, n_atoms = %d]", j_atom_begin, j_atom_begin + n_tile - 1,
               n_atoms);
  if (P.n_valid == 0)
    return;
  FRPA_REQUIRE(avn && wave && gwave && hpsi && vrho && vsigma && gradient_rho && partition_tab && H, FRPA_ERR_INVALID,
               "H bias part 2: null array");
  const HBiasIn in{avn, P.d_map.ptr, wave, gwave, hpsi, vrho, vsigma, gradient_rho, partition_tab,
                   n_atoms, j_atom_begin, n_tile, j_coord, P.nmb, P.n_basis};
  hipLaunchKernelGGL(h1_bias2_kernel, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr, in, H);
  check_launch("h1_bias2_kernel");
}

} // namespace frpa
