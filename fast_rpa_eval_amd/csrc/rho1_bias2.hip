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
// One workgroup per batch covers the whole atom tile.  The rows of all V_t (disjoint, at most nc; a
// single atom's functions in most batches) are stacked into 16-row MFMA tiles, up to two per sweep
// over K.  The batch's points are split in 16-point groups dealt round-robin to the four waves, so
// every wave streams its own points' phi straight from global memory into MFMA operands: no LDS
// staging and no workgroup barrier in the K loop, operands of chunk c + 1 loaded while chunk c is
// multiplied.  K is visited in a permuted order (lane quad lk of k-step s takes column
// 16 c + 4 lk + s) so that each lane reads four consecutive phi values of one point (one 32-byte
// load when the wave blocks start on 4-double boundaries).  W = A phi lives in
// v_mfma_f64_16x16x4f64 accumulators and the product with phi' is the epilogue, summed per tile atom
// across lanes with shuffles into wave-private LDS columns (deterministic).  FP64 throughout.
//
// The A operand (DM + DM^T)(i_basis(V), i_basis(:)): per call, the basis functions that occur as
// rows of the tile are marked and numbered (b2_mark_kernel, b2_scan_kernel) and
// tab(l, c) = DM(row_l, c) + DM(c, row_l) is written for all columns c (b2_table_kernel, an LDS
// transpose; R x n_basis, a few MB), so the GEMM's A gathers are one load of 16 consecutive rows
// instead of two gathers through the n_basis^2 matrix (whose transposed half touches one line per
// lane).  When the tile has more rows than the table holds, the kernel gathers from DM directly;
// both give the same sums bit for bit.
//
// The same kernels compute the inline bias of the v6 variant of the rho1 function
// (task4_common_calculation_eliminated/version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:
// 179-231): acc_bias(p, t) = sum_{k in V_t} phi'(k, p) sum_i DM(k, i) phi(i, p), the one-sided matrix
// (sides = 1: A(k, i) = DM(k, i) only), ADDED to the bias array it is given (accumulate), so that a
// copy of the bias input becomes the input of the forward.
#include <algorithm>
#include <cstdlib>

#include "rho1_common.hpp"

namespace frpa {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int B2_TMAX = 8;               // atoms per tile
constexpr int B2_KC = 16;                // K chunk
constexpr int B2_RT = 2;                 // 16-row tiles per sweep over K
constexpr int B2_GW = 4;                 // 16-point groups per wave per pass
constexpr int B2_PASS = 4 * B2_GW * 16;  // points per pass (256)
constexpr int B2_TAB_ROWS = 256;         // default row capacity of the per-call table

// Compute rows [vs, ve) of tile atom t in batch d, clamped to the batch's rows (a malformed prefix
// cannot index outside its phi block); the marking pass and the GEMM use the same ranges.
__device__ __forceinline__ void tile_rows(const BatchDesc &d, const int *__restrict__ avn, int n_atoms,
                                          int j_atom_begin, int t, int &vs, int &ve) {
  const int *av = avn + (size_t)(n_atoms + 1) * d.b;
  vs = min(max(av[j_atom_begin - 1 + t], 0), d.nc);
  ve = min(max(av[j_atom_begin + t], vs), d.nc);
}

// rowmap[mu] = 1 for every basis function mu that is a row of the tile in some batch (rowmap zeroed).
__global__ __launch_bounds__(64) void b2_mark_kernel(const BatchDesc *__restrict__ descs,
                                                    const int *__restrict__ ibasis0, const int *__restrict__ avn,
                                                    int n_atoms, int j_atom_begin, int n_tile, int *rowmap) {
  const BatchDesc d = descs[blockIdx.x];
  const int *__restrict__ ib = ibasis0 + d.ib_off;
  for (int t = 0; t < n_tile; t++) {
    int vs, ve;
    tile_rows(d, avn, n_atoms, j_atom_begin, t, vs, ve);
    for (int k = vs + (int)threadIdx.x; k < ve; k += 64)
      rowmap[ib[k]] = 1;
  }
}

// Numbers the marked functions in ascending order: rowmap[mu] = local row or -1, rows[0] = R,
// rows[1 + l] = the basis function of local row l.  One workgroup.
__global__ __launch_bounds__(1024) void b2_scan_kernel(int *__restrict__ rowmap, int *__restrict__ rows,
                                                      int n_basis) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int per = (n_basis + 1023) / 1024;
  const int lo = min(tid * per, n_basis), hi = min(lo + per, n_basis);
  int cnt = 0;
  for (int r = lo; r < hi; r++)
    cnt += rowmap[r] > 0;
  part[tid] = cnt;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int l = part[tid] - cnt;
  for (int r = lo; r < hi; r++) {
    if (rowmap[r] > 0) {
      rowmap[r] = l;
      rows[1 + l] = r;
      l++;
    } else {
      rowmap[r] = -1;
    }
  }
  if (tid == 1023)
    rows[0] = part[1023];
}

// tab[c * cap + l] = DM(rows_l, c) + DM(c, rows_l) (sides = 2) or DM(rows_l, c) (sides = 1) for l < R,
// every column c; 64 x 64 tiles, the DM(c, rows_l) half read along c and transposed through LDS.
// Nothing is written when R > cap.
__global__ __launch_bounds__(256) void b2_table_kernel(const double *__restrict__ DM, int n_basis,
                                                      const int *__restrict__ rows, double *__restrict__ tab,
                                                      int cap, int sides) {
  __shared__ double sT[64][65];
  const int R = rows[0];
  const int c0 = blockIdx.x * 64, l0 = blockIdx.y * 64;
  if (R > cap || l0 >= R)
    return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int li = ty; li < 64; li += 4) {
    const int l = l0 + li, c = c0 + tx;
    sT[li][tx] = (sides == 2 && l < R && c < n_basis) ? DM[c + (size_t)n_basis * rows[1 + l]] : 0.0;
  }
  __syncthreads();
  for (int ci = ty; ci < 64; ci += 4) {
    const int c = c0 + ci, l = l0 + tx;
    if (l < R && c < n_basis)
      tab[(size_t)c * cap + l] = sides == 2 ? DM[rows[1 + l] + (size_t)n_basis * c] + sT[tx][ci]
                                            : DM[rows[1 + l] + (size_t)n_basis * c];
  }
}

// QUADS: every wave block starts on a 4-double boundary, so a lane's four phi values are one load.
// sides: 2 for DM + DM^T, 1 for DM; accumulate: add the sums to bias instead of storing them.
template <bool QUADS>
__global__ __launch_bounds__(256, 2) void bias2_kernel(const BatchDesc *__restrict__ descs,
                                                   const int *__restrict__ ibasis0, const int *__restrict__ avn,
                                                   int n_atoms, int j_atom_begin, int n_tile, int j_coord,
                                                   const double *__restrict__ wave, const double *__restrict__ gwave,
                                                   const double *__restrict__ DM, int n_basis,
                                                   const double *__restrict__ tab, const int *__restrict__ rowmap,
                                                   const int *__restrict__ rows, int cap, int sides,
                                                   int accumulate, double *__restrict__ bias, int nmb,
                                                   int n_batches_all) {
  __shared__ double red[B2_TMAX][B2_PASS];
  __shared__ int s_vs[B2_TMAX], s_off[B2_TMAX + 1];
  const BatchDesc d = descs[blockIdx.x];
  const int nc = d.nc, ld = (nc + 3) & ~3;
  const int *__restrict__ ib = ibasis0 + d.ib_off;
  const double *__restrict__ W = wave + d.wave_off;
  const double *__restrict__ G = gwave + 3 * d.wave_off + (size_t)ld * (j_coord - 1);  // component j_coord
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid == 0) {
    int o = 0;
    for (int t = 0; t < n_tile; t++) {
      int vs, ve;
      tile_rows(d, avn, n_atoms, j_atom_begin, t, vs, ve);
      s_vs[t] = vs;
      s_off[t] = o;
      o += ve - vs;
    }
    s_off[n_tile] = o;
  }
  __syncthreads();
  const int nR = s_off[n_tile];
  const bool use_tab = tab != nullptr && rows[0] <= cap;
  // stacked row r (< nR) -> (tile atom t, compute row k)
  auto row_of = [&](int r, int &t, int &k) {
    t = 0;
    while (t + 1 < n_tile && s_off[t + 1] <= r)
      t++;
    k = s_vs[t] + r - s_off[t];
  };
  const int nchunk = (nc + B2_KC - 1) / B2_KC;

  for (int pc = 0; pc < d.np; pc += B2_PASS) {
    const int npp = min(B2_PASS, d.np - pc);
    const int ng = (npp + 15) >> 4;
    const int gw = ng > w ? (ng - w + 3) >> 2 : 0;  // this wave's groups: w, w + 4, ...
    // this wave's points of the pass: 16 (w + 4 q) + lr, q < gw; their red columns are its own
    for (int t = lk; t < n_tile; t += 4)
      for (int q = 0; q < gw; q++)
        red[t][16 * (w + 4 * q) + lr] = 0.0;
    __syncthreads();

    for (int r0 = 0; r0 < nR && gw > 0; r0 += 16 * B2_RT) {
      const int nrt = min(B2_RT, (nR - r0 + 15) >> 4);
      // A rows of this lane: stacked rows r0 + 16 rt + lr (basis function, table row)
      long long arow[B2_RT];
      int trow[B2_RT];
#pragma unroll
      for (int rt = 0; rt < B2_RT; rt++) {
        const int r = r0 + 16 * rt + lr;
        arow[rt] = -1;
        trow[rt] = 0;
        if (r < nR) {
          int t, k;
          row_of(r, t, k);
          arow[rt] = ib[k];
          if (use_tab)
            trow[rt] = max(rowmap[arow[rt]], 0);  // marked by construction
        }
      }
      // B columns of this lane: points pc + 16 (w + 4 q) + lr
      const double *bcol[B2_GW];
      bool pok[B2_GW];
#pragma unroll
      for (int q = 0; q < B2_GW; q++) {
        const int p = 16 * (w + 4 * q) + lr;
        pok[q] = q < gw && p < npp;
        bcol[q] = W + (size_t)ld * (pc + (pok[q] ? p : 0));
      }
      v4d acc[B2_RT][B2_GW];
#pragma unroll
      for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
        for (int q = 0; q < B2_GW; q++)
          acc[rt][q] = v4d{0.0, 0.0, 0.0, 0.0};

      double an[B2_RT][4], bn[B2_GW][4];
      int jn[4];  // basis indices of chunk c + 1's columns of this lane, -1 past nc
      auto load_ib = [&](int c, int *jj) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int j = B2_KC * c + 4 * lk + s;
          jj[s] = j < nc ? ib[j] : -1;
        }
      };
      auto load_ab = [&](int c, const int *jj) {
        if (use_tab) {
#pragma unroll
          for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
              const long long jb = jj[s];
              an[rt][s] = (rt < nrt && arow[rt] >= 0 && jb >= 0) ? tab[jb * cap + trow[rt]] : 0.0;
            }
        } else {
#pragma unroll
          for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
              const long long jb = jj[s];
              an[rt][s] = (rt < nrt && arow[rt] >= 0 && jb >= 0)
                              ? (sides == 2 ? DM[arow[rt] + (size_t)n_basis * jb] + DM[jb + (size_t)n_basis * arow[rt]]
                                            : DM[arow[rt] + (size_t)n_basis * jb])
                              : 0.0;
            }
        }
        const int j0 = B2_KC * c + 4 * lk;
#pragma unroll
        for (int q = 0; q < B2_GW; q++) {
          if (QUADS) {
            // j0 < nc: the quad lies inside the batch's padded column (ld = nc rounded up to 4)
            v4d v = v4d{0.0, 0.0, 0.0, 0.0};
            if (pok[q] && j0 < nc)
              v = *reinterpret_cast<const v4d *>(bcol[q] + j0);
#pragma unroll
            for (int s = 0; s < 4; s++)
              bn[q][s] = j0 + s < nc ? v[s] : 0.0;
          } else {
#pragma unroll
            for (int s = 0; s < 4; s++)
              bn[q][s] = (pok[q] && j0 + s < nc) ? bcol[q][j0 + s] : 0.0;
          }
        }
      };
      {
        int j0[4];
        load_ib(0, j0);
        load_ab(0, j0);
        if (nchunk > 1)
          load_ib(1, jn);
      }
      for (int c = 0; c < nchunk; c++) {
        double a[B2_RT][4], b[B2_GW][4];
#pragma unroll
        for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
          for (int s = 0; s < 4; s++)
            a[rt][s] = an[rt][s];
#pragma unroll
        for (int q = 0; q < B2_GW; q++)
#pragma unroll
          for (int s = 0; s < 4; s++)
            b[q][s] = bn[q][s];
        if (c + 1 < nchunk) {
          int jj[4];
#pragma unroll
          for (int s = 0; s < 4; s++)
            jj[s] = jn[s];
          load_ab(c + 1, jj);
          if (c + 2 < nchunk)
            load_ib(c + 2, jn);
        }
#pragma unroll
        for (int q = 0; q < B2_GW; q++) {
          if (q < gw) {
#pragma unroll
            for (int rt = 0; rt < B2_RT; rt++) {
              if (rt < nrt) {
#pragma unroll
                for (int s = 0; s < 4; s++)
                  acc[rt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][s], b[q][s], acc[rt][q], 0, 0, 0);
              }
            }
          }
        }
      }
      // acc[rt][q][i] = W(stacked row r0 + 16 rt + 4 i + lk, point pc + 16 (w + 4 q) + lr)
      int rtt[B2_RT][4], rkk[B2_RT][4];
#pragma unroll
      for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int rr = r0 + 16 * rt + 4 * i + lk;
          rtt[rt][i] = -1;
          rkk[rt][i] = 0;
          if (rr < nR)
            row_of(rr, rtt[rt][i], rkk[rt][i]);
        }
      int t_lo, t_hi, kdummy;
      row_of(r0, t_lo, kdummy);
      row_of(min(r0 + 16 * B2_RT, nR) - 1, t_hi, kdummy);
#pragma unroll
      for (int q = 0; q < B2_GW; q++) {
        if (q >= gw)
          continue;
        const int p = 16 * (w + 4 * q) + lr;
        double v[B2_RT][4];
#pragma unroll
        for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
          for (int i = 0; i < 4; i++)
            v[rt][i] = (rtt[rt][i] >= 0 && p < npp) ? -acc[rt][q][i] * G[rkk[rt][i] + (size_t)3 * ld * (pc + p)] : 0.0;
        for (int t = t_lo; t <= t_hi; t++) {
          double part = 0.0;
#pragma unroll
          for (int rt = 0; rt < B2_RT; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
              part += rtt[rt][i] == t ? v[rt][i] : 0.0;
          part += __shfl_xor(part, 16);
          part += __shfl_xor(part, 32);
          if (lk == 0)
            red[t][p] += part;
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < n_tile * npp; e += 256) {
      const int t = e / npp, p = e % npp;
      double *o = bias + (size_t)(pc + p) + (size_t)nmb * (d.b + (size_t)n_batches_all * t);
      *o = accumulate ? *o + red[t][p] : red[t][p];
    }
    __syncthreads();
  }
}

} // namespace

void rho1_bias_part2(Rho1Plan &P, int T, int j_atom_begin, int n_tile, int j_coord, int n_atoms, const int *avn,
                     const double *wave, const double *gwave, const double *DM, double *bias, hipStream_t s,
                     int sides, bool accumulate) {
  FRPA_REQUIRE(sides == 1 || sides == 2, FRPA_ERR_INVALID, "bias: sides %d is not 1 or 2", sides);
  FRPA_REQUIRE(n_tile >= 1 && n_tile <= T && n_tile <= B2_TMAX, FRPA_ERR_INVALID,
               "bias part 2: n_tile %d outside [1, min(T = %d, %d)]", n_tile, T, B2_TMAX);
  FRPA_REQUIRE(j_coord >= 1 && j_coord <= 3, FRPA_ERR_INVALID, "bias part 2: j_coord %d outside [1, 3]", j_coord);
  FRPA_REQUIRE(j_atom_begin >= 1 && j_atom_begin + n_tile - 1 <= n_atoms, FRPA_ERR_INVALID,
               "bias part 2: atoms %d..%d outside [1, n_atoms = %d]", j_atom_begin, j_atom_begin + n_tile - 1,
               n_atoms);
  if (P.n_valid == 0)
    return;
  FRPA_REQUIRE(avn && wave && gwave && DM && bias, FRPA_ERR_INVALID, "bias part 2: null array");
  // row capacity of the table: FRPA_BIAS2_TABLE_ROWS (0 = gather from DM directly), a multiple of 64,
  // at most n_basis rounded up and at most 1 GiB of table
  int cap = B2_TAB_ROWS;
  if (const char *e = std::getenv("FRPA_BIAS2_TABLE_ROWS"))
    cap = std::atoi(e);
  cap = std::min(std::max(cap, 0), (P.n_basis + 63) / 64 * 64);
  cap = (cap + 63) / 64 * 64;
  while (cap > 0 && (double)cap * P.n_basis * sizeof(double) > (double)(1ull << 30))
    cap -= 64;
  const double *tab = nullptr;
  if (cap > 0) {
    P.d_b2_rowmap.alloc(P.n_basis);
    P.d_b2_rows.alloc(1 + (size_t)P.n_basis);
    P.d_b2_tab.alloc((size_t)cap * P.n_basis);
    P.b2_cap = cap;
    fill_i32(P.d_b2_rowmap.ptr, P.n_basis, 0, s);
    hipLaunchKernelGGL(b2_mark_kernel, dim3(P.n_valid), dim3(64), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr, avn, n_atoms,
                       j_atom_begin, n_tile, P.d_b2_rowmap.ptr);
    check_launch("b2_mark_kernel");
    hipLaunchKernelGGL(b2_scan_kernel, dim3(1), dim3(1024), 0, s, P.d_b2_rowmap.ptr, P.d_b2_rows.ptr, P.n_basis);
    check_launch("b2_scan_kernel");
    hipLaunchKernelGGL(b2_table_kernel, dim3((P.n_basis + 63) / 64, cap / 64), dim3(256), 0, s, DM, P.n_basis,
                       P.d_b2_rows.ptr, P.d_b2_tab.ptr, cap, sides);
    check_launch("b2_table_kernel");
    tab = P.d_b2_tab.ptr;
  }
  const bool quads = P.wave_off_quads && reinterpret_cast<uintptr_t>(wave) % (4 * sizeof(double)) == 0;
  auto kern = quads ? bias2_kernel<true> : bias2_kernel<false>;
  hipLaunchKernelGGL(kern, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr, avn, n_atoms,
                     j_atom_begin, n_tile, j_coord, wave, gwave, DM, P.n_basis, tab, P.d_b2_rowmap.ptr,
                     P.d_b2_rows.ptr, cap, sides, accumulate ? 1 : 0, bias, P.nmb, P.n_batches);
  check_launch("bias2_kernel");
}

} // namespace frpa
