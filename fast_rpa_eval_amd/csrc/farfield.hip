// Far-field Hartree multipole sum on gfx950 (FP64 throughout).
//
// Reference: far_distance_real_hartree_potential_single_atom_p2_fused_fp_cu and its helpers
// (fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.cu:62-166,373-543) -- one thread per point,
// serial over centres, l_pot_max fixed to 6, coefficients read from global memory.
//
// Layout here: one lane per valid point (coordinates gathered once through the valid -> full map);
// the loop over hartree centres is wave-uniform, so each centre's record and coefficient block are
// fetched with scalar loads and fed to v_fma_f64 as SGPR operands (no LDS, no per-lane loads).
// A pair's l_max depends on the point (outer_potential_radius reduction); lanes are grouped by their
// l with ballots and each group runs code specialised for that l (AOT-style unrolling, l <= 10).
//
// Two formulations of the same sum:
//  * FRPA_FAR_TERMS: the reference's term-by-term Cartesian dot
//        sum_n x^i y^j z^k Fp(nn) multipole_c(n)
//    with the exponents compiled in (the plain FP64 baseline).
//  * FRPA_FAR_HORNER: terms sharing (i, j, k, nn) merged per atom once per call; with
//    u = 1/r^2 and X = x u, Y = y u, Z = z u the pair sum is
//        (1/r) sum_q u^q Q_q(X, Y, Z),   Q_q of degree l - 2q in X, Y, Z (X degree <= 1),
//    evaluated by nested Horner: one FMA per merged coefficient (165 at l = 8 vs 425 terms).
#include <cstdlib>
#include <mutex>
#include <utility>

#include "farfield_common.hpp"

// Coefficient table of frpa_far_forward_noalloc: a fixed array in every device's memory (held until
// the process ends), used in place of the plan's per-call buffer so that those calls allocate
// nothing.  Host code never passes this name to a kernel (on the host it is a stand-in, not a
// device address); far_scratch() returns the device's own address of it.
static __device__ __align__(256) double scratch[FRPA_FAR_NOALLOC_SCRATCH_DOUBLES];

namespace frpa {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------------------------
// Per-call coefficient preparation
// ---------------------------------------------------------------------------------------------

// multipole_c(n, a, t) of sum_up_whole_potential.cpp:358-378 / init_multipole_c_cu (.cu:501-543).
// Entries n >= n_cc_lm_ijk(lmfd(a, t)) are set to 0 (the reference leaves them unset).
__global__ void far_init_mc_kernel(int n_atoms, int mc_dim0, int dim0, int l_pot_max, int lama,
                                   const int *__restrict__ index_cc, const double *__restrict__ cc,
                                   const int *__restrict__ ncc, const int *__restrict__ lmfd,
                                   const double *__restrict__ M, double *__restrict__ mc) {
  const int a = blockIdx.x, t = blockIdx.y;
  const int lm = clampi(lmfd[a + n_atoms * t], 0, lama);
  const int nlim = ncc[lm];
  const int lpp = (l_pot_max + 1) * (l_pot_max + 1);
  for (int n = threadIdx.x; n < mc_dim0; n += blockDim.x) {
    double v = 0.0;
    if (n < nlim) {
      const int l = index_cc[n], m = index_cc[n + dim0];
      if (l <= lm && l <= l_pot_max) {
        const double Mv = M[(l * l + l + m) + (size_t)lpp * (a + (size_t)n_atoms * t)];
        if (fabs(Mv) > 1e-10)
          v = cc[n] * Mv;
      }
    }
    mc[n + (size_t)mc_dim0 * (a + (size_t)n_atoms * t)] = v;
  }
}

// Merged nested-Horner coefficients: entry e of the (a, t) block is sum over the listed terms n of
// (-1)^nn (2nn-1)!! cc(n) M(lm(n), a, t), with the same masks as multipole_c.
__global__ void far_init_horner_kernel(int n_atoms, int cstride, int dim0, int l_pot_max, int lama,
                                       const int *__restrict__ index_cc, const int *__restrict__ ncc,
                                       const int *__restrict__ lmfd, const double *__restrict__ M,
                                       const int *__restrict__ hc_ptr, const int *__restrict__ hc_term,
                                       const double *__restrict__ hc_w, double *__restrict__ coef) {
  const int a = blockIdx.x, t = blockIdx.y;
  const int lm = clampi(lmfd[a + n_atoms * t], 0, lama);
  const int nlim = ncc[lm];
  const int lpp = (l_pot_max + 1) * (l_pot_max + 1);
  const double *Ma = M + (size_t)lpp * (a + (size_t)n_atoms * t);
  for (int e = threadIdx.x; e < cstride; e += blockDim.x) {
    double s = 0.0;
    for (int q = hc_ptr[e]; q < hc_ptr[e + 1]; q++) {
      const int n = hc_term[q];
      if (n >= nlim)
        continue;
      const int l = index_cc[n], m = index_cc[n + dim0];
      if (l > lm || l > l_pot_max)
        continue;
      const double Mv = Ma[l * l + l + m];
      if (fabs(Mv) > 1e-10)
        s += hc_w[q] * Mv;
    }
    coef[e + (size_t)cstride * (a + (size_t)n_atoms * t)] = s;
  }
}

// ---------------------------------------------------------------------------------------------
// Pair evaluators (compile-time l)
// ---------------------------------------------------------------------------------------------

template <int L, int Q, int I, int J>
__device__ __forceinline__ double hz(const double *__restrict__ c, double Z) {
  constexpr int K = L - 2 * Q - I - J;
  double s = c[horner_index(L, Q, I, J, K)];
#pragma unroll
  for (int k = K - 1; k >= 0; k--)
    s = fma(s, Z, c[horner_index(L, Q, I, J, k)]);
  return s;
}
template <int L, int Q, int I, int J>
__device__ __forceinline__ double hy(const double *__restrict__ c, double Y, double Z) {
  if constexpr (J == L - 2 * Q - I)
    return hz<L, Q, I, J>(c, Z);
  else
    return fma(hy<L, Q, I, J + 1>(c, Y, Z), Y, hz<L, Q, I, J>(c, Z));
}
template <int L, int Q>
__device__ __forceinline__ double hq(const double *__restrict__ c, double X, double Y, double Z) {
  if constexpr (L - 2 * Q == 0)
    return hy<L, Q, 0, 0>(c, Y, Z);
  else
    return fma(hy<L, Q, 1, 0>(c, Y, Z), X, hy<L, Q, 0, 0>(c, Y, Z));
}
template <int L, int Q>
__device__ __forceinline__ double hu(const double *__restrict__ c, double X, double Y, double Z, double u) {
  if constexpr (2 * (Q + 1) > L)
    return hq<L, Q>(c, X, Y, Z);
  else
    return fma(hu<L, Q + 1>(c, X, Y, Z, u), u, hq<L, Q>(c, X, Y, Z));
}

// sum_n x^i y^j z^k Fp(nn) mc(n) over n < n_cc(L), reference multiplication order
template <int N>
__device__ __forceinline__ void term_acc(double &acc, const double *px, const double *py, const double *pz,
                                         const double *Fp, const double *__restrict__ mc) {
  constexpr FarTerm T = FAR_TERMS[N];
  acc += px[T.i] * py[T.j] * pz[T.k] * Fp[T.nn] * mc[N];
}
template <int... Ns>
__device__ __forceinline__ double terms_dot(std::integer_sequence<int, Ns...>, const double *px, const double *py,
                                            const double *pz, const double *Fp, const double *__restrict__ mc) {
  double acc = 0.0;
  (term_acc<Ns>(acc, px, py, pz, Fp, mc), ...);
  return acc;
}

template <int L>
__device__ __forceinline__ double pair_terms(const double *__restrict__ mc, double mx, double my, double mz,
                                             double d2) {
  // far_distance_hartree_Fp_cluster_single_atom_p2 (sum_up_whole_potential.cpp:141-149)
  const double dist = sqrt(d2);
  const double dsq = dist * dist;
  double Fp[L + 1];
  Fp[0] = 1.0 / dist;
#pragma unroll
  for (int k = 1; k <= L; k++)
    Fp[k] = Fp[k - 1] * (double)(1 - 2 * k) / dsq;
  double px[2] = {1.0, mx};
  double py[L + 1], pz[L + 1];
  py[0] = pz[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= L; k++) {
    py[k] = my * py[k - 1];
    pz[k] = mz * pz[k - 1];
  }
  return terms_dot(std::make_integer_sequence<int, FAR_NCC[L]>{}, px, py, pz, Fp, mc);
}

template <int L>
__device__ __forceinline__ double pair_horner(const double *__restrict__ c, double mx, double my, double mz,
                                              double d2) {
  const double u = 1.0 / d2;
  const double rinv = sqrt(u);
  return hu<L, 0>(c + horner_prefix(L), mx * u, my * u, mz * u, u) * rinv;
}

template <int VARIANT, int L>
__device__ __forceinline__ double pair_eval(const double *__restrict__ c, double mx, double my, double mz, double d2) {
  if constexpr (VARIANT == FRPA_FAR_HORNER)
    return pair_horner<L>(c, mx, my, mz, d2);
  else
    return pair_terms<L>(c, mx, my, mz, d2);
}

template <int VARIANT, int... Ls>
__device__ __forceinline__ double pair_switch(std::integer_sequence<int, Ls...>, int L, const double *__restrict__ c,
                                              double mx, double my, double mz, double d2) {
  double v = 0.0;
  ((L == Ls ? (void)(v = pair_eval<VARIANT, Ls>(c, mx, my, mz, d2)) : (void)0), ...);
  return v;
}

// ---------------------------------------------------------------------------------------------
// Forward kernel
// ---------------------------------------------------------------------------------------------

template <int VARIANT, int LMAXK>
__global__ __launch_bounds__(256) void far_forward_kernel(
    const CenterRec *__restrict__ centers, int n_hc, const double *__restrict__ cp, const int *__restrict__ map,
    int n_valid, const double *__restrict__ opr, int lp1, const double *__restrict__ mrsq, int n_atoms,
    const double *__restrict__ coefs, int cstride, double *__restrict__ dv, int n_full) {
  const int iv = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  const bool active = iv < n_valid;
  const int p = active ? map[iv] : 0;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (active) {
    rx = cp[3 * (size_t)p];
    ry = cp[3 * (size_t)p + 1];
    rz = cp[3 * (size_t)p + 2];
  }
  double V = 0.0;
  for (int ic = 0; ic < n_hc; ic++) {
    const CenterRec c = centers[ic];
    const double ax = rx - c.dx, ay = ry - c.dy, az = rz - c.dz;
    const double d2 = ax * ax + ay * ay + az * az;
    const double *op = opr + (size_t)lp1 * (c.atom + (size_t)n_atoms * t);
    int l = c.l_const;
    while (op[l] < d2 && l > 0)
      l--;
    const bool far = active && !(d2 < mrsq[c.atom + (size_t)n_atoms * t]) && d2 < 1e8;
    unsigned long long pending = __ballot(far);
    if (!pending)
      continue;
    const double mx = rx - c.mx, my = ry - c.my, mz = rz - c.mz;
    const double *cb = coefs + (size_t)cstride * (c.mc_atom + (size_t)n_atoms * t);
    while (pending) {
      const int lead = __ffsll((long long)pending) - 1;
      const int L = __builtin_amdgcn_readlane(l, lead);
      const bool mine = far && l == L;
      pending &= ~__ballot(mine);
      if (mine)
        V += pair_switch<VARIANT>(std::make_integer_sequence<int, LMAXK + 1>{}, L, cb, mx, my, mz, d2);
    }
  }
  if (active)
    dv[(size_t)p + (size_t)n_full * t] += V;
}

// ---------------------------------------------------------------------------------------------
// Horner forward with several points per lane: one scalar load of a coefficient feeds PPL
// independent FMA chains (the coefficient block of one centre, 165 doubles at l = 8, does not fit
// the SGPRs, so the single-point kernel waits on scalar loads between FMAs).
// ---------------------------------------------------------------------------------------------

template <int N>
struct DV {
  double v[N];
};
template <int N>
__device__ __forceinline__ DV<N> splat(double x) {
  DV<N> r;
#pragma unroll
  for (int q = 0; q < N; q++)
    r.v[q] = x;
  return r;
}
template <int N>
__device__ __forceinline__ DV<N> fmav(const DV<N> &a, const DV<N> &b, const DV<N> &c) {
  DV<N> r;
#pragma unroll
  for (int q = 0; q < N; q++)
    r.v[q] = fma(a.v[q], b.v[q], c.v[q]);
  return r;
}
template <int N>
__device__ __forceinline__ DV<N> fmas(const DV<N> &a, const DV<N> &b, double c) {
  DV<N> r;
#pragma unroll
  for (int q = 0; q < N; q++)
    r.v[q] = fma(a.v[q], b.v[q], c);
  return r;
}

template <int N, int L, int Q, int I, int J>
__device__ __forceinline__ DV<N> hzv(const double *__restrict__ c, const DV<N> &Z) {
  constexpr int K = L - 2 * Q - I - J;
  DV<N> s = splat<N>(c[horner_index(L, Q, I, J, K)]);
#pragma unroll
  for (int k = K - 1; k >= 0; k--)
    s = fmas(s, Z, c[horner_index(L, Q, I, J, k)]);
  return s;
}
template <int N, int L, int Q, int I, int J>
__device__ __forceinline__ DV<N> hyv(const double *__restrict__ c, const DV<N> &Y, const DV<N> &Z) {
  if constexpr (J == L - 2 * Q - I)
    return hzv<N, L, Q, I, J>(c, Z);
  else
    return fmav(hyv<N, L, Q, I, J + 1>(c, Y, Z), Y, hzv<N, L, Q, I, J>(c, Z));
}
template <int N, int L, int Q>
__device__ __forceinline__ DV<N> hqv(const double *__restrict__ c, const DV<N> &X, const DV<N> &Y, const DV<N> &Z) {
  if constexpr (L - 2 * Q == 0)
    return hyv<N, L, Q, 0, 0>(c, Y, Z);
  else
    return fmav(hyv<N, L, Q, 1, 0>(c, Y, Z), X, hyv<N, L, Q, 0, 0>(c, Y, Z));
}
template <int N, int L, int Q>
__device__ __forceinline__ DV<N> huv(const double *__restrict__ c, const DV<N> &X, const DV<N> &Y, const DV<N> &Z,
                                     const DV<N> &u) {
  if constexpr (2 * (Q + 1) > L)
    return hqv<N, L, Q>(c, X, Y, Z);
  else
    return fmav(huv<N, L, Q + 1>(c, X, Y, Z, u), u, hqv<N, L, Q>(c, X, Y, Z));
}

// 1 / sqrt(x) from v_rsq_f64 and two Newton steps r += r (1 - x r^2) / 2 (quadratic convergence:
// within a few ulp of the correctly rounded value, for the x > 0 of far pairs)
template <int N, int L>
__device__ __forceinline__ DV<N> pair_horner_v(const double *__restrict__ c, const DV<N> &mx, const DV<N> &my,
                                               const DV<N> &mz, const DV<N> &d2) {
  DV<N> u, X, Y, Z, rinv;
#pragma unroll
  for (int q = 0; q < N; q++) {
    rinv.v[q] = rsqrt_f64(d2.v[q]);
    u.v[q] = rinv.v[q] * rinv.v[q];
    X.v[q] = mx.v[q] * u.v[q];
    Y.v[q] = my.v[q] * u.v[q];
    Z.v[q] = mz.v[q] * u.v[q];
  }
  DV<N> s = huv<N, L, 0>(c + horner_prefix(L), X, Y, Z, u);
#pragma unroll
  for (int q = 0; q < N; q++)
    s.v[q] *= rinv.v[q];
  return s;
}

template <int N, int... Ls>
__device__ __forceinline__ DV<N> pair_switch_v(std::integer_sequence<int, Ls...>, int L, const double *__restrict__ c,
                                               const DV<N> &mx, const DV<N> &my, const DV<N> &mz, const DV<N> &d2) {
  DV<N> v = splat<N>(0.0);
  ((L == Ls ? (void)(v = pair_horner_v<N, Ls>(c, mx, my, mz, d2)) : (void)0), ...);
  return v;
}

// Point q of lane `tid` of block b is valid point b * 256 N + 256 q + tid.
template <int LMAXK, int N>
__global__ __launch_bounds__(256) void far_horner_kernel(
    const CenterRec *__restrict__ centers, int n_hc, const double *__restrict__ cp, const int *__restrict__ map,
    int n_valid, const double *__restrict__ opr, int lp1, const double *__restrict__ mrsq, int n_atoms,
    const double *__restrict__ coefs, int cstride, double *__restrict__ dv, int n_full) {
  const int t = blockIdx.y;
  bool active[N];
  int p[N];
  double rx[N], ry[N], rz[N], V[N];
#pragma unroll
  for (int q = 0; q < N; q++) {
    const int iv = blockIdx.x * 256 * N + 256 * q + threadIdx.x;
    active[q] = iv < n_valid;
    p[q] = active[q] ? map[iv] : 0;
    rx[q] = active[q] ? cp[3 * (size_t)p[q]] : 0.0;
    ry[q] = active[q] ? cp[3 * (size_t)p[q] + 1] : 0.0;
    rz[q] = active[q] ? cp[3 * (size_t)p[q] + 2] : 0.0;
    V[q] = 0.0;
  }
  for (int ic = 0; ic < n_hc; ic++) {
    const CenterRec c = centers[ic];
    const double *op = opr + (size_t)lp1 * (c.atom + (size_t)n_atoms * t);
    const double rsq = mrsq[c.atom + (size_t)n_atoms * t];
    const double op_top = op[c.l_const];  // wave-uniform: one scalar load per centre
    DV<N> d2;
    int l[N];
    bool far[N];
    unsigned long long pend[N], any = 0;
#pragma unroll
    for (int q = 0; q < N; q++) {
      const double ax = rx[q] - c.dx, ay = ry[q] - c.dy, az = rz[q] - c.dz;
      d2.v[q] = ax * ax + ay * ay + az * az;
      // sum_up_whole_potential.cpp:404-407: l = l_hartree; while (opr(l) < d^2 && l > 0) l--
      l[q] = c.l_const;
      if (op_top < d2.v[q] && l[q] > 0) {
        l[q]--;
        while (op[l[q]] < d2.v[q] && l[q] > 0)
          l[q]--;
      }
      far[q] = active[q] && !(d2.v[q] < rsq) && d2.v[q] < 1e8;
      pend[q] = __ballot(far[q]);
      any |= pend[q];
    }
    if (!any)
      continue;
    DV<N> mx, my, mz;
#pragma unroll
    for (int q = 0; q < N; q++) {
      mx.v[q] = rx[q] - c.mx;
      my.v[q] = ry[q] - c.my;
      mz.v[q] = rz[q] - c.mz;
    }
    const double *cb = coefs + (size_t)cstride * (c.mc_atom + (size_t)n_atoms * t);
    for (;;) {
      int L = -1;
#pragma unroll
      for (int q = 0; q < N; q++)
        if (L < 0 && pend[q])
          L = __builtin_amdgcn_readlane(l[q], __ffsll((long long)pend[q]) - 1);
      if (L < 0)
        break;
      bool mine[N], anymine = false;
#pragma unroll
      for (int q = 0; q < N; q++) {
        mine[q] = far[q] && l[q] == L;
        pend[q] &= ~__ballot(mine[q]);
        anymine |= mine[q];
      }
      if (anymine) {
        const DV<N> r = pair_switch_v<N>(std::make_integer_sequence<int, LMAXK + 1>{}, L, cb, mx, my, mz, d2);
#pragma unroll
        for (int q = 0; q < N; q++)
          if (mine[q])
            V[q] += r.v[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < N; q++)
    if (active[q])
      dv[(size_t)p[q] + (size_t)n_full * t] += V[q];
}

// ---------------------------------------------------------------------------------------------
// Adjoint: d/dM of sum_p gV(p, t) delta_v(p, t)
// ---------------------------------------------------------------------------------------------

template <int N, int LM>
__device__ __forceinline__ void adj_term(double *acc, int l, double g, const double *px, const double *py,
                                         const double *pz, const double *Fp, const double *__restrict__ cc) {
  constexpr FarTerm T = FAR_TERMS[N];
  constexpr int lm = T.l * T.l + T.l + T.m;
  if (T.l <= l)
    acc[lm] += g * cc[N] * (px[T.i] * py[T.j] * pz[T.k] * Fp[T.nn]);
}
template <int LM, int... Ns>
__device__ __forceinline__ void adj_terms(std::integer_sequence<int, Ns...>, double *acc, int l, double g,
                                          const double *px, const double *py, const double *pz, const double *Fp,
                                          const double *__restrict__ cc) {
  (adj_term<Ns, LM>(acc, l, g, px, py, pz, Fp, cc), ...);
}

template <int LM>
__global__ __launch_bounds__(256) void far_adjoint_kernel(
    const CenterRec *__restrict__ centers, const double *__restrict__ cp, const int *__restrict__ map, int n_valid,
    const double *__restrict__ opr, int lp1, const double *__restrict__ mrsq, int n_atoms,
    const double *__restrict__ cc, const double *__restrict__ gV, int n_full, double *__restrict__ partial) {
  constexpr int NLM = (LM + 1) * (LM + 1);
  __shared__ double red[4][NLM];
  const int ic = blockIdx.y, t = blockIdx.z, nbx = gridDim.x;
  const CenterRec c = centers[ic];
  const double *op = opr + (size_t)lp1 * (c.atom + (size_t)n_atoms * t);
  const double rsq = mrsq[c.atom + (size_t)n_atoms * t];
  double acc[NLM];
#pragma unroll
  for (int k = 0; k < NLM; k++)
    acc[k] = 0.0;
  for (int iv = blockIdx.x * 256 + threadIdx.x; iv < n_valid; iv += nbx * 256) {
    const int p = map[iv];
    const double rx = cp[3 * (size_t)p], ry = cp[3 * (size_t)p + 1], rz = cp[3 * (size_t)p + 2];
    const double ax = rx - c.dx, ay = ry - c.dy, az = rz - c.dz;
    const double d2 = ax * ax + ay * ay + az * az;
    int l = c.l_const;
    while (op[l] < d2 && l > 0)
      l--;
    if (d2 < rsq || !(d2 < 1e8))
      continue;
    const double g = gV[(size_t)p + (size_t)n_full * t];
    const double dist = sqrt(d2);
    const double dsq = dist * dist;
    double Fp[LM + 1], py[LM + 1], pz[LM + 1];
    const double px[2] = {1.0, rx - c.mx};
    const double my = ry - c.my, mz = rz - c.mz;
    Fp[0] = 1.0 / dist;
    py[0] = pz[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= LM; k++) {
      Fp[k] = Fp[k - 1] * (double)(1 - 2 * k) / dsq;
      py[k] = my * py[k - 1];
      pz[k] = mz * pz[k - 1];
    }
    adj_terms<LM>(std::make_integer_sequence<int, FAR_NCC[LM]>{}, acc, l, g, px, py, pz, Fp, cc);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NLM; k++) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      v += __shfl_xor(v, o);
    if (lane == 0)
      red[w][k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NLM; k += 256)
    partial[((size_t)(t * gridDim.y + ic) * nbx + blockIdx.x) * NLM + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// grad_M(lm, a, t) += sum over centres ic with mc_atom(ic) == a and blocks of the partial sums,
// masked as the forward masks multipole_c.
__global__ void far_adjoint_reduce_kernel(int n_atoms, int n_hc, int nbx, int nlm_k, int l_pot_max, int lama,
                                          const int *__restrict__ atom_ic_ptr, const int *__restrict__ atom_ic,
                                          const int *__restrict__ lmfd, const double *__restrict__ M,
                                          const double *__restrict__ partial, double *__restrict__ gM) {
  const int a = blockIdx.x, t = blockIdx.y;
  const int lpp = (l_pot_max + 1) * (l_pot_max + 1);
  const int lm_max = clampi(lmfd[a + n_atoms * t], 0, lama);
  for (int k = threadIdx.x; k < lpp; k += blockDim.x) {
    int l = 0;
    while ((l + 1) * (l + 1) <= k)
      l++;
    double s = 0.0;
    for (int q = atom_ic_ptr[a]; q < atom_ic_ptr[a + 1]; q++) {
      const int ic = atom_ic[q];
      for (int bx = 0; bx < nbx; bx++)
        s += partial[((size_t)(t * n_hc + ic) * nbx + bx) * nlm_k + k];
    }
    const size_t mi = k + (size_t)lpp * (a + (size_t)n_atoms * t);
    if (l <= lm_max && fabs(M[mi]) > 1e-10)
      gM[mi] += s;
  }
}

template <int N>
void launch_horner(FarPlan &P, const FarCall &c, const double *coef, double *dv, hipStream_t s) {
  dim3 grid((c.n_valid + 256 * N - 1) / (256 * N), c.n_tile);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, P.d_centers.ptr, P.n_hc, c.coord_points, c.map, c.n_valid, c.opr,
                       P.l_pot_max + 1, c.mrsq, P.n_atoms, coef, P.coef_stride, dv, c.n_full);
  };
  switch (P.l_pot_max) {
#define FRPA_CASE(LL)                                                                                          \
  case LL:                                                                                                     \
    go(far_horner_kernel<LL, N>);                                                                              \
    break;
    FRPA_CASE(0) FRPA_CASE(1) FRPA_CASE(2) FRPA_CASE(3) FRPA_CASE(4) FRPA_CASE(5) FRPA_CASE(6) FRPA_CASE(7)
    FRPA_CASE(8) FRPA_CASE(9) FRPA_CASE(10)
#undef FRPA_CASE
  default:
    FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "l_pot_max %d > %d", P.l_pot_max, FAR_LMAX);
  }
  check_launch("far_horner_kernel");
}

template <int VARIANT>
void launch_forward(FarPlan &P, const FarCall &c, const double *coefs, double *dv, hipStream_t s) {
  dim3 grid((c.n_valid + 255) / 256, c.n_tile);
  const int cstride = VARIANT == FRPA_FAR_HORNER ? P.coef_stride : P.mc_dim0;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, P.d_centers.ptr, P.n_hc, c.coord_points, c.map, c.n_valid, c.opr,
                       P.l_pot_max + 1, c.mrsq, P.n_atoms, coefs, cstride, dv, c.n_full);
  };
  // kernels are specialised for the largest l they may meet: l_pot_max
  switch (P.l_pot_max) {
#define FRPA_CASE(LL)                                                                                          \
  case LL:                                                                                                     \
    go(far_forward_kernel<VARIANT, LL>);                                                                       \
    break;
    FRPA_CASE(0) FRPA_CASE(1) FRPA_CASE(2) FRPA_CASE(3) FRPA_CASE(4) FRPA_CASE(5) FRPA_CASE(6) FRPA_CASE(7)
    FRPA_CASE(8) FRPA_CASE(9) FRPA_CASE(10)
#undef FRPA_CASE
  default:
    FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "l_pot_max %d > %d", P.l_pot_max, FAR_LMAX);
  }
  check_launch("far_forward_kernel");
}

} // namespace

size_t far_coef_doubles(const FarPlan &P, int variant, int T) {
  const size_t stride = variant == FRPA_FAR_TERMS ? (size_t)P.mc_dim0 : (size_t)P.coef_stride;
  return stride * (size_t)P.n_atoms * (size_t)T;
}

void far_forward(FarPlan &P, const FarCall &c, int variant, double *dv, hipStream_t s, double *coef_buf) {
  if (c.n_valid == 0 || c.n_tile == 0 || P.n_hc == 0)
    return;
  const int lama = P.lama;
  if (variant == FRPA_FAR_TERMS) {
    FRPA_REQUIRE(P.terms_ok, FRPA_ERR_UNSUPPORTED,
                 "index_cc differs from the compiled term table; use FRPA_FAR_HORNER");
    double *mc = coef_buf;
    if (!mc) {
      P.d_mc.grow(far_coef_doubles(P, variant, c.T));
      mc = P.d_mc.ptr;
    }
    hipLaunchKernelGGL(far_init_mc_kernel, dim3(P.n_atoms, c.n_tile), dim3(256), 0, s, P.n_atoms, P.mc_dim0, P.dim0,
                       P.l_pot_max, lama, P.d_index_cc.ptr, P.d_cc.ptr, P.d_ncc.ptr, c.lmfd, c.moments, mc);
    check_launch("far_init_mc_kernel");
    launch_forward<FRPA_FAR_TERMS>(P, c, mc, dv, s);
  } else if (variant == FRPA_FAR_HORNER) {
    double *coef = coef_buf;
    if (!coef) {
      P.d_coef.grow(far_coef_doubles(P, variant, c.T));
      coef = P.d_coef.ptr;
    }
    hipLaunchKernelGGL(far_init_horner_kernel, dim3(P.n_atoms, c.n_tile), dim3(256), 0, s, P.n_atoms,
                       P.coef_stride, P.dim0, P.l_pot_max, lama, P.d_index_cc.ptr, P.d_ncc.ptr, c.lmfd, c.moments,
                       P.d_hc_ptr.ptr, P.d_hc_term.ptr, P.d_hc_w.ptr, coef);
    check_launch("far_init_horner_kernel");
    // points per lane of the Horner kernel (FRPA_FAR_POINTS_PER_LANE = 1, 2 or 4; default 4: l = 8
    // over 6M points x 1024 centres runs 73 / 46 / 45 ms)
    const char *e = getenv("FRPA_FAR_POINTS_PER_LANE");
    const int ppl = e ? atoi(e) : 4;
    if (ppl == 1)
      launch_forward<FRPA_FAR_HORNER>(P, c, coef, dv, s);
    else if (ppl == 2)
      launch_horner<2>(P, c, coef, dv, s);
    else
      launch_horner<4>(P, c, coef, dv, s);
  } else {
    FRPA_REQUIRE(false, FRPA_ERR_INVALID, "unknown far-field variant %d", variant);
  }
}

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[SCRATCH_MAX_DEVICES];  // device address of `scratch` per device, resolved once

// Address of the fixed coefficient array on the current device (looked up on the first call on it,
// with that device current).
double *far_scratch() {
  int dev = -1;
  FRPA_HIP(hipGetDevice(&dev));
  FRPA_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, FRPA_ERR_UNSUPPORTED, "device %d: at most %d devices", dev,
               SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    FRPA_REQUIRE(e == hipSuccess && p, FRPA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                 hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
  }
  return g_scratch[dev];
}
} // namespace

void far_forward_noalloc(FarPlan &P, const FarCall &c, int variant, double *dv, hipStream_t s) {
  FRPA_REQUIRE(variant == FRPA_FAR_TERMS || variant == FRPA_FAR_HORNER, FRPA_ERR_INVALID,
               "unknown far-field variant %d", variant);
  const size_t need = far_coef_doubles(P, variant, c.T);
  FRPA_REQUIRE(need <= (size_t)FRPA_FAR_NOALLOC_SCRATCH_DOUBLES, FRPA_ERR_UNSUPPORTED,
               "far forward without allocation: the coefficient table needs %zu doubles (%d per atom x %d atoms x "
               "atom_tile_size %d), the fixed scratch holds %d",
               need, variant == FRPA_FAR_TERMS ? P.mc_dim0 : P.coef_stride, P.n_atoms, c.T,
               FRPA_FAR_NOALLOC_SCRATCH_DOUBLES);
  if (c.n_valid == 0 || c.n_tile == 0 || P.n_hc == 0)
    return;
  far_forward(P, c, variant, dv, s, far_scratch());
}

void far_adjoint(FarPlan &P, const FarCall &c, const double *gV, double *gM, hipStream_t s) {
  if (c.n_valid == 0 || c.n_tile == 0 || P.n_hc == 0)
    return;
  FRPA_REQUIRE(P.terms_ok, FRPA_ERR_UNSUPPORTED, "far adjoint needs the compiled term table");
  const int nbx = std::max(1, std::min(64, (c.n_valid + 4095) / 4096));
  int LM = P.l_pot_max <= 4 ? 4 : (P.l_pot_max <= 6 ? 6 : (P.l_pot_max <= 8 ? 8 : 10));
  FRPA_REQUIRE(P.l_pot_max <= 10, FRPA_ERR_UNSUPPORTED, "l_pot_max %d > 10", P.l_pot_max);
  const int nlm = (LM + 1) * (LM + 1);
  P.d_partial.grow((size_t)c.n_tile * P.n_hc * nbx * nlm);
  dim3 grid(nbx, P.n_hc, c.n_tile);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, P.d_centers.ptr, c.coord_points, c.map, c.n_valid, c.opr,
                       P.l_pot_max + 1, c.mrsq, P.n_atoms, P.d_cc.ptr, gV, c.n_full, P.d_partial.ptr);
  };
  if (LM == 4)
    go(far_adjoint_kernel<4>);
  else if (LM == 6)
    go(far_adjoint_kernel<6>);
  else if (LM == 8)
    go(far_adjoint_kernel<8>);
  else
    go(far_adjoint_kernel<10>);
  check_launch("far_adjoint_kernel");
  hipLaunchKernelGGL(far_adjoint_reduce_kernel, dim3(P.n_atoms, c.n_tile), dim3(128), 0, s, P.n_atoms, P.n_hc, nbx,
                     nlm, P.l_pot_max, P.lama, P.d_atom_ic_ptr.ptr, P.d_atom_ic.ptr, c.lmfd, c.moments,
                     P.d_partial.ptr, gM);
  check_launch("far_adjoint_reduce_kernel");
}

} // namespace frpa
