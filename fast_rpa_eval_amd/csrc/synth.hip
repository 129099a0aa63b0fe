// Synthetic wave_batches_compress generated in place on the device (benchmark inputs of tens of GB
// are not built on the host).  phi_mu(r) = exp(-alpha_mu |r - R_a|^2) Y_{l m}((r - R_a)^), l <= 2,
// Y in the SHEval convention (fast_rap_src_dfpt_gpu_eval/sheval.hpp:9-60):
//   m = 0: K_l0 P_l(cos t);  m > 0: (-1)^m sqrt2 K_lm P_l^m cos(m phi);  m < 0: sqrt2 K_l|m| P_l^|m| sin(|m| phi),
//   K_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), P_l^m without the Condon-Shortley phase.
#include <cmath>

#include "rho1_common.hpp"

namespace frpa {
namespace {

__device__ double ylm_l2(int l, int m, double x, double y, double z) {
  const double pi = 3.14159265358979323846;
  if (l == 0)
    return sqrt(1.0 / (4.0 * pi));
  if (l == 1) {
    const double n = sqrt(3.0 / (4.0 * pi));
    return m < 0 ? n * y : (m == 0 ? n * z : -n * x);
  }
  // l == 2
  const double k20 = sqrt(5.0 / (4.0 * pi));
  const double k21 = sqrt(5.0 / (4.0 * pi) / 6.0);
  const double k22 = sqrt(5.0 / (4.0 * pi) / 24.0);
  const double s2 = sqrt(2.0);
  switch (m) {
  case -2: return s2 * k22 * 3.0 * 2.0 * x * y;
  case -1: return s2 * k21 * 3.0 * z * y;
  case 0: return k20 * 0.5 * (3.0 * z * z - 1.0);
  case 1: return -s2 * k21 * 3.0 * z * x;
  default: return s2 * k22 * 3.0 * (x * x - y * y);
  }
}

__global__ __launch_bounds__(256) void synth_wave_kernel(const BatchDesc *__restrict__ descs,
                                                         const int *__restrict__ ibasis0,
                                                         const int *__restrict__ map, const double *__restrict__ cp,
                                                         const double *__restrict__ axyz,
                                                         const int *__restrict__ batom, const int *__restrict__ bl,
                                                         const int *__restrict__ bm,
                                                         const double *__restrict__ balpha, double *__restrict__ wave) {
  const BatchDesc d = descs[blockIdx.x];
  const int ncp = (d.nc + 3) & ~3;
  const int *ib = ibasis0 + d.ib_off;
  double *out = wave + d.wave_off;
  for (int e = threadIdx.x; e < ncp * d.np; e += blockDim.x) {
    const int i = e % ncp, p = e / ncp;
    if (i >= d.nc) {
      out[e] = __builtin_nan("");
      continue;
    }
    const int mu = ib[i];
    const int a = batom[mu];
    const int q = map[d.pstart + p];
    const double dx = cp[3 * (size_t)q] - axyz[3 * a], dy = cp[3 * (size_t)q + 1] - axyz[3 * a + 1],
                 dz = cp[3 * (size_t)q + 2] - axyz[3 * a + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double inv = r > 0.0 ? 1.0 / r : 0.0;
    out[e] = exp(-balpha[mu] * r2) * ylm_l2(bl[mu], bm[mu], dx * inv, dy * inv, dz * inv);
  }
}

} // namespace

void synth_wave(const Rho1Plan &P, int n_atoms, const double *atom_xyz, const int *basis_atom, const int *basis_l,
                const int *basis_m, const double *basis_alpha, int n_basis, const double *coord_points, double *wave,
                hipStream_t s) {
  for (int mu = 0; mu < n_basis; mu++)
    FRPA_REQUIRE(basis_atom[mu] >= 0 && basis_atom[mu] < n_atoms && basis_l[mu] >= 0 && basis_l[mu] <= 2 &&
                     basis_m[mu] >= -basis_l[mu] && basis_m[mu] <= basis_l[mu],
                 FRPA_ERR_INVALID, "synthetic basis function %d: bad atom/l/m", mu);
  DevBuf<double> axyz, balpha;
  DevBuf<int> batom, bl, bm;
  axyz.upload(atom_xyz, (size_t)3 * n_atoms);
  balpha.upload(basis_alpha, n_basis);
  batom.upload(basis_atom, n_basis);
  bl.upload(basis_l, n_basis);
  bm.upload(basis_m, n_basis);
  if (P.n_valid > 0) {
    hipLaunchKernelGGL(synth_wave_kernel, dim3(P.n_valid), dim3(256), 0, s, P.d_desc.ptr, P.d_ibasis0.ptr,
                       P.d_map.ptr, coord_points, axyz.ptr, batom.ptr, bl.ptr, bm.ptr, balpha.ptr, wave);
    check_launch("synth_wave_kernel");
  }
  FRPA_HIP(hipStreamSynchronize(s));  // the small parameter buffers above are freed on return
}

} // namespace frpa
