// INT8-split (Ozaki) rho1: placeholder until the split kernels land.
#include "rho1_common.hpp"

namespace frpa {

struct OzakiState {};

void ozaki_release(OzakiState *st) { delete st; }

long long ozaki_prepare_wave(Rho1Plan &, const double *, hipStream_t) {
  FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "FRPA_PREC_OZAKI is not available in this build");
  return 0;
}

void rho1_forward_ozaki(Rho1Plan &, const Rho1Call &, hipStream_t) {
  FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "FRPA_PREC_OZAKI is not available in this build");
}

} // namespace frpa
