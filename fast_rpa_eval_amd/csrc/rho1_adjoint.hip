// rho1 adjoint (first-order H integration): placeholder until the kernel lands.
#include "rho1_common.hpp"

namespace frpa {

void rho1_adjoint(Rho1Plan &, int, int, const double *, const double *, double *, hipStream_t) {
  FRPA_REQUIRE(false, FRPA_ERR_UNSUPPORTED, "rho1 adjoint is not available in this build");
}

} // namespace frpa
