#include "mod_initFlow.hh"

#include <cstdint>
#include <stdexcept>

#include "mod_device.hh"
#include "mod_parameters.hh"

namespace mod_initFlow {

void initFlow(Array3DH3& u, Array3DH3& v, Array3DH3& w) {
  if (u.slot() != PLL_U || v.slot() != PLL_V || w.slot() != PLL_W)
    throw std::invalid_argument("initFlow: expects the context's u, v, w");
  int profile;
  fp vel0 = u0;
  if (initial_field == "uni") {
    profile = PLL_PROFILE_UNI;
  } else if (initial_field == "poi") {
    profile = PLL_PROFILE_POI;
  } else if (initial_field == "log") {
    profile = PLL_PROFILE_LOG;
  } else if (initial_field == "mpi") {
    profile = PLL_PROFILE_UNI;
    vel0 = 0.0;  // the rank number of the single rank
  } else {
    throw std::invalid_argument("initFlow: unknown initial_field <" + initial_field + ">");
  }
  const uint32_t seeds[3] = {123u, 456u, 789u};
  mod_device::check(pll_init_flow(mod_device::context(), profile, vel0, init_with_noise ? 1 : 0,
                                  noise_intensity, seeds),
                    "initFlow");
  mod_device::sync();
}

}  // namespace mod_initFlow
