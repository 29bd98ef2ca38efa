#include "mod_variables.hh"

#include <stdexcept>

#include "mod_parameters.hh"

namespace mod_variables {

void allocVariables(std::array<int, 3> sz) {
  if (sz[0] != nx || sz[1] != ny || sz[2] != nz)
    throw std::runtime_error("allocVariables: one GPU holds the whole (nx, ny, nz) domain");
  u.allocate(PLL_U, sz);
  v.allocate(PLL_V, sz);
  w.allocate(PLL_W, sz);
  u1.allocate(PLL_U1, sz);
  v1.allocate(PLL_V1, sz);
  w1.allocate(PLL_W1, sz);
  p.allocate(PLL_P, sz);
}

void freeVariables() {
  for (Array3D* f : {&u, &v, &w, &u1, &v1, &w1, &p}) f->deallocate();
}

}  // namespace mod_variables
