// mod_initFlow.hh -- initial velocity field (src/mod_initFlow.hh,
// src/mod_initFlow.cc:16-102).
#pragma once

#include "mod_type.hh"

namespace mod_initFlow {
// initial_field "uni" | "poi" | "log" ("mpi" is "uni" with the rank, 0, as
// the velocity); noise seeds 123/456/789 for u/v/w (src/mod_initFlow.cc:86-88)
// drawn from the library's documented stream (pll_init_flow).
void initFlow(Array3DH3& u, Array3DH3& v, Array3DH3& w);
}  // namespace mod_initFlow
