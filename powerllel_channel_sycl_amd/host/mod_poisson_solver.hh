// mod_poisson_solver.hh -- FFT(x, y) + tridiagonal(z) pressure solver
// (src/mod_poisson_solver.hh:5-10).  init_poisson_solver creates the device
// context: it owns the fields, the FFT plans and the factored tridiagonal
// systems, so the run parameters (readInputParam) and the mesh (initMesh)
// must be set first, as in src/main.cc:36-57.
#pragma once

#include <array>

#include "mod_type.hh"

void init_poisson_solver(int nx_global, int ny_global, int nz_global, double dx, double dy,
                         double* dzf_global, const char bctype_x[], const char bctype_y[],
                         const char bctype_z[], std::array<std::array<int, 6>, 3>& neighbor_xyz);
void execute_poisson_solver(Array3DH1& p);
void free_poisson_solver();
