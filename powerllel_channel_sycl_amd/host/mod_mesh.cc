#include "mod_mesh.hh"

#include <cstdio>
#include <fstream>
#include <stdexcept>

#include "mod_dataIO.hh"
#include "mod_device.hh"
#include "mod_parameters.hh"

namespace mod_mesh {

void initMesh() {
  xc_global.allocate(nx);
  yc_global.allocate(ny);
  zc_global.allocate(nz);
  dzf_global.allocate(nz);
  dzc.allocate(nz);
  dzc_inv.allocate(nz);
  dzf.allocate(nz);
  dzf_inv.allocate(nz);
  dzflzi.allocate(nz);
  visc_dzf_inv.allocate(nz);

  dx = lx / nx;
  dy = ly / ny;
  dx_inv = 1.0 / dx;
  dy_inv = 1.0 / dy;
  xc_global(0) = -0.5 * dx;
  for (int i = 1; i <= nx + 1; ++i) xc_global(i) = xc_global(0) + i * dx;
  yc_global(0) = -0.5 * dy;
  for (int j = 1; j <= ny + 1; ++j) yc_global(j) = yc_global(0) + j * dy;

  if (read_mesh)
    inputMesh(nz, dzf_global);
  else
    initMeshByFunc(mesh_type, stretch_ratio, nz, lz, dzf_global);

  zc_global(0) = -0.5 * dzf_global(0);
  for (int k = 1; k <= nz + 1; ++k)
    zc_global(k) = zc_global(k - 1) + 0.5 * (dzf_global(k - 1) + dzf_global(k));
  for (int k = 0; k <= nz + 1; ++k) {
    dzf(k) = dzf_global(k);
    dzf_inv(k) = 1.0 / dzf(k);
    visc_dzf_inv(k) = 1.0 / dzf(k);
  }
  for (int k = 0; k <= nz; ++k) {
    dzc(k) = 0.5 * (dzf(k) + dzf(k + 1));
    dzc_inv(k) = 1.0 / dzc(k);
  }
  for (int k = 0; k <= nz + 1; ++k) dzflzi(k) = dzf(k) / lz;
  // smooth_wall_visc only changes visc_dzf_inv, which no kernel reads
  // (src/mod_mesh.cc:57,67-73).
  if (smooth_wall_visc) {
    visc_dzf_inv(1) = 0.5 * (dzc(0) + dzc(1));
    visc_dzf_inv(nz) = 0.5 * (dzc(nz - 1) + dzc(nz));
  }
}

void inputMesh(int nz_global, Array1DH1& dzf_g) {
  std::ifstream ifile("mesh.in");
  if (!ifile) throw std::runtime_error("PowerLLEL.ERROR.inputMesh: File mesh.in doesn't exist!");
  for (int k = 1; k <= nz_global; ++k) {
    if (!(ifile >> dzf_g(k)))
      throw std::runtime_error("PowerLLEL.ERROR.inputMesh: mesh.in holds fewer than nz values");
  }
  dzf_g(0) = dzf_g(1);
  dzf_g(nz_global + 1) = dzf_g(nz_global);
}

void initMeshByFunc(int mtype, fp ratio, int nz_global, fp lz_, Array1DH1& dzf_g) {
  mod_device::check(pll_mesh_by_func(mtype, stretch_func, ratio, nz_global, lz_, dzf_g.data()),
                    "pll_mesh_by_func");
}

void outputMesh() {
  char name[64];
  std::snprintf(name, sizeof(name), "mesh_%d-%d-%d.h5", nx, ny, nz);
  mod_dataIO::writeMesh(name);
  std::printf("PowerLLEL.NOTE.outputMesh: Finish writing file %s!\n", name);
  std::snprintf(name, sizeof(name), "mesh_%05d.out", nz);
  FILE* f = std::fopen(name, "w");
  if (!f) throw std::runtime_error(std::string("PowerLLEL.ERROR.outputMesh: cannot write ") + name);
  std::fprintf(f, "%4s%s%21s%s%22s%s\n", "", "k", "", "dzf", "", "zc");
  for (int k = 1; k <= nz; ++k)
    std::fprintf(f, "%05d %23.15E %23.15E\n", k, dzf_global(k), zc_global(k));
  std::fclose(f);
  std::printf("PowerLLEL.NOTE.outputMesh: Finish writing file %s!\n", name);
}

void freeMesh() {
  xc_global.deallocate();
  yc_global.deallocate();
  zc_global.deallocate();
  dzf_global.deallocate();
  dzc.deallocate();
  dzc_inv.deallocate();
  dzf.deallocate();
  dzf_inv.deallocate();
  dzflzi.deallocate();
  visc_dzf_inv.deallocate();
}

}  // namespace mod_mesh
