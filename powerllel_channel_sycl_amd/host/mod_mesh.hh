// mod_mesh.hh -- grid spacings and wall-normal metrics (src/mod_mesh.hh,
// src/mod_mesh.cc:13-147).  Host copies; the device context derives its own
// from dzf_global at creation with the same formulas.
#pragma once

#include "mod_type.hh"

namespace mod_mesh {
inline fp dx, dy;
inline fp dx_inv, dy_inv;
inline Array1DH1 xc_global;
inline Array1DH1 yc_global;
inline Array1DH1 zc_global;
inline Array1DH1 dzf_global;
inline Array1DH1 dzc, dzc_inv;
inline Array1DH1 dzf, dzf_inv;
inline Array1DH1 dzflzi;
inline Array1DH1 visc_dzf_inv;

void initMesh();
// Reads nz_global spacings from mesh.in (src/mod_mesh.cc:78-93).
void inputMesh(int nz_global, Array1DH1& dzf_global);
void initMeshByFunc(int mesh_type, fp stretch_ratio, int nz_global, fp lz, Array1DH1& dzf_global);
// Writes mesh_<nx>-<ny>-<nz>.h5 (datasets xc, yc, zc) and mesh_<nz>.out (k,
// dzf, zc), src/mod_mesh.cc:127-145.
void outputMesh();
void freeMesh();
}  // namespace mod_mesh
