// mod_dataIO.hh -- checkpoint and field output (src/mod_dataIO.hh,
// src/mod_dataIO.cc:68-330) in the reference's HDF5 layout: one file per
// field, dataset <tag> of shape (nz, ny, nx) in FP64 holding the interior,
// and a one-element int dataset "nt" (statistics files: "nts", "nte", "nspl").
#pragma once

#include "mod_type.hh"

namespace mod_dataIO {
// Reads <fn_prefix_input_inst>_{u,v,w}.h5; returns their common nt.
int inputData(Array3DH3& u, Array3DH3& v, Array3DH3& w);
// Instantaneous fields every nt_out_inst, checkpoints every nt_out_save,
// statistics after nt_init_stat (src/mod_dataIO.cc:130-226).
void outputData(int nt, fp u_crf, Array3DH3& u, const Array3DH3& v, const Array3DH3& w,
                const Array3DH1& p);
void inputStatData(int nt_in_inst);
void outputStatData(int nt);

// The mesh file of outputMesh (src/mod_mesh.cc:127-136): datasets xc, yc, zc
// of the interior cell-centre coordinates (nx, ny, nz values).
void writeMesh(const char* filename);

// One field file (exposed for tests): write the interior of var (host mirror
// must be current) / read into the mirror's interior.
void writeField(const char* filename, const char* tag, int nt, const Array3D& var, fp shift = 0.0);
int readField(const char* filename, const char* tag, Array3D& var);
}  // namespace mod_dataIO
