// decomp_2d.hh -- 2D pencil decomposition and transposes (src/decomp_2d.hh:1-63).
//
// The p_row x p_col ranks are virtual: every rank's pencil lives on this GPU,
// one after another in rank order (rank = coords[0]*p_col + coords[1]).  The
// transposes take device pointers to those whole buffers; src_size and
// dst_size, when given, must equal the pencil sizes of rank decomp_2d_nrank
// as in the reference, and are checked rather than used.
#pragma once

#include <array>
#include <vector>

#include "pll.h"

typedef double DECOMP_2D_REAL;

inline int decomp_2d_nx_global, decomp_2d_ny_global, decomp_2d_nz_global;
inline int decomp_2d_periodic_x, decomp_2d_periodic_y, decomp_2d_periodic_z;
inline int decomp_2d_nrank = 0, decomp_2d_nproc = 1;
inline int dims[2] = {1, 1};

struct decomp_2d_info {
  pll_decomp* handle = nullptr;
  bool even = false;
  // starting/ending index (1-based) and size held by rank decomp_2d_nrank
  std::array<int, 3> xst{}, xen{}, xsz{};
  std::array<int, 3> yst{}, yen{}, ysz{};
  std::array<int, 3> zst{}, zen{}, zsz{};
  // how each dimension is distributed along pencils
  std::vector<int> x1st, x1en, x1dist, y1st, y1en, y1dist, y2st, y2en, y2dist, z2st, z2en, z2dist;
};
inline decomp_2d_info decomp_main;

void decomp_2d_init(int nx, int ny, int nz, int p_row, int p_col, std::array<bool, 3> periodic_bc);
void decomp_2d_finalize();
[[noreturn]] void decomp_2d_abort(int errorcode, const char* msg);
void decomp_info_init(decomp_2d_info* decomp, int nx, int ny, int nz);
void decomp_info_finalize(decomp_2d_info* decomp);
void transpose_x_to_y_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp);
void transpose_y_to_x_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp);
void transpose_y_to_z_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp);
void transpose_z_to_y_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp);
