#include "decomp_2d.hh"

#include <string>

#include "mod_device.hh"

void decomp_2d_abort(int errorcode, const char* msg) {
  throw mod_device::Error(PLL_ERR_ARG, "DECOMP_2D ERROR - errorcode: " + std::to_string(errorcode) +
                                           ", message: " + msg);
}

void decomp_2d_init(int nx, int ny, int nz, int p_row, int p_col, std::array<bool, 3> periodic_bc) {
  decomp_2d_nx_global = nx;
  decomp_2d_ny_global = ny;
  decomp_2d_nz_global = nz;
  decomp_2d_periodic_x = periodic_bc[0] ? 1 : 0;
  decomp_2d_periodic_y = periodic_bc[1] ? 1 : 0;
  decomp_2d_periodic_z = periodic_bc[2] ? 1 : 0;
  if (p_row < 1 || p_col < 1) decomp_2d_abort(1, "Invalid 2D processor grid - nproc /= p_row*p_col");
  dims[0] = p_row;
  dims[1] = p_col;
  decomp_2d_nproc = p_row * p_col;
  decomp_2d_nrank = 0;
  decomp_info_init(&decomp_main, nx, ny, nz);
}

void decomp_2d_finalize() { decomp_info_finalize(&decomp_main); }

void decomp_info_init(decomp_2d_info* d, int nx, int ny, int nz) {
  if (nx < dims[0] || ny < dims[0] || ny < dims[1] || nz < dims[1])
    decomp_2d_abort(6, "Invalid 2D processor grid. Make sure that min(nx,ny) >= p_row and "
                       "min(ny,nz) >= p_col");
  decomp_info_finalize(d);
  mod_device::check(pll_decomp_create(nx, ny, nz, dims[0], dims[1], &d->handle), "decomp_info_init");
  d->even = nx % dims[0] == 0 && ny % dims[0] == 0 && ny % dims[1] == 0 && nz % dims[1] == 0;
  int st[3], sz[3];
  long off;
  std::array<int, 3>* S[3] = {&d->xst, &d->yst, &d->zst};
  std::array<int, 3>* E[3] = {&d->xen, &d->yen, &d->zen};
  std::array<int, 3>* Z[3] = {&d->xsz, &d->ysz, &d->zsz};
  for (int p = 0; p < 3; ++p) {
    mod_device::check(pll_decomp_pencil(d->handle, p, decomp_2d_nrank, st, sz, &off), "decomp_info_init");
    for (int a = 0; a < 3; ++a) {
      (*S[p])[a] = st[a];
      (*Z[p])[a] = sz[a];
      (*E[p])[a] = st[a] + sz[a] - 1;
    }
  }
  // the distributions, read back from the pencils of the ranks on one row / column
  auto fill = [&](int pencil, int axis, bool by_row, std::vector<int>& vst, std::vector<int>& ven,
                  std::vector<int>& vdist) {
    const int n = by_row ? dims[0] : dims[1];
    vst.resize(n);
    ven.resize(n);
    vdist.resize(n);
    for (int m = 0; m < n; ++m) {
      const int rank = by_row ? m * dims[1] : m;
      mod_device::check(pll_decomp_pencil(d->handle, pencil, rank, st, sz, &off), "decomp_info_init");
      vst[m] = st[axis];
      vdist[m] = sz[axis];
      ven[m] = st[axis] + sz[axis] - 1;
    }
  };
  fill(1, 0, true, d->x1st, d->x1en, d->x1dist);   // y-pencil x over p_row
  fill(0, 1, true, d->y1st, d->y1en, d->y1dist);   // x-pencil y over p_row
  fill(2, 1, false, d->y2st, d->y2en, d->y2dist);  // z-pencil y over p_col
  fill(0, 2, false, d->z2st, d->z2en, d->z2dist);  // x-pencil z over p_col
}

void decomp_info_finalize(decomp_2d_info* d) {
  if (d->handle) pll_decomp_destroy(d->handle);
  d->handle = nullptr;
}

namespace {
void check_size(const int* given, const std::array<int, 3>& want, const char* what) {
  if (given && (given[0] != want[0] || given[1] != want[1] || given[2] != want[2]))
    decomp_2d_abort(2, what);
}
decomp_2d_info* info(decomp_2d_info* d) {
  if (!d || !d->handle) decomp_2d_abort(3, "decomposition not initialised");
  return d;
}
}  // namespace

void transpose_x_to_y_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp) {
  decomp_2d_info* d = info(decomp);
  check_size(src_size, d->xsz, "transpose_x_to_y_real: src_size is not the x-pencil size");
  check_size(dst_size, d->ysz, "transpose_x_to_y_real: dst_size is not the y-pencil size");
  mod_device::check(pll_transpose_x_to_y_real(d->handle, src, dst, sizeof(DECOMP_2D_REAL)),
                    "transpose_x_to_y_real");
}

void transpose_y_to_x_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp) {
  decomp_2d_info* d = info(decomp);
  check_size(src_size, d->ysz, "transpose_y_to_x_real: src_size is not the y-pencil size");
  check_size(dst_size, d->xsz, "transpose_y_to_x_real: dst_size is not the x-pencil size");
  mod_device::check(pll_transpose_y_to_x_real(d->handle, src, dst, sizeof(DECOMP_2D_REAL)),
                    "transpose_y_to_x_real");
}

void transpose_y_to_z_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp) {
  decomp_2d_info* d = info(decomp);
  check_size(src_size, d->ysz, "transpose_y_to_z_real: src_size is not the y-pencil size");
  check_size(dst_size, d->zsz, "transpose_y_to_z_real: dst_size is not the z-pencil size");
  mod_device::check(pll_transpose_y_to_z_real(d->handle, src, dst, sizeof(DECOMP_2D_REAL)),
                    "transpose_y_to_z_real");
}

void transpose_z_to_y_real(DECOMP_2D_REAL* src, DECOMP_2D_REAL* dst, int* src_size, int* dst_size,
                           decomp_2d_info* decomp) {
  decomp_2d_info* d = info(decomp);
  check_size(src_size, d->zsz, "transpose_z_to_y_real: src_size is not the z-pencil size");
  check_size(dst_size, d->ysz, "transpose_z_to_y_real: dst_size is not the y-pencil size");
  mod_device::check(pll_transpose_z_to_y_real(d->handle, src, dst, sizeof(DECOMP_2D_REAL)),
                    "transpose_z_to_y_real");
}
