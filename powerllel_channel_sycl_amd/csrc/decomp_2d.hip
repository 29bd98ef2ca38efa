// decomp_2d.hip -- the 2D pencil decomposition and its transposes
// (src/decomp_2d.cc:41-589) for p_row x p_col ranks held on one GPU.
//
// Bookkeeping follows decomp_2d_init / decomp_info_init / distribute /
// partition (src/decomp_2d.cc:41-207,257-335): rank r has coordinates
// (r / p_col, r % p_col) and owns one x-, y- and z-pencil, each stored with x
// fastest, then y, then z.  The device buffers hold every rank's pencil one
// after another in rank order.
//
// A transpose moves each value to the pencil that holds the same global
// (x, y, z) -- what the reference does with mem_split, MPI_Alltoallv and
// mem_merge.  Because every pencil keeps x fastest, each destination row is a
// contiguous run of a source row (x -> y, y -> z, z -> y) or at most p_row
// runs (y -> x), so one coalesced gather kernel writes every destination value
// exactly once; it moves 2 values per point, the HBM roofline of a permutation.
#include <algorithm>
#include <vector>

#include "pll_internal.hpp"

struct pll_decomp {
  int nx, ny, nz, p_row, p_col;
  std::vector<int> x1st, x1sz, y1st, y1sz, y2st, y2sz, z2st, z2sz;
  // per rank and pencil: start (1-based), size, element offset
  std::vector<int> st, sz;  // [(r*3 + pencil)*3 + axis]
  std::vector<long> off;    // [r*3 + pencil]
  std::vector<long> rows;   // [pencil*(P+1) + r]: first x-row of rank r in the pencil
  long total[3];
  std::vector<int> meta;    // for each rank and pencil st[3], sz[3]
  std::vector<int> dists;   // y1st, x1st, z2st, y2st one after another
  int dist_at[4];           // where each starts in dists, by transpose kind
  int* d_meta = nullptr;    // device copies, made by the first transpose
  long* d_off = nullptr;
  long* d_rows = nullptr;
  int* d_dists = nullptr;
};

namespace pll {
namespace {

// Block split of 1..num over nprocs ranks, the same split as src/decomp_2d.cc:257-279:
// every rank holds floor(num/nprocs) indices and the last num % nprocs ranks one more.
// Closed form per rank r, with `first_long` = the first rank that holds the extra index:
//   count(r) = base + [r >= first_long],  start(r) = 1 + r*base + max(0, r - first_long).
void distribute(int num, int nprocs, std::vector<int>& st, std::vector<int>& sz) {
  const int base = num / nprocs;
  const int first_long = nprocs - num % nprocs;
  st.resize(nprocs);
  sz.resize(nprocs);
  for (int r = 0; r < nprocs; ++r) {
    const int extra_before = r > first_long ? r - first_long : 0;
    sz[r] = base + (r >= first_long ? 1 : 0);
    st[r] = 1 + r * base + extra_before;
  }
}

// The last index m < n with t[m] <= key of a non-decreasing table (t[0] <= key):
// binary search, log2(n) dependent loads instead of up to n.
template <typename V, typename K>
__device__ __forceinline__ int last_le(const V* t, int n, K key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t[mid] <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Owner index of global coordinate g (1-based) in a distribution (the last
// rank whose start is <= g, so ranks of size 0 are passed over).
__device__ __forceinline__ int owner(const int* st, int n, int g) { return last_le(st, n, g); }

struct XposeArgs {
  int p_row, p_col;
  int col;          // 1: comm_col (ranks differ in coords[0]); 0: comm_row
  int ps, pd;       // source and destination pencil
  int split_axis;   // axis whose distribution picks the source rank
  int dist_len;     // length of that distribution
  long rows;        // destination row count (rows of x)
};

// One wave per destination row (fixed y, z of one destination rank), lanes
// along x: the rank and row bookkeeping is done once per row, and both the
// loads and the stores of a wave are contiguous runs.  rows_off[r] is the
// first row of rank r in the destination; meta: [(r*3+pencil)*6 + {st, sz}].
template <typename E>
__global__ __launch_bounds__(256) void k_transpose(XposeArgs a, const int* __restrict__ meta,
                                                   const long* __restrict__ off,
                                                   const long* __restrict__ rows_off,
                                                   const int* __restrict__ dist_st,
                                                   const E* __restrict__ src, E* __restrict__ dst,
                                                   int stage) {
  const int P = a.p_row * a.p_col;
  const int lane = threadIdx.x & 63;
  // The bookkeeping tables go to LDS once per block (when they fit), so a
  // row's rank search and metadata reads are LDS reads, and the row's only
  // global latency is its data.
  extern __shared__ long xs_lds[];
  if (stage) {
    long* s_rows = xs_lds;                                        // P + 1
    long* s_off = s_rows + (P + 1);                               // 3P
    int* s_meta = reinterpret_cast<int*>(s_off + 3 * P);          // 18P
    int* s_dist = s_meta + 18 * P;                                // dist_len
    for (int t = threadIdx.x; t < P + 1; t += blockDim.x) s_rows[t] = rows_off[t];
    for (int t = threadIdx.x; t < 3 * P; t += blockDim.x) s_off[t] = off[t];
    for (int t = threadIdx.x; t < 18 * P; t += blockDim.x) s_meta[t] = meta[t];
    for (int t = threadIdx.x; t < a.dist_len; t += blockDim.x) s_dist[t] = dist_st[t];
    __syncthreads();
    rows_off = s_rows;
    off = s_off;
    meta = s_meta;
    dist_st = s_dist;
  }
  const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
  for (long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); row < a.rows;
       row += nwaves) {
    const int r = last_le(rows_off, P, row);  // destination rank of the row
    const int* dm = meta + (r * 3 + a.pd) * 6;
    const long lr = row - rows_off[r];
    const int lj = (int)(lr % dm[4]);
    const int lk = (int)(lr / dm[4]);
    const int gj = dm[1] + lj, gk = dm[2] + lk;
    E* drow = dst + off[r * 3 + a.pd] + lr * dm[3];
    const int c0 = r / a.p_col, c1 = r % a.p_col;
    // each lane moves U elements per pass: all loads are issued before the stores
    constexpr int U = 4;
    const int n = dm[3];
    if (a.split_axis == 0) {
      // y -> x: the x row gathers from p_row source pencils, one run each
      for (int l0 = 0; l0 < n; l0 += 64 * U) {
        E t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int li = l0 + u * 64 + lane;
          if (li < n) {
            const int gi = dm[0] + li;
            const int m = owner(dist_st, a.dist_len, gi);
            const int rs = m * a.p_col + c1;
            const int* sm = meta + (rs * 3 + a.ps) * 6;
            t[u] = src[off[rs * 3 + a.ps] + (gi - sm[0]) +
                       (long)sm[3] * ((gj - sm[1]) + (long)sm[4] * (gk - sm[2]))];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int li = l0 + u * 64 + lane;
          if (li < n) drow[li] = t[u];
        }
      }
    } else {
      const int m = owner(dist_st, a.dist_len, a.split_axis == 1 ? gj : gk);
      const int rs = a.col ? (m * a.p_col + c1) : (c0 * a.p_col + m);
      const int* sm = meta + (rs * 3 + a.ps) * 6;
      const E* srow = src + off[rs * 3 + a.ps] + (dm[0] - sm[0]) +
                      (long)sm[3] * ((gj - sm[1]) + (long)sm[4] * (gk - sm[2]));
      int l0 = 0;
      for (; l0 + 64 * U <= n; l0 += 64 * U) {  // whole passes: no guards between the stores
        E t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = srow[l0 + u * 64 + lane];
#pragma unroll
        for (int u = 0; u < U; ++u) drow[l0 + u * 64 + lane] = t[u];
      }
      if (l0 < n) {  // the rest of the row: guarded, loads still issued together
        E t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int li = l0 + u * 64 + lane;
          if (li < n) t[u] = srow[li];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int li = l0 + u * 64 + lane;
          if (li < n) drow[li] = t[u];
        }
      }
    }
  }
}

}  // namespace
}  // namespace pll

using namespace pll;

extern "C" {

int pll_decomp_create(int nx, int ny, int nz, int p_row, int p_col, pll_decomp** out) {
  if (!out) return fail(PLL_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || p_row < 1 || p_col < 1)
    return fail(PLL_ERR_ARG, "grid sizes and the processor grid must be positive");
  // src/decomp_2d.cc:117-124
  if (nx < p_row || ny < p_row || ny < p_col || nz < p_col)
    return fail(PLL_ERR_ARG,
                "Invalid 2D processor grid. Make sure that min(nx,ny) >= p_row and min(ny,nz) >= p_col");
  pll_decomp* d = new pll_decomp();
  d->nx = nx;
  d->ny = ny;
  d->nz = nz;
  d->p_row = p_row;
  d->p_col = p_col;
  distribute(nx, p_row, d->x1st, d->x1sz);
  distribute(ny, p_row, d->y1st, d->y1sz);
  distribute(ny, p_col, d->y2st, d->y2sz);
  distribute(nz, p_col, d->z2st, d->z2sz);
  const int P = p_row * p_col;
  d->st.assign(P * 9, 0);
  d->sz.assign(P * 9, 0);
  d->off.assign(P * 3, 0);
  d->total[0] = d->total[1] = d->total[2] = 0;
  std::vector<int> meta(P * 18);
  for (int r = 0; r < P; ++r) {
    const int c0 = r / p_col, c1 = r % p_col;
    // partition (src/decomp_2d.cc:309-335): pdim {0,1,2}, {1,0,2}, {1,2,0}
    const int st[3][3] = {{1, d->y1st[c0], d->z2st[c1]},
                          {d->x1st[c0], 1, d->z2st[c1]},
                          {d->x1st[c0], d->y2st[c1], 1}};
    const int sz[3][3] = {{nx, d->y1sz[c0], d->z2sz[c1]},
                          {d->x1sz[c0], ny, d->z2sz[c1]},
                          {d->x1sz[c0], d->y2sz[c1], nz}};
    for (int p = 0; p < 3; ++p) {
      for (int a = 0; a < 3; ++a) {
        d->st[(r * 3 + p) * 3 + a] = st[p][a];
        d->sz[(r * 3 + p) * 3 + a] = sz[p][a];
        meta[(r * 3 + p) * 6 + a] = st[p][a];
        meta[(r * 3 + p) * 6 + 3 + a] = sz[p][a];
      }
      d->off[r * 3 + p] = d->total[p];
      d->total[p] += (long)sz[p][0] * sz[p][1] * sz[p][2];
    }
  }
  d->meta = meta;
  d->rows.assign(3 * (P + 1), 0);
  for (int p = 0; p < 3; ++p)
    for (int r = 0; r < P; ++r)
      d->rows[p * (P + 1) + r + 1] =
          d->rows[p * (P + 1) + r] + (long)d->sz[(r * 3 + p) * 3 + 1] * d->sz[(r * 3 + p) * 3 + 2];
  const std::vector<int>* by_kind[4] = {&d->y1st, &d->x1st, &d->z2st, &d->y2st};
  for (int k = 0; k < 4; ++k) {
    d->dist_at[k] = (int)d->dists.size();
    d->dists.insert(d->dists.end(), by_kind[k]->begin(), by_kind[k]->end());
  }
  *out = d;
  return PLL_OK;
}

int pll_decomp_destroy(pll_decomp* d) {
  if (!d) return PLL_OK;
  if (d->d_meta) (void)hipFree(d->d_meta);
  if (d->d_off) (void)hipFree(d->d_off);
  if (d->d_dists) (void)hipFree(d->d_dists);
  if (d->d_rows) (void)hipFree(d->d_rows);
  delete d;
  return PLL_OK;
}

int pll_decomp_pencil(const pll_decomp* d, int pencil, int rank, int st[3], int sz[3],
                      long* offset) {
  if (!d) return fail(PLL_ERR_ARG, "null decomposition");
  if (pencil < 0 || pencil > 2 || rank < 0 || rank >= d->p_row * d->p_col)
    return fail(PLL_ERR_ARG, "pencil or rank out of range");
  for (int a = 0; a < 3; ++a) {
    if (st) st[a] = d->st[(rank * 3 + pencil) * 3 + a];
    if (sz) sz[a] = d->sz[(rank * 3 + pencil) * 3 + a];
  }
  if (offset) *offset = d->off[rank * 3 + pencil];
  return PLL_OK;
}

int pll_decomp_total(const pll_decomp* d, int pencil, long* total) {
  if (!d || !total || pencil < 0 || pencil > 2) return fail(PLL_ERR_ARG, "bad arguments");
  *total = d->total[pencil];
  return PLL_OK;
}

static int transpose(pll_decomp* d, int kind, const void* src, void* dst, int elem_bytes) {
  if (!d || !src || !dst) return fail(PLL_ERR_ARG, "null argument");
  if (elem_bytes != 8 && elem_bytes != 4) return fail(PLL_ERR_ARG, "elem_bytes must be 4 or 8");
  XposeArgs a;
  a.p_row = d->p_row;
  a.p_col = d->p_col;
  switch (kind) {
    case 0:  // x -> y within comm_col: source rank by the y1 owner of global y
      a.ps = 0; a.pd = 1; a.col = 1; a.split_axis = 1; a.dist_len = d->p_row; break;
    case 1:  // y -> x within comm_col: source rank by the x1 owner of global x
      a.ps = 1; a.pd = 0; a.col = 1; a.split_axis = 0; a.dist_len = d->p_row; break;
    case 2:  // y -> z within comm_row: source rank by the z2 owner of global z
      a.ps = 1; a.pd = 2; a.col = 0; a.split_axis = 2; a.dist_len = d->p_col; break;
    default:  // z -> y within comm_row: source rank by the y2 owner of global y
      a.ps = 2; a.pd = 1; a.col = 0; a.split_axis = 1; a.dist_len = d->p_col; break;
  }
  const int P = d->p_row * d->p_col;
  a.rows = d->rows[a.pd * (P + 1) + P];
  if (!d->d_meta) {
    PLL_HIP_TRY(hipMalloc(&d->d_meta, d->meta.size() * sizeof(int)));
    PLL_HIP_TRY(hipMalloc(&d->d_off, d->off.size() * sizeof(long)));
    PLL_HIP_TRY(hipMalloc(&d->d_dists, d->dists.size() * sizeof(int)));
    PLL_HIP_TRY(hipMalloc(&d->d_rows, d->rows.size() * sizeof(long)));
    PLL_HIP_TRY(hipMemcpy(d->d_rows, d->rows.data(), d->rows.size() * sizeof(long),
                          hipMemcpyHostToDevice));
    PLL_HIP_TRY(hipMemcpy(d->d_meta, d->meta.data(), d->meta.size() * sizeof(int),
                          hipMemcpyHostToDevice));
    PLL_HIP_TRY(hipMemcpy(d->d_off, d->off.data(), d->off.size() * sizeof(long),
                          hipMemcpyHostToDevice));
    PLL_HIP_TRY(hipMemcpy(d->d_dists, d->dists.data(), d->dists.size() * sizeof(int),
                          hipMemcpyHostToDevice));
  }
  if (a.rows == 0) return PLL_OK;
  const int* dist = d->d_dists + d->dist_at[kind];
  const long* rows_off = d->d_rows + a.pd * (P + 1);
  const unsigned blocks = (unsigned)std::min<long>((a.rows + 3) / 4, 65536L);
  // LDS copy of the tables: rows (P+1) and offsets (3P) as long, metadata (18P)
  // and the distribution as int; staged when it stays small
  const size_t lds = (size_t)(P + 1 + 3 * P) * sizeof(long) + (size_t)(18 * P + a.dist_len) * sizeof(int);
  const int stage = lds <= 16384 ? 1 : 0;
  const size_t shm = stage ? lds : 0;
  if (elem_bytes == 8)
    hipLaunchKernelGGL((k_transpose<double>), dim3(blocks), dim3(256), shm, 0, a, d->d_meta,
                       d->d_off, rows_off, dist, static_cast<const double*>(src), static_cast<double*>(dst),
                       stage);
  else
    hipLaunchKernelGGL((k_transpose<float>), dim3(blocks), dim3(256), shm, 0, a, d->d_meta,
                       d->d_off, rows_off, dist, static_cast<const float*>(src), static_cast<float*>(dst),
                       stage);
  PLL_HIP_TRY(hipGetLastError());
  PLL_HIP_TRY(hipDeviceSynchronize());
  return PLL_OK;
}

// transpose_x_to_y_real (src/decomp_2d.cc:522-536)
int pll_transpose_x_to_y_real(pll_decomp* d, const void* src, void* dst, int elem_bytes) {
  return transpose(d, 0, src, dst, elem_bytes);
}
// transpose_y_to_x_real (src/decomp_2d.cc:539-553)
int pll_transpose_y_to_x_real(pll_decomp* d, const void* src, void* dst, int elem_bytes) {
  return transpose(d, 1, src, dst, elem_bytes);
}
// transpose_y_to_z_real (src/decomp_2d.cc:556-570)
int pll_transpose_y_to_z_real(pll_decomp* d, const void* src, void* dst, int elem_bytes) {
  return transpose(d, 2, src, dst, elem_bytes);
}
// transpose_z_to_y_real (src/decomp_2d.cc:573-589)
int pll_transpose_z_to_y_real(pll_decomp* d, const void* src, void* dst, int elem_bytes) {
  return transpose(d, 3, src, dst, elem_bytes);
}

}  // extern "C"
