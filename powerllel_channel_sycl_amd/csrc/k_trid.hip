// k_trid.hip -- the wall-normal tridiagonal solve of the pressure Poisson
// equation for every (kx, ky) mode (ssolve_trid, src/mod_poisson_solver.cc:317-357).
//
// The reference stores the forward-eliminated pivots as a full nx*ny*nz
// array (src/mod_poisson_solver.cc:255-282).  The pivots of a column depend
// only on its eigenvalue lambda_x(i) + lambda_y(j), and the perm packing gives
// the real and imaginary parts of a mode the same eigenvalue, so four columns
// share each pivot column.  The table here keeps one pivot column per
// distinct pair: (nx/2+1)*(ny/2+1)*nz values, built on the host with the
// reference's arithmetic, so every pivot is bit-identical to the reference's.
//
// One thread per (i,j) column, coalesced over i, sequential in k; many
// columns per CU hide the latency of the dependent sweeps.  (Staging whole
// columns in LDS saves the second pass over the work array but leaves a CU
// only ~50 FP64 columns of 384 planes: the sweeps then run latency-bound,
// 3x slower at 768x384x384.)  The arithmetic is FP64 in both precision modes
// (the FP32-mixed path stores the work array in FP32 but eliminates in FP64,
// which covers the kx=ky=0 and the near-singular low-wavenumber columns).
// FP64 keeps the reference's divisions (bit-exact); FP32-mixed stores
// reciprocal pivots and multiplies (RECIP).
#include "pll_internal.hpp"

#include <cmath>
#include <type_traits>

namespace pll {

namespace {

constexpr int kTx = 64;
constexpr int kTy = 4;

// perm index -> compact eigenvalue index: 0 -> 0, 1 -> 1 (Nyquist), 2m/2m+1 -> m+1.
__host__ __device__ __forceinline__ int compact_index(int i) { return i <= 1 ? i : (i >> 1) + 1; }

// Loads are issued a chunk of D planes ahead of the sequential recurrence, so
// each sweep waits on HBM latency once per D planes rather than once per plane.
constexpr int kD = 8;

template <typename T, bool RECIP>
__global__ __launch_bounds__(kTx* kTy) void k_trid(int nx, int ny, int nz, T* __restrict__ var,
                                                   const double* __restrict__ a,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ piv, int cxn,
                                                   int cyn) {
  const int i = blockIdx.x * kTx + threadIdx.x;
  const int j = blockIdx.y * kTy + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const long nxy = (long)nx * ny;
  const long P = (long)cxn * cyn;
  // plane bases are uniform across the block; each lane keeps one 32-bit offset
  const int off = j * nx + i;
  const int poff = compact_index(j) * cxn + compact_index(i);
  auto V = [&](int k) -> T& { return (var + (long)k * nxy)[off]; };
  auto B = [&](int k) -> double { return (piv + (long)k * P)[poff]; };

  // forward elimination, src/mod_poisson_solver.cc:334-339:
  // var[k] -= (a[k] / b[k-1]) * var[k-1], k = 1..nz-1
  double prev = (double)V(0);
  T xv[2][kD];
  double xb[2][kD];
  auto load_fwd = [&](int buf, int kb) {
#pragma unroll
    for (int d = 0; d < kD; ++d) {
      const int k = kb + d;
      if (k < nz) {
        xv[buf][d] = V(k);
        xb[buf][d] = B(k - 1);
      }
    }
  };
  auto run_fwd = [&](int buf, int kb) {
#pragma unroll
    for (int d = 0; d < kD; ++d) {
      const int k = kb + d;
      if (k < nz) {
        const double a_tmp = RECIP ? a[k] * xb[buf][d] : a[k] / xb[buf][d];
        const double cur = (double)xv[buf][d] - a_tmp * prev;
        V(k) = (T)cur;
        prev = cur;
      }
    }
  };
  load_fwd(0, 1);
  for (int kb = 1; kb < nz; kb += 2 * kD) {
    if (kb + kD < nz) load_fwd(1, kb + kD);
    run_fwd(0, kb);
    if (kb + 2 * kD < nz) load_fwd(0, kb + 2 * kD);
    if (kb + kD < nz) run_fwd(1, kb + kD);
  }
  // last row, src/mod_poisson_solver.cc:341-346
  const double b_val = B(nz - 1);
  if (RECIP)
    prev = prev * b_val;  // the reciprocal table holds 0 for a zero pivot
  else
    prev = (b_val != 0) ? prev / b_val : 0.0;
  V(nz - 1) = (T)prev;
  // back substitution, src/mod_poisson_solver.cc:348-354:
  // var[k] = (var[k] - c[k] * var[k+1]) / b[k], k = nz-2..0
  auto load_bwd = [&](int buf, int kt) {
#pragma unroll
    for (int d = 0; d < kD; ++d) {
      const int k = kt - d;
      if (k >= 0) {
        xv[buf][d] = V(k);
        xb[buf][d] = B(k);
      }
    }
  };
  auto run_bwd = [&](int buf, int kt) {
#pragma unroll
    for (int d = 0; d < kD; ++d) {
      const int k = kt - d;
      if (k >= 0) {
        const double r = (double)xv[buf][d] - c[k] * prev;
        const double cur = RECIP ? r * xb[buf][d] : r / xb[buf][d];
        V(k) = (T)cur;
        prev = cur;
      }
    }
  };
  load_bwd(0, nz - 2);
  for (int kt = nz - 2; kt >= 0; kt -= 2 * kD) {
    if (kt - kD >= 0) load_bwd(1, kt - kD);
    run_bwd(0, kt);
    if (kt - 2 * kD >= 0) load_bwd(0, kt - 2 * kD);
    if (kt - kD >= 0) run_bwd(1, kt - kD);
  }
}


// ---------------------------------------------------------------------------
// The same solve inside the step, partitioned so that the work array makes one
// HBM pass each way.  A block owns kCw adjacent columns of one row j; each
// column is cut into S segments of M planes, one thread per segment, and the
// segment's right-hand side and reciprocal pivots stay in registers.  The
// pivots are not read from a table: each segment continues the elimination's
// pivot recurrence (init_poisson_solver, src/mod_poisson_solver.cc:273-282)
// from the pivot of the plane below it, which the table `pstart` holds, with
// the reference's operations (FP64 division), so every pivot is the one the
// full table holds, bit for bit; the reciprocal is 0 for a zero pivot:
//   A. forward elimination of the segment from a zero incoming value, with the
//      running product g of the negated multipliers (the segment end's
//      dependence on the incoming value);
//   exchange: the true segment-end values, sequentially over the S segments
//      (Y_s = yhat_s + g_s * Y_{s-1}), through LDS;
//   B. the forward elimination again from the true incoming value (the last
//      row divided by its pivot, src/mod_poisson_solver.cc:341-346);
//   C, exchange, D: the same for the back substitution (:348-354) from the top.
// The segment boundaries are combined with other roundings than the
// sequential sweep, and the pivots are applied as reciprocals, so the result
// differs from k_trid's by FP64 round-off (it is not bit-identical).
// FP32-mixed runs the recurrence with less than one division's work per plane
// instead of two divisions: the multiplier is a[k] times the reciprocal of the
// pivot below (already formed), and the reciprocal is the hardware estimate
// refined by one Newton step (which squares its relative error: far below the
// 6e-8 rounding of the FP32 work array).  The recurrence is the kernel's
// longest dependent chain, and this form of it saves 16 % of the FP32-mixed
// solve at 768x384x384 (375 -> 316 us; a second Newton step costs 7 us); the
// FP64 solve keeps the divisions (its pivots stay bit-exact, and a reciprocal
// saved it nothing).
__device__ __forceinline__ double rcp_newton(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return __builtin_fma(r, __builtin_fma(-q, r, 1.0), r);
}
constexpr int kSegPlanes = 768;  // the deepest column the segments hold
// Planes per segment M (8, 16 or 24) and segments per column (at most S = 16,
// or 32 for columns deeper than 384 planes): short segments where they suffice
// keep the registers of a thread (its M values and M+1 reciprocal pivots) low.
constexpr int seg_planes(int nz) {
  return nz <= 128 ? 8 : (nz <= 256 ? 16 : (nz <= 384 ? 24 : (nz <= 512 ? 16 : 24)));
}
constexpr int seg_count_max(int nz) { return nz <= 384 ? 16 : 32; }
static_assert(seg_planes(kSegPlanes) * seg_count_max(kSegPlanes) >= kSegPlanes &&
                  seg_planes(512) * seg_count_max(512) >= 512 &&
                  seg_planes(384) * seg_count_max(384) >= 384,
              "the segments hold every column they take");

// The sequential combination of the S segment-end values, Y_q = hat_q + dep_q *
// Y_{q-1} (forward, REV false) or from the top (REV true), by the threads of
// one segment (one per column); the LDS reads are issued four at a time ahead
// of the dependent chain.
template <int CW, int S, bool REV>
__device__ __forceinline__ void combine(const double (*hat)[CW], const double (*dep)[CW],
                                        double (*inc)[CW], int nseg, int col) {
  double y = 0.0;
#pragma unroll
  for (int q0 = 0; q0 < S; q0 += 4) {
    double h[4], d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = REV ? nseg - 1 - (q0 + u) : q0 + u;
      if (q0 + u < nseg) {
        h[u] = hat[q][col];
        d[u] = dep[q][col];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = REV ? nseg - 1 - (q0 + u) : q0 + u;
      if (q0 + u < nseg) {
        y = h[u] + d[u] * y;
        inc[q][col] = y;
      }
    }
  }
}

// SegCoef: the diagonal before the eigenvalue b (nz), the compact eigenvalues
// of x (cxn) and y (cyn), and the pivots of the planes below the segments,
// pstart[s * cxn * cyn + cy * cxn + cx] = pivot of plane s*M - 1 (s >= 1).
struct SegCoef {
  const double* b;
  const double* elx;
  const double* ely;
  const double* pstart;
  int pin;  // the (0,0) column's last pivot is pinned to 0
};

template <typename T, int M, int CW, int S>
__global__ __launch_bounds__(CW* S) void k_trid_seg(int nx, int ny, int nz, int nseg,
                                                        T* __restrict__ var,
                                                        const double* __restrict__ a,
                                                        const double* __restrict__ c,
                                                        SegCoef sg, int cxn, int cyn) {
  __shared__ double hat[S][CW];  // segment-end value from a zero incoming value
  __shared__ double dep[S][CW];  // its dependence on the incoming value
  __shared__ double inc[S][CW];  // the true value at the segment end
  __shared__ double sa[kSegPlanes], sb[kSegPlanes], sc[kSegPlanes];
  const int t = threadIdx.x, col = t % CW, s = t / CW;
  const int i = blockIdx.x * CW + col, j = blockIdx.y;
  for (int k = t; k < nz; k += CW * nseg) {
    sa[k] = a[k];
    sb[k] = sg.b[k];
    sc[k] = c[k];
  }
  const bool on = i < nx;
  const long nxy = (long)nx * ny;
  const int k0 = s * M;
  const int n = on ? min(M, nz - k0) : 0;
  T* v = var + (long)j * nx + i;
  const int ci = compact_index(on ? i : 0), cj = compact_index(j);
  // the segment's values (in the field precision: FP32-mixed rounds the
  // forward-eliminated values to FP32 between the sweeps, as k_trid does),
  // loaded first; then the reciprocal pivots of planes k0-1 .. k0+n-1
  T x[M];
  double rb[M + 1];
#pragma unroll
  for (int m = 0; m < M; ++m)
    if (m < n) x[m] = v[(long)(k0 + m) * nxy];
  const double lxy = sg.elx[ci] + sg.ely[cj];
  double piv = k0 > 0 && n > 0 ? sg.pstart[(long)s * cxn * cyn + (long)cj * cxn + ci] : 0.0;
  rb[0] = piv != 0.0 ? 1.0 / piv : 0.0;
  __syncthreads();  // sa, sb, sc
  // init_poisson_solver's recurrence (src/mod_poisson_solver.cc:273-282); a
  // zero pivot only ends a column, so rb[m - 1] is 1/piv wherever k > 0
  constexpr bool kMixed = std::is_same<T, float>::value;
#pragma unroll
  for (int m = 1; m <= M; ++m)
    if (m <= n) {
      const int k = k0 + m - 1;
      double q = sb[k] + lxy;
      if (k > 0) {
        const double a_tmp = kMixed ? sa[k] * rb[m - 1] : sa[k] / piv;
        q -= (a_tmp * sc[k - 1]);
      }
      if (sg.pin && k == nz - 1 && ci == 0 && cj == 0) q = 0.0;
      rb[m] = q != 0.0 ? (kMixed ? rcp_newton(q) : 1.0 / q) : 0.0;
      piv = q;
    }
  // A: forward elimination from a zero incoming value
  {
    double y = 0.0, g = 1.0;
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (m < n) {
        const int k = k0 + m;
        if (k == 0) {
          y = (double)x[0];
          g = 0.0;
        } else {
          const double l = sa[k] * rb[m];
          y = (double)x[m] - l * y;
          g = -l * g;
        }
      }
    if (on) {
      hat[s][col] = y;
      dep[s][col] = g;
    }
  }
  __syncthreads();
  if (s == 0) combine<CW, S, false>(hat, dep, inc, nseg, col);
  __syncthreads();
  // B: forward elimination from the true incoming value; the last row
  {
    double y = s > 0 && on ? inc[s - 1][col] : 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (m < n) {
        const int k = k0 + m;
        y = k == 0 ? (double)x[0] : (double)x[m] - (sa[k] * rb[m]) * y;
        if (k == nz - 1) y = y * rb[m + 1];  // the reciprocal table holds 0 for a zero pivot
        x[m] = (T)y;
      }
  }
  __syncthreads();  // inc is rewritten below
  // C: back substitution from a zero value above the segment
  {
    double xv = 0.0, h = 1.0;
#pragma unroll
    for (int m = M - 1; m >= 0; --m)
      if (m < n) {
        const int k = k0 + m;
        if (k == nz - 1) {
          xv = (double)x[m];
          h = 0.0;
        } else {
          const double q = sc[k] * rb[m + 1];
          xv = (double)x[m] * rb[m + 1] - q * xv;
          h = -q * h;
        }
      }
    if (on) {
      hat[s][col] = xv;
      dep[s][col] = h;
    }
  }
  __syncthreads();
  if (s == 0) combine<CW, S, true>(hat, dep, inc, nseg, col);
  __syncthreads();
  // D: back substitution from the true value above the segment
  {
    double xv = s + 1 < nseg && on ? inc[s + 1][col] : 0.0;
#pragma unroll
    for (int m = M - 1; m >= 0; --m)
      if (m < n) {
        const int k = k0 + m;
        xv = k == nz - 1 ? (double)x[m] : (double)x[m] * rb[m + 1] - (sc[k] * rb[m + 1]) * xv;
        v[(long)k * nxy] = (T)xv;
      }
  }
}

}  // namespace

// init_poisson_solver's eigenvalue, coefficient and pivot set-up
// (src/mod_poisson_solver.cc:230-282, src/mod_fft.cc:196-222) at one GPU,
// boundary types "PP","PP","NN" (src/main.cc:55-57).
int trid_setup(pll_context* c) {
  const int nx = c->g.nx, ny = c->g.ny, nz = c->g.nz;
  std::vector<double> lx(nx), ly(ny);
  pll_eigen_values(nx, lx.data());
  for (int i = 0; i < nx; ++i) lx[i] /= (c->dx * c->dx);
  pll_eigen_values(ny, ly.data());
  for (int j = 0; j < ny; ++j) ly[j] /= (c->dy * c->dy);

  // set_trid_coeff('c', "NN") with dzf0[k] = dzf_global(k), k = 0..nz+1
  // (src/mod_poisson_solver.cc:35-98,258-264)
  const double* dzf0 = c->dzf_global.data();
  std::vector<double> a(nz), b(nz), cc(nz);
  for (int k = 0; k < nz; ++k) {
    a[k] = 2.0 / (dzf0[k + 1] * (dzf0[k] + dzf0[k + 1]));
    cc[k] = 2.0 / (dzf0[k + 1] * (dzf0[k + 2] + dzf0[k + 1]));
  }
  for (int k = 0; k < nz; ++k) b[k] = -a[k] - cc[k];
  const double factor = 1.0;  // 'N'
  b[0] += factor * a[0];
  a[0] *= std::fabs(factor) - 1.0;
  b[nz - 1] += factor * cc[nz - 1];
  cc[nz - 1] *= std::fabs(factor) - 1.0;

  c->cxn = nx / 2 + 1;
  c->cyn = ny / 2 + 1;
  const long P = (long)c->cxn * c->cyn;
  std::vector<double> piv((size_t)P * nz);
  auto rep = [](int ci) { return ci <= 1 ? ci : 2 * (ci - 1); };  // a perm index of that class
  for (int cy = 0; cy < c->cyn; ++cy)
    for (int cx = 0; cx < c->cxn; ++cx) {
      const double lxy = lx[rep(cx)] + ly[rep(cy)];
      double* col = piv.data() + (long)cy * c->cxn + cx;
      for (int k = 0; k < nz; ++k) col[k * P] = b[k] + lxy;
      for (int k = 1; k < nz; ++k) {
        const double a_tmp = a[k] / col[(k - 1) * P];
        col[k * P] -= (a_tmp * cc[k - 1]);
      }
    }
  // FP32-mixed always pins: the reference's round-off pivot would divide the
  // FP32 round-off of the mean right-hand side into a constant of ~1e6 in p,
  // which an FP32 p cannot carry next to O(1) pressure differences
  if (c->cfg.pin_singular || c->prec != PLL_FP64) piv[(size_t)(nz - 1) * P] = 0.0;
  // the partitioned solve of the step: b, the compact eigenvalues and the
  // pivots of the planes below its segments (it rebuilds the others)
  std::vector<double> seg;
  if (nz <= kSegPlanes) {
    const int M = seg_planes(nz), nseg = (nz + M - 1) / M;
    seg.assign(nz + c->cxn + c->cyn + (size_t)nseg * P, 0.0);
    for (int k = 0; k < nz; ++k) seg[k] = b[k];
    for (int cx = 0; cx < c->cxn; ++cx) seg[nz + cx] = lx[rep(cx)];
    for (int cy = 0; cy < c->cyn; ++cy) seg[nz + c->cxn + cy] = ly[rep(cy)];
    double* ps = seg.data() + nz + c->cxn + c->cyn;
    for (int s = 1; s < nseg; ++s)
      for (long q = 0; q < P; ++q) ps[s * P + q] = piv[(size_t)(s * M - 1) * P + q];
    c->seg_pin = (c->cfg.pin_singular || c->prec != PLL_FP64) ? 1 : 0;
  }
  // FP32-mixed multiplies by reciprocal pivots (0 stands for a zero pivot:
  // the reference's b == 0 branch gives 0, src/mod_poisson_solver.cc:341-346)
  if (c->prec != PLL_FP64)
    for (double& x : piv) x = x != 0.0 ? 1.0 / x : 0.0;

  PLL_HIP_TRY(ctx_malloc(c, &c->d_a, nz * sizeof(double)));
  PLL_HIP_TRY(ctx_malloc(c, &c->d_c, nz * sizeof(double)));
  PLL_HIP_TRY(ctx_malloc(c, &c->d_piv, piv.size() * sizeof(double)));
  PLL_HIP_TRY(hipMemcpy(c->d_a, a.data(), nz * sizeof(double), hipMemcpyHostToDevice));
  PLL_HIP_TRY(hipMemcpy(c->d_c, cc.data(), nz * sizeof(double), hipMemcpyHostToDevice));
  PLL_HIP_TRY(hipMemcpy(c->d_piv, piv.data(), piv.size() * sizeof(double), hipMemcpyHostToDevice));
  if (!seg.empty()) {
    PLL_HIP_TRY(ctx_malloc(c, &c->d_seg, seg.size() * sizeof(double)));
    PLL_HIP_TRY(hipMemcpy(c->d_seg, seg.data(), seg.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return PLL_OK;
}

int launch_trid(pll_context* c) {
  TimerScope ts(c, K_TRID);
  const dim3 grid((c->g.nx + kTx - 1) / kTx, (c->g.ny + kTy - 1) / kTy);
  if (c->prec == PLL_FP64)
    hipLaunchKernelGGL((k_trid<double, false>), grid, dim3(kTx, kTy), 0, c->stream, c->g.nx,
                       c->g.ny, c->g.nz, static_cast<double*>(c->spec), c->d_a, c->d_c, c->d_piv,
                       c->cxn, c->cyn);
  else
    hipLaunchKernelGGL((k_trid<float, true>), grid, dim3(kTx, kTy), 0, c->stream, c->g.nx,
                       c->g.ny, c->g.nz, static_cast<float*>(c->spec), c->d_a, c->d_c, c->d_piv,
                       c->cxn, c->cyn);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

namespace {
// Columns per block, planes per segment and segments per column of the
// partitioned solve: FP32-mixed takes 32 columns of up to 16 segments (128-byte
// rows; 16 columns ran 374 against 316 us at 768x384x384) and 16 columns of 32
// segments (512-thread blocks; 32 columns, 1024 threads, ran 9 % slower at
// 1536x768x512), FP64 16 columns.
template <typename T, int S>
constexpr int seg_cols() {
  return std::is_same<T, float>::value && S <= 16 ? 32 : 16;
}
template <typename T, int M, int S>
void trid_seg(pll_context* c) {
  constexpr int kCw = seg_cols<T, S>();
  const int nseg = (c->g.nz + M - 1) / M;
  const dim3 grid((c->g.nx + kCw - 1) / kCw, c->g.ny);
  const int nz = c->g.nz;
  const SegCoef sg{c->d_seg, c->d_seg + nz, c->d_seg + nz + c->cxn,
                   c->d_seg + nz + c->cxn + c->cyn, c->seg_pin};
  hipLaunchKernelGGL((k_trid_seg<T, M, kCw, S>), grid, dim3(kCw * nseg), 0, c->stream, c->g.nx,
                     c->g.ny, nz, nseg, static_cast<T*>(c->spec), c->d_a, c->d_c, sg, c->cxn,
                     c->cyn);
}
template <typename T>
void trid_seg_t(pll_context* c) {
  const int nz = c->g.nz;
  if (nz <= 384) {
    switch (seg_planes(nz)) {
      case 8: trid_seg<T, 8, 16>(c); break;
      case 16: trid_seg<T, 16, 16>(c); break;
      default: trid_seg<T, 24, 16>(c); break;
    }
  } else if (nz <= 512) {
    trid_seg<T, 16, 32>(c);
  } else {
    trid_seg<T, 24, 32>(c);
  }
}
}  // namespace

int launch_trid_step(pll_context* c) {
  // columns deeper than the segments hold take the sequential kernel
  // (FP32-mixed at 1536x768x512: 2397 us partitioned against 2533 us
  // sequential, since its recurrence runs on reciprocals)
  if (!c->d_seg || c->g.nz > kSegPlanes) return launch_trid(c);
  TimerScope ts(c, K_TRID);
  if (c->prec == PLL_FP64)
    trid_seg_t<double>(c);
  else
    trid_seg_t<float>(c);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

}  // namespace pll
