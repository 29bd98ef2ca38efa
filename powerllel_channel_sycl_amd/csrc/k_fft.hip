// k_fft.hip -- the pressure-Poisson transforms (src/mod_fft.cc:60-187) as
// hand-written LDS kernels.
//
// x direction: each row of nx reals is transformed as a complex sequence of
// nx/2 points (even samples real, odd samples imaginary) followed by the
// split step that yields the half spectrum; the result is written straight in
// the reference's perm packing [X0, X(n/2), Re X1, Im X1, ...].  The inverse
// runs the same steps backwards and folds in the 1/(nx*ny) scale.
//
// y direction: the work array keeps its [k][j][i] order (no transpose pass).
// A block loads a tile of adjacent perm columns of one k plane; columns
// (2q, 2q+1) form one complex sequence along j, whose DFT is split into the two
// real-column DFTs and written back in perm packing along j.  This is the
// reference's per-column real DFT (src/mod_fft.cc:122-187) computed two
// columns at a time.
//
// Complex transforms use the Stockham autosort formulation with radix 4, 2,
// 3 and 5 passes ping-ponging between two LDS buffers.  Twiddles are computed
// on the host in long double and rounded to the field precision.
#include "fft_radix.hpp"
#include "pll_internal.hpp"

#include <cmath>
#include <complex>
#include <cstdlib>

namespace pll {

namespace {

// One Stockham butterfly: sequence q, butterfly jj of a radix-R pass with span Ns.
template <typename T, int R, int S>
__device__ __forceinline__ void butterfly(const cpx<T>* __restrict__ a, cpx<T>* __restrict__ b,
                                          const cpx<T>* __restrict__ tw, int jj, int q, int Ns,
                                          int LR, int es, int ss) {
  const int k = jj % Ns;
  cpx<T> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = a[q * ss + (jj + r * LR) * es];
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cpx<T> w = tw[k * (R - 1) + r - 1];
      if (S > 0) w.im = -w.im;
      v[r] = cmul(v[r], w);
    }
  }
  dft<T, R, S>(v);
  const int base = (jj / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) b[q * ss + (base + r * Ns) * es] = v[r];
}

// All passes of a batch of nseq complex transforms of length pl.L in LDS.
// Element m of sequence q lives at q*ss + m*es.  With seq_fast, consecutive
// threads take consecutive sequences (for the y layout, where sequences are
// adjacent in LDS); otherwise consecutive butterflies of one sequence.
// Returns the buffer that holds the result.
template <typename T, int S>
__device__ cpx<T>* lds_fft(cpx<T>* a, cpx<T>* b, const FftPlan& pl, const cpx<T>* __restrict__ tw,
                           int nseq, int es, int ss, bool seq_fast) {
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p], Ns = pl.ns[p], LR = pl.L / R;
    const cpx<T>* twp = tw + pl.twoff[p];
    const int total = nseq * LR;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      int q, jj;
      if (seq_fast) {
        q = t % nseq;
        jj = t / nseq;
      } else {
        jj = t % LR;
        q = t / LR;
      }
      switch (R) {
        case 4: butterfly<T, 4, S>(a, b, twp, jj, q, Ns, LR, es, ss); break;
        case 2: butterfly<T, 2, S>(a, b, twp, jj, q, Ns, LR, es, ss); break;
        case 3: butterfly<T, 3, S>(a, b, twp, jj, q, Ns, LR, es, ss); break;
        default: butterfly<T, 5, S>(a, b, twp, jj, q, Ns, LR, es, ss); break;
      }
    }
    __syncthreads();
    cpx<T>* t = a;
    a = b;
    b = t;
  }
  return a;
}

constexpr int kFftThreads = 256;

// ---- x direction ------------------------------------------------------------------
// execute_fft_fwd_xpen (src/mod_fft.cc:60-88): rows r = j + ny*k of p's interior.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_fwd(Geom g, const T* __restrict__ p,
                                                           T* __restrict__ spec, FftPlan pl,
                                                           const cpx<T>* __restrict__ tw,
                                                           const cpx<T>* __restrict__ rtw,
                                                           int batch) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.nx, h = n / 2;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + (long)batch * h;
  const long nrows = (long)g.ny * g.nz;
  const long r0 = (long)blockIdx.x * batch;
  T* Ar = reinterpret_cast<T*>(A);
  for (int e = threadIdx.x; e < batch * n; e += blockDim.x) {
    const int b = e / n, i = e - b * n;
    const long r = r0 + b;
    T val = T(0);
    if (r < nrows) {
      const int j = (int)(r % g.ny), k = (int)(r / g.ny);
      val = p[g.idx(i + 1, j + 1, k + 1)];
    }
    Ar[e] = val;
  }
  __syncthreads();
  const cpx<T>* Z = lds_fft<T, -1>(A, B, pl, tw, batch, 1, h, false);
  for (int e = threadIdx.x; e < batch * h; e += blockDim.x) {
    const int b = e / h, kk = e - b * h;
    const long r = r0 + b;
    if (r >= nrows) continue;
    T* out = spec + r * n;
    const cpx<T> zk = Z[b * h + kk];
    if (kk == 0) {
      out[0] = zk.re + zk.im;  // X0
      out[1] = zk.re - zk.im;  // X(n/2)
    } else {
      const cpx<T> zm = Z[b * h + (h - kk)];
      const T er = T(0.5) * (zk.re + zm.re), ei = T(0.5) * (zk.im - zm.im);
      const T orr = T(0.5) * (zk.im + zm.im), oi = T(-0.5) * (zk.re - zm.re);
      const cpx<T> w = rtw[kk];
      out[2 * kk] = er + (w.re * orr - w.im * oi);
      out[2 * kk + 1] = ei + (w.re * oi + w.im * orr);
    }
  }
}

// execute_fft_bwd_xpen (src/mod_fft.cc:90-120): perm rows -> p interior * norm.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_x_bwd(Geom g, const T* __restrict__ spec,
                                                           T* __restrict__ p, FftPlan pl,
                                                           const cpx<T>* __restrict__ tw,
                                                           const cpx<T>* __restrict__ rtw,
                                                           int batch, T norm) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.nx, h = n / 2;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + (long)batch * h;
  const long nrows = (long)g.ny * g.nz;
  const long r0 = (long)blockIdx.x * batch;
  T* Ar = reinterpret_cast<T*>(A);
  for (int e = threadIdx.x; e < batch * n; e += blockDim.x) {
    const int b = e / n;
    const long r = r0 + b;
    Ar[e] = (r < nrows) ? spec[r * n + (e - b * n)] : T(0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < batch * h; e += blockDim.x) {
    const int b = e / h, kk = e - b * h;
    const T* pr = Ar + (long)b * n;
    cpx<T> xk, xm;
    if (kk == 0) {
      xk = cpx<T>{pr[0], T(0)};
      xm = cpx<T>{pr[1], T(0)};
    } else {
      xk = cpx<T>{pr[2 * kk], pr[2 * kk + 1]};
      xm = cpx<T>{pr[2 * (h - kk)], pr[2 * (h - kk) + 1]};
    }
    const cpx<T> s{xk.re + xm.re, xk.im - xm.im};  // Xk + conj(X(h-k))
    const cpx<T> d{xk.re - xm.re, xk.im + xm.im};  // Xk - conj(X(h-k))
    const cpx<T> w{rtw[kk].re, -rtw[kk].im};       // exp(+2 pi i k/n)
    const cpx<T> wd = cmul(w, d);
    B[b * h + kk] = cpx<T>{s.re - wd.im, s.im + wd.re};
  }
  __syncthreads();
  const cpx<T>* z = lds_fft<T, +1>(B, A, pl, tw, batch, 1, h, false);
  const T* zr = reinterpret_cast<const T*>(z);
  for (int e = threadIdx.x; e < batch * n; e += blockDim.x) {
    const int b = e / n, i = e - b * n;
    const long r = r0 + b;
    if (r >= nrows) continue;
    const int j = (int)(r % g.ny), k = (int)(r / g.ny);
    p[g.idx(i + 1, j + 1, k + 1)] = norm * zr[e];
  }
}

// ---- y direction ------------------------------------------------------------------
// execute_fft_fwd_ypen (src/mod_fft.cc:122-153) on a tile of `tile` perm columns.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_y_fwd(Geom g, T* __restrict__ spec,
                                                           FftPlan pl,
                                                           const cpx<T>* __restrict__ tw,
                                                           int tile) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.ny, Q = tile / 2;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + (long)n * Q;
  const int k = blockIdx.y, i0 = blockIdx.x * tile;
  T* plane = spec + (long)k * g.nx * g.ny + i0;
  T* Ar = reinterpret_cast<T*>(A);
  for (int e = threadIdx.x; e < n * tile; e += blockDim.x) {
    const int j = e / tile, cidx = e - j * tile;
    Ar[e] = plane[(long)j * g.nx + cidx];
  }
  __syncthreads();
  const cpx<T>* Z = lds_fft<T, -1>(A, B, pl, tw, Q, Q, 1, true);
  T* O = reinterpret_cast<T*>(Z == A ? B : A);
  for (int e = threadIdx.x; e < Q * (n / 2 + 1); e += blockDim.x) {
    const int q = e % Q, m = e / Q;
    const cpx<T> zm = Z[m * Q + q];
    const cpx<T> zc = Z[((n - m) % n) * Q + q];
    const T ar = T(0.5) * (zm.re + zc.re), ai = T(0.5) * (zm.im - zc.im);
    const T br = T(0.5) * (zm.im + zc.im), bi = T(-0.5) * (zm.re - zc.re);
    if (m == 0) {
      O[2 * q] = ar;
      O[2 * q + 1] = br;
    } else if (m == n / 2) {
      O[tile + 2 * q] = ar;
      O[tile + 2 * q + 1] = br;
    } else {
      O[(2 * m) * tile + 2 * q] = ar;
      O[(2 * m + 1) * tile + 2 * q] = ai;
      O[(2 * m) * tile + 2 * q + 1] = br;
      O[(2 * m + 1) * tile + 2 * q + 1] = bi;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * tile; e += blockDim.x) {
    const int j = e / tile, cidx = e - j * tile;
    plane[(long)j * g.nx + cidx] = O[e];
  }
}

// execute_fft_bwd_ypen (src/mod_fft.cc:155-187), unscaled.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_y_bwd(Geom g, T* __restrict__ spec,
                                                           FftPlan pl,
                                                           const cpx<T>* __restrict__ tw,
                                                           int tile) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.ny, Q = tile / 2;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + (long)n * Q;
  const int k = blockIdx.y, i0 = blockIdx.x * tile;
  T* plane = spec + (long)k * g.nx * g.ny + i0;
  T* Ar = reinterpret_cast<T*>(A);
  for (int e = threadIdx.x; e < n * tile; e += blockDim.x) {
    const int j = e / tile, cidx = e - j * tile;
    Ar[e] = plane[(long)j * g.nx + cidx];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < Q * n; e += blockDim.x) {
    const int q = e % Q, m = e / Q;
    const T* pa = Ar + 2 * q;  // perm column 2q, stride tile
    const T* pb = pa + 1;      // perm column 2q+1
    cpx<T> z;
    if (m == 0) {
      z = cpx<T>{pa[0], pb[0]};
    } else if (m == n / 2) {
      z = cpx<T>{pa[tile], pb[tile]};
    } else if (m < n / 2) {
      const T are = pa[(2 * m) * tile], aim = pa[(2 * m + 1) * tile];
      const T bre = pb[(2 * m) * tile], bim = pb[(2 * m + 1) * tile];
      z = cpx<T>{are - bim, aim + bre};
    } else {
      const int mm = n - m;
      const T are = pa[(2 * mm) * tile], aim = pa[(2 * mm + 1) * tile];
      const T bre = pb[(2 * mm) * tile], bim = pb[(2 * mm + 1) * tile];
      z = cpx<T>{are + bim, bre - aim};
    }
    B[m * Q + q] = z;
  }
  __syncthreads();
  const cpx<T>* z = lds_fft<T, +1>(B, A, pl, tw, Q, Q, 1, true);
  const T* zr = reinterpret_cast<const T*>(z);
  for (int e = threadIdx.x; e < n * tile; e += blockDim.x) {
    const int j = e / tile, cidx = e - j * tile;
    plane[(long)j * g.nx + cidx] = zr[e];
  }
}

// ============================================================================
// Compile-time transforms for the lengths of the supported grids.  One
// padded LDS buffer per block; each pass reads its butterfly inputs into
// registers, synchronises, and writes the outputs back in place, so every
// thread does at most one radix-R butterfly per pass.  Index arithmetic is
// constant-folded and the 6/8/12/16 butterflies keep the pass count at 2-3.
// ============================================================================
template <int L>
struct Shape {
  static constexpr int np = 0;
  static constexpr int r[1] = {0};
};
#define PLL_SHAPE(LEN, ...)                          \
  template <>                                        \
  struct Shape<LEN> {                                \
    static constexpr int r[] = {__VA_ARGS__};        \
    static constexpr int np = sizeof(r) / sizeof(int); \
  };
PLL_SHAPE(4, 4)
PLL_SHAPE(6, 6)
PLL_SHAPE(12, 12)
PLL_SHAPE(20, 4, 5)
PLL_SHAPE(32, 8, 4)
PLL_SHAPE(48, 8, 6)
PLL_SHAPE(64, 8, 8)
PLL_SHAPE(128, 16, 8)
PLL_SHAPE(192, 16, 12)
PLL_SHAPE(256, 16, 16)
PLL_SHAPE(384, 8, 8, 6)
PLL_SHAPE(768, 8, 8, 12)
#undef PLL_SHAPE

template <class SH>
constexpr int shape_rmin() {
  int m = 1 << 30;
  for (int p = 0; p < SH::np; ++p) m = SH::r[p] < m ? SH::r[p] : m;
  return m;
}
template <class SH, int P>
constexpr int shape_span() {
  int s = 1;
  for (int p = 0; p < P; ++p) s *= SH::r[p];
  return s;
}
// threads per sequence: every thread owns one butterfly of the smallest radix
template <int L>
constexpr int tps() {
  return L / shape_rmin<Shape<L>>();
}
// one pad slot per 16 complex values spreads strided accesses over the banks
__device__ __forceinline__ int pad(int x) { return x + (x >> 4); }
constexpr int padded(int n) { return n + (n >> 4) + 1; }
constexpr int kLdsMax = 64 * 1024;

// rows per block of the x transforms (about 256 threads, LDS <= 64 KiB)
template <typename T, int L>
constexpr int x_rows() {
  int nb = tps<L>() >= 256 ? 1 : 256 / tps<L>();
  while (nb > 1 && padded(nb * L) * 2 * (int)sizeof(T) > kLdsMax) nb /= 2;
  return nb;
}
// sequences per block of the y transforms: 128-byte rows, at most 1024 threads
// and 64 KiB of LDS
template <typename T, int L>
constexpr int y_seqs() {
  int q = 128 / (2 * (int)sizeof(T));
  while (q > 1 && (q * tps<L>() > 1024 || padded(L * q) * 2 * (int)sizeof(T) > kLdsMax)) q /= 2;
  return q;
}


template <typename T, int S, int L, int R, int NS, int ES>
__device__ __forceinline__ void stockham_pass(cpx<T>* buf, const cpx<T>* __restrict__ twL,
                                              int soff, int jj) {
  constexpr int LR = L / R;
  cpx<T> v[R];
  const bool active = jj < LR;
  if (active) {
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = buf[pad(soff + (jj + r * LR) * ES)];
  }
  __syncthreads();
  if (active) {
    const int k = jj % NS;
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = twL[r * k * (L / (NS * R))];
        if (S > 0) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
    }
    dft<T, R, S>(v);
    const int base = (jj / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[pad(soff + (base + r * NS) * ES)] = v[r];
  }
  __syncthreads();
}

template <typename T, int S, int L, int ES, int P = 0>
__device__ __forceinline__ void stockham(cpx<T>* buf, const cpx<T>* __restrict__ twL, int soff,
                                         int jj) {
  using SH = Shape<L>;
  if constexpr (P < SH::np) {
    stockham_pass<T, S, L, SH::r[P], shape_span<SH, P>(), ES>(buf, twL, soff, jj);
    stockham<T, S, L, ES, P + 1>(buf, twL, soff, jj);
  }
}

template <typename T>
__device__ __forceinline__ T& real_at(cpx<T>* buf, int e) {
  return reinterpret_cast<T*>(buf)[2 * pad(e >> 1) + (e & 1)];
}

// the largest divisor of n that is at most m
constexpr int chunk_of(int n, int m) {
  int c = m < n ? m : n;
  while (n % c) --c;
  return c;
}

// The velocity divergence of calcRHS (src/mod_calcRHS.cc:7-33), for a forward
// x transform that reads it straight from u, v, w instead of from p.
template <typename T>
struct RhsSrc {
  const T* u;
  const T* v;
  const T* w;
  const T* dzf_inv;
  T dti, dtidxi, dtidyi;
  // the velocity's boundary update did not run: the backward neighbours in
  // the halo are read from their periodic images and w is the wall value 0
  // on planes 0 and nz (src/mod_updateBound.cc:32-144)
  int bcfold;
};

// execute_fft_fwd_xpen (src/mod_fft.cc:60-88), L = nx/2 points per row.  With
// RHS the rows are calcRHS's output computed on load (nothing is stored in p).
template <typename T, int L, bool RHS>
__global__ __launch_bounds__((x_rows<T, L>() * tps<L>())) void k_xf(Geom g, const T* __restrict__ p,
                                                               T* __restrict__ spec,
                                                               const cpx<T>* __restrict__ twL,
                                                               const cpx<T>* __restrict__ rtw,
                                                               long nrows, RhsSrc<T> rs) {
  constexpr int TPS = tps<L>(), NB = x_rows<T, L>(), NT = NB * TPS, N = 2 * L;
  __shared__ cpx<T> buf[padded(NB * L)];
  __shared__ long rowbase[NB];
  __shared__ T rowdz[NB];
  __shared__ long rowvs[NB];     // offset of the v row below (j-1, or its image)
  __shared__ unsigned char roww[NB];  // bit 0: w below is the wall 0; bit 1: w is the wall 0
  const long row0 = (long)xcd_remap(blockIdx.x, gridDim.x) * NB;
  if (threadIdx.x < NB) {
    const long r = row0 + threadIdx.x;
    const int j = (int)(r % g.ny), k = (int)(r / g.ny);
    rowbase[threadIdx.x] = r < nrows ? g.idx(1, j + 1, k + 1) : -1;
    if (RHS) {
      rowdz[threadIdx.x] = r < nrows ? rs.dti * rs.dzf_inv[k + 1] : T(0);
      const bool f = rs.bcfold != 0;
      rowvs[threadIdx.x] = (f && j == 0) ? (long)(g.ny - 1) * g.sx : -g.sx;
      roww[threadIdx.x] = (unsigned char)((f && k == 0 ? 1 : 0) | (f && k + 1 == g.nz ? 2 : 0));
    }
  }
  __syncthreads();
  // NB*N values, NLD per thread, loaded in chunks of CH whose loads are all
  // issued before the first is used
  constexpr int NLD = NB * N / NT, CH = RHS ? chunk_of(NLD, 6) : NLD;
  static_assert((NB * N) % NT == 0 && NLD % CH == 0, "whole rows per block");
#pragma unroll
  for (int u0 = 0; u0 < NLD; u0 += CH) {
    T a0[CH], a1[CH], a2[CH], a3[CH], a4[CH], a5[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int e = threadIdx.x + (u0 + u) * NT;
      const int s = e / N, i = e % N;
      const long b = rowbase[s];
      a0[u] = a1[u] = a2[u] = a3[u] = a4[u] = a5[u] = T(0);
      if (b >= 0) {
        const long c = b + i;
        if (RHS) {
          const long cw = (i == 0 && rs.bcfold) ? b + (N - 1) : c - 1;  // u(0) = u(nx)
          const unsigned char wf = roww[s];
          a0[u] = rs.u[c];
          a1[u] = rs.u[cw];
          a2[u] = rs.v[c];
          a3[u] = rs.v[c + rowvs[s]];
          if (!(wf & 2)) a4[u] = rs.w[c];
          if (!(wf & 1)) a5[u] = rs.w[c - g.sxy];
        } else {
          a0[u] = p[c];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int e = threadIdx.x + (u0 + u) * NT;
      const int s = e / N, i = e % N;
      T val = a0[u];
      if (RHS) {
        // src/mod_calcRHS.cc:21-28, same association as k_calc_rhs
        val = (a0[u] - a1[u]) * rs.dtidxi + (a2[u] - a3[u]) * rs.dtidyi + (a4[u] - a5[u]) * rowdz[s];
      }
      real_at(buf, s * L * 2 + i) = val;
    }
  }
  __syncthreads();
  const int s = threadIdx.x / TPS, jj = threadIdx.x % TPS;
  stockham<T, -1, L, 1>(buf, twL, s * L, jj);
  const long r = row0 + s;
  if (r >= nrows) return;
  cpx<T>* out = reinterpret_cast<cpx<T>*>(spec + r * N);
  for (int kk = jj; kk < L; kk += TPS) {
    const cpx<T> zk = buf[pad(s * L + kk)];
    if (kk == 0) {
      out[0] = cpx<T>{zk.re + zk.im, zk.re - zk.im};  // X0, X(n/2)
    } else {
      const cpx<T> zm = buf[pad(s * L + (L - kk))];
      const T er = T(0.5) * (zk.re + zm.re), ei = T(0.5) * (zk.im - zm.im);
      const T orr = T(0.5) * (zk.im + zm.im), oi = T(-0.5) * (zk.re - zm.re);
      const cpx<T> w = rtw[kk];
      out[kk] = cpx<T>{er + (w.re * orr - w.im * oi), ei + (w.re * oi + w.im * orr)};
    }
  }
}

// execute_fft_bwd_xpen (src/mod_fft.cc:90-120): perm rows -> p interior * norm.
// With HALO (the step sequence, where updateBoundP follows at once) a row is
// stored as aligned pairs p(0..nx+1), the two x-halo cells holding their
// periodic values p(nx) and p(1) -- what updateBoundP writes there.
template <typename T, int L, bool HALO>
__global__ __launch_bounds__((x_rows<T, L>() * tps<L>())) void k_xb(Geom g, const T* __restrict__ spec,
                                                               T* __restrict__ p,
                                                               const cpx<T>* __restrict__ twL,
                                                               const cpx<T>* __restrict__ rtw,
                                                               long nrows, T norm) {
  constexpr int TPS = tps<L>(), NB = x_rows<T, L>(), NT = NB * TPS, N = 2 * L;
  constexpr int RMIN = shape_rmin<Shape<L>>();
  __shared__ cpx<T> buf[padded(NB * L)];
  __shared__ long rowbase[NB];
  const long row0 = (long)xcd_remap(blockIdx.x, gridDim.x) * NB;
  if (threadIdx.x < NB) {
    const long r = row0 + threadIdx.x;
    const int j = (int)(r % g.ny), k = (int)(r / g.ny);
    rowbase[threadIdx.x] = r < nrows ? g.idx(HALO ? 0 : 1, j + 1, k + 1) : -1;
  }
  const cpx<T>* specc = reinterpret_cast<const cpx<T>*>(spec);
  {
    constexpr int NLD = NB * L / NT;
    static_assert((NB * L) % NT == 0, "whole rows per block");
    cpx<T> t[NLD];
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int e = threadIdx.x + u * NT;
      const long r = row0 + e / L;
      t[u] = r < nrows ? specc[r * L + e % L] : cpx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int u = 0; u < NLD; ++u) buf[pad(threadIdx.x + u * NT)] = t[u];
  }
  __syncthreads();
  const int s = threadIdx.x / TPS, jj = threadIdx.x % TPS;
  // perm slot kk holds (Re Xkk, Im Xkk); slot 0 holds (X0, X(n/2))
  cpx<T> z[RMIN];
#pragma unroll
  for (int u = 0; u < RMIN; ++u) {
    const int kk = jj + u * TPS;
    const cpx<T> a = buf[pad(s * L + kk)];
    cpx<T> xk, xm;
    if (kk == 0) {
      xk = cpx<T>{a.re, T(0)};
      xm = cpx<T>{a.im, T(0)};
    } else {
      xk = a;
      xm = buf[pad(s * L + (L - kk))];
    }
    const cpx<T> sm{xk.re + xm.re, xk.im - xm.im};  // Xk + conj(X(h-k))
    const cpx<T> d{xk.re - xm.re, xk.im + xm.im};   // Xk - conj(X(h-k))
    const cpx<T> w{rtw[kk].re, -rtw[kk].im};        // exp(+2 pi i k/n)
    const cpx<T> wd = cmul(w, d);
    z[u] = cpx<T>{sm.re - wd.im, sm.im + wd.re};
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < RMIN; ++u) buf[pad(s * L + jj + u * TPS)] = z[u];
  __syncthreads();
  stockham<T, +1, L, 1>(buf, twL, s * L, jj);
  if (HALO) {
    // pair q = (p(2q), p(2q+1)); p(i) is element i-1 of the row, p(0) = p(nx), p(nx+1) = p(1)
    constexpr int NP = L + 1, NST = (NB * NP + NT - 1) / NT;
    cpx<T> t[NST];
#pragma unroll
    for (int u = 0; u < NST; ++u) {
      const int e = threadIdx.x + u * NT;
      const int s2 = e / NP, q = e % NP;
      if (e < NB * NP) {
        const int e0 = q == 0 ? N - 1 : 2 * q - 1, e1 = q == L ? 0 : 2 * q;
        t[u] = cpx<T>{norm * real_at(buf, s2 * N + e0), norm * real_at(buf, s2 * N + e1)};
      }
    }
#pragma unroll
    for (int u = 0; u < NST; ++u) {
      const int e = threadIdx.x + u * NT;
      const int s2 = e / NP, q = e % NP;
      if (e < NB * NP && rowbase[s2] >= 0) reinterpret_cast<cpx<T>*>(p + rowbase[s2])[q] = t[u];
    }
  } else {
    constexpr int NST = NB * N / NT;
    static_assert((NB * N) % NT == 0, "whole rows per block");
    T t[NST];
#pragma unroll
    for (int u = 0; u < NST; ++u) t[u] = norm * real_at(buf, threadIdx.x + u * NT);
#pragma unroll
    for (int u = 0; u < NST; ++u) {
      const int e = threadIdx.x + u * NT;
      const long b = rowbase[e / N];
      if (b >= 0) p[b + e % N] = t[u];
    }
  }
}

// A y tile between global memory and LDS: NE complex slots, slot e = j*Q + q2
// is the pair of columns (2 q2, 2 q2 + 1) of row j.  Each thread first issues
// all of its loads (or LDS reads), then all of its LDS writes (or stores), so
// a thread keeps ceil(NE/NT) memory operations in flight instead of one.
template <typename T, int NE, int NT, int Q>
__device__ __forceinline__ void load_tile(cpx<T>* buf, const T* plane, int nx) {
  constexpr int NLD = (NE + NT - 1) / NT;
  cpx<T> t[NLD];
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = threadIdx.x + u * NT;
    if (NE % NT == 0 || e < NE)
      t[u] = reinterpret_cast<const cpx<T>*>(plane + (long)(e / Q) * nx)[e % Q];
  }
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = threadIdx.x + u * NT;
    if (NE % NT == 0 || e < NE) buf[pad(e)] = t[u];
  }
}
template <typename T, int NE, int NT, int Q>
__device__ __forceinline__ void store_tile(const cpx<T>* buf, T* plane, int nx) {
  constexpr int NLD = (NE + NT - 1) / NT;
  cpx<T> t[NLD];
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = threadIdx.x + u * NT;
    if (NE % NT == 0 || e < NE) t[u] = buf[pad(e)];
  }
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = threadIdx.x + u * NT;
    if (NE % NT == 0 || e < NE) reinterpret_cast<cpx<T>*>(plane + (long)(e / Q) * nx)[e % Q] = t[u];
  }
}

// execute_fft_fwd_ypen (src/mod_fft.cc:122-153) on a tile of 2Q perm columns.
template <typename T, int L>
__global__ __launch_bounds__((y_seqs<T, L>() * tps<L>())) void k_yf(Geom g, T* __restrict__ spec,
                                                                  const cpx<T>* __restrict__ twL,
                                                                  int ntiles) {
  constexpr int TPS = tps<L>(), Q = y_seqs<T, L>(), NT = Q * TPS, W = 2 * Q;
  __shared__ cpx<T> buf[padded(L * Q)];
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = (int)(b % (unsigned)ntiles), k = (int)(b / (unsigned)ntiles);
  T* plane = spec + (long)k * g.nx * g.ny + (long)tile * W;
  // a slot holds the values of columns (2q, 2q+1) of one row: one 2-element load
  load_tile<T, L * Q, NT, Q>(buf, plane, g.nx);
  __syncthreads();
  const int q = threadIdx.x % Q, jj = threadIdx.x / Q;
  stockham<T, -1, L, Q>(buf, twL, q, jj);
#pragma unroll
  for (int mu = 0; mu <= L / 2 / TPS; ++mu) {
    const int m = jj + mu * TPS;
    if (m > L / 2) break;
    const cpx<T> zm = buf[pad(m * Q + q)];
    const cpx<T> zc = buf[pad(((L - m) % L) * Q + q)];
    const T ar = T(0.5) * (zm.re + zc.re), ai = T(0.5) * (zm.im - zc.im);
    const T br = T(0.5) * (zm.im + zc.im), bi = T(-0.5) * (zm.re - zc.re);
    if (m == 0) {
      reinterpret_cast<cpx<T>*>(plane)[q] = cpx<T>{ar, br};
    } else if (m == L / 2) {
      reinterpret_cast<cpx<T>*>(plane + g.nx)[q] = cpx<T>{ar, br};
    } else {
      reinterpret_cast<cpx<T>*>(plane + (long)(2 * m) * g.nx)[q] = cpx<T>{ar, br};
      reinterpret_cast<cpx<T>*>(plane + (long)(2 * m + 1) * g.nx)[q] = cpx<T>{ai, bi};
    }
  }
}

// execute_fft_bwd_ypen (src/mod_fft.cc:155-187), unscaled.
template <typename T, int L>
__global__ __launch_bounds__((y_seqs<T, L>() * tps<L>())) void k_yb(Geom g, T* __restrict__ spec,
                                                                  const cpx<T>* __restrict__ twL,
                                                                  int ntiles) {
  constexpr int TPS = tps<L>(), Q = y_seqs<T, L>(), NT = Q * TPS, W = 2 * Q;
  __shared__ cpx<T> buf[padded(L * Q)];
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = (int)(b % (unsigned)ntiles), k = (int)(b / (unsigned)ntiles);
  T* plane = spec + (long)k * g.nx * g.ny + (long)tile * W;
  // Global row j, column pair q holds (perm_a[j], perm_b[j]) of columns 2q and
  // 2q+1.  Rows 2m and 2m+1 (m = 1..L/2-1) give the complex inputs Z[m] and
  // Z[L-m] of the pair's transform: with A = (e0.re, e1.re), B = (e0.im,
  // e1.im), Z = A + iB below L/2 and conj(A) + i conj(B) above; rows 0 and 1
  // hold Z[0] and Z[L/2].  Each thread loads its row pairs and writes the
  // unpacked values to LDS directly.
  {
    constexpr int NE = (L / 2) * Q, NLD = (NE + NT - 1) / NT;
    cpx<T> e0[NLD], e1[NLD];
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int e = threadIdx.x + u * NT;
      if (NE % NT == 0 || e < NE) {
        const int mm = e / Q, q2 = e % Q;
        e0[u] = reinterpret_cast<const cpx<T>*>(plane + (long)(2 * mm) * g.nx)[q2];
        e1[u] = reinterpret_cast<const cpx<T>*>(plane + (long)(2 * mm + 1) * g.nx)[q2];
      }
    }
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int e = threadIdx.x + u * NT;
      if (NE % NT == 0 || e < NE) {
        const int mm = e / Q, q2 = e % Q;
        if (mm == 0) {
          buf[pad(q2)] = e0[u];
          buf[pad((L / 2) * Q + q2)] = e1[u];
        } else {
          buf[pad(mm * Q + q2)] = cpx<T>{e0[u].re - e1[u].im, e1[u].re + e0[u].im};
          buf[pad((L - mm) * Q + q2)] = cpx<T>{e0[u].re + e1[u].im, e0[u].im - e1[u].re};
        }
      }
    }
  }
  __syncthreads();
  const int q = threadIdx.x % Q, jj = threadIdx.x / Q;
  stockham<T, +1, L, Q>(buf, twL, q, jj);
  store_tile<T, L * Q, NT, Q>(buf, plane, g.nx);
}

// ---- host side -------------------------------------------------------------------
bool build_plan(int L, FftPlan& pl, std::vector<std::complex<long double>>& tw) {
  pl.L = L;
  pl.npass = 0;
  int m = L;
  std::vector<int> rad;
  while (m % 4 == 0) { rad.push_back(4); m /= 4; }
  while (m % 2 == 0) { rad.push_back(2); m /= 2; }
  while (m % 3 == 0) { rad.push_back(3); m /= 3; }
  while (m % 5 == 0) { rad.push_back(5); m /= 5; }
  if (m != 1 || (int)rad.size() > kMaxPass) return false;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  int ns = 1;
  for (int R : rad) {
    const int p = pl.npass++;
    pl.radix[p] = R;
    pl.ns[p] = ns;
    pl.twoff[p] = (int)tw.size();
    for (int k = 0; k < ns; ++k)
      for (int r = 1; r < R; ++r) {
        const long t = ((long)r * k) % ((long)ns * R);
        const long double ang = -two_pi * (long double)t / (long double)(ns * R);
        tw.emplace_back(cosl(ang), sinl(ang));
      }
    ns *= R;
  }
  if (tw.empty()) tw.emplace_back(1.0L, 0.0L);
  return true;
}

template <typename T>
int upload_table(pll_context* c, const std::vector<std::complex<long double>>& src, void** dst) {
  std::vector<cpx<T>> h(src.size());
  for (size_t t = 0; t < src.size(); ++t) h[t] = cpx<T>{(T)src[t].real(), (T)src[t].imag()};
  PLL_HIP_TRY(ctx_malloc(c, dst, h.size() * sizeof(cpx<T>)));
  PLL_HIP_TRY(hipMemcpy(*dst, h.data(), h.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
  return PLL_OK;
}

constexpr size_t kLdsBudget = 64 * 1024;  // bytes per block for the two ping-pong buffers

template <typename T>
int setup_t(pll_context* c) {
  const int nx = c->g.nx, ny = c->g.ny;
  std::vector<std::complex<long double>> twx, twy, rtw;
  if (!build_plan(nx / 2, c->plan_x, twx)) return fail(PLL_ERR_ARG, "nx/2 must factor into 2, 3, 5");
  if (!build_plan(ny, c->plan_y, twy)) return fail(PLL_ERR_ARG, "ny must factor into 2, 3, 5");
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k <= nx / 2; ++k) {
    const long double ang = -two_pi * (long double)k / (long double)nx;
    rtw.emplace_back(cosl(ang), sinl(ang));
  }
  int rc;
  if ((rc = upload_table<T>(c, twx, &c->d_tw_x))) return rc;
  if ((rc = upload_table<T>(c, twy, &c->d_tw_y))) return rc;
  if ((rc = upload_table<T>(c, rtw, &c->d_rtw_x))) return rc;
  const size_t rowbytes = (size_t)nx * sizeof(T) * 2;  // two buffers
  long batch = (long)(kLdsBudget / rowbytes);
  if (batch < 1) return fail(PLL_ERR_ARG, "nx too large for one LDS row transform");
  if (batch > 64) batch = 64;
  c->x_batch = (int)batch;
  const size_t colbytes = (size_t)ny * sizeof(T) * 2;
  int tile = 2;
  while (tile * 2 <= 64 && (size_t)(tile * 2) * colbytes <= kLdsBudget && nx % (tile * 2) == 0)
    tile *= 2;
  if ((size_t)tile * colbytes > kLdsBudget) return fail(PLL_ERR_ARG, "ny too large for one LDS column transform");
  c->y_tile = tile;
  return PLL_OK;
}

template <typename T>
int x_fwd_generic(pll_context* c) {
  const long nrows = (long)c->g.ny * c->g.nz;
  const unsigned blocks = (unsigned)((nrows + c->x_batch - 1) / c->x_batch);
  const size_t sh = (size_t)c->x_batch * c->g.nx * sizeof(T) * 2;
  TimerScope ts(c, K_FFT_X_FWD);
  hipLaunchKernelGGL((k_fft_x_fwd<T>), dim3(blocks), dim3(kFftThreads), sh, c->stream, c->g,
                     static_cast<const T*>(c->fld[PLL_P]), static_cast<T*>(c->spec), c->plan_x,
                     static_cast<const cpx<T>*>(c->d_tw_x), static_cast<const cpx<T>*>(c->d_rtw_x),
                     c->x_batch);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int x_bwd_generic(pll_context* c) {
  const long nrows = (long)c->g.ny * c->g.nz;
  const unsigned blocks = (unsigned)((nrows + c->x_batch - 1) / c->x_batch);
  const size_t sh = (size_t)c->x_batch * c->g.nx * sizeof(T) * 2;
  const double norm = 1.0 / (c->g.nx * c->g.ny);  // src/mod_fft.cc:57
  TimerScope ts(c, K_FFT_X_BWD);
  hipLaunchKernelGGL((k_fft_x_bwd<T>), dim3(blocks), dim3(kFftThreads), sh, c->stream, c->g,
                     static_cast<const T*>(c->spec), static_cast<T*>(c->fld[PLL_P]), c->plan_x,
                     static_cast<const cpx<T>*>(c->d_tw_x), static_cast<const cpx<T>*>(c->d_rtw_x),
                     c->x_batch, (T)norm);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int y_generic(pll_context* c, bool fwd) {
  const dim3 grid(c->g.nx / c->y_tile, c->g.nz);
  const size_t sh = (size_t)c->y_tile * c->g.ny * sizeof(T) * 2;
  if (fwd) {
    TimerScope ts(c, K_FFT_Y_FWD);
    hipLaunchKernelGGL((k_fft_y_fwd<T>), grid, dim3(kFftThreads), sh, c->stream, c->g,
                       static_cast<T*>(c->spec), c->plan_y, static_cast<const cpx<T>*>(c->d_tw_y),
                       c->y_tile);
  } else {
    TimerScope ts(c, K_FFT_Y_BWD);
    hipLaunchKernelGGL((k_fft_y_bwd<T>), grid, dim3(kFftThreads), sh, c->stream, c->g,
                       static_cast<T*>(c->spec), c->plan_y, static_cast<const cpx<T>*>(c->d_tw_y),
                       c->y_tile);
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

// ---- fixed-length kernels ------------------------------------------------------------
template <typename T, int L>
int x_fixed(pll_context* c, bool fwd, bool rhs = false) {
  constexpr int NB = x_rows<T, L>(), NT = NB * tps<L>();
  const long nrows = (long)c->g.ny * c->g.nz;
  const unsigned blocks = (unsigned)((nrows + NB - 1) / NB);
  const cpx<T>* tw = static_cast<const cpx<T>*>(c->d_twl_x);
  const cpx<T>* rtw = static_cast<const cpx<T>*>(c->d_rtw_x);
  if (fwd && rhs) {
    const double dti = 1.0 / c->dt;  // src/mod_calcRHS.cc:11-13
    RhsSrc<T> rs{static_cast<const T*>(c->fld[PLL_U]), static_cast<const T*>(c->fld[PLL_V]),
                 static_cast<const T*>(c->fld[PLL_W]),
                 static_cast<const T*>(c->d_mesh) + (c->g.nz + 2),  // dzf_inv
                 (T)dti, (T)(dti * c->dx_inv), (T)(dti * c->dy_inv), c->rhs_bcfold};
    TimerScope ts(c, K_FFT_X_FWD_RHS);
    hipLaunchKernelGGL((k_xf<T, L, true>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                       static_cast<const T*>(nullptr), static_cast<T*>(c->spec), tw, rtw, nrows, rs);
  } else if (fwd) {
    TimerScope ts(c, K_FFT_X_FWD);
    hipLaunchKernelGGL((k_xf<T, L, false>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                       static_cast<const T*>(c->fld[PLL_P]), static_cast<T*>(c->spec), tw, rtw,
                       nrows, RhsSrc<T>{});
  } else {
    const double norm = 1.0 / (c->g.nx * c->g.ny);  // src/mod_fft.cc:57
    TimerScope ts(c, K_FFT_X_BWD);
    if (rhs) {  // inside the step: updateBoundP follows, x halos may be written
      hipLaunchKernelGGL((k_xb<T, L, true>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                         static_cast<const T*>(c->spec), static_cast<T*>(c->fld[PLL_P]), tw, rtw,
                         nrows, (T)norm);
    } else {
      hipLaunchKernelGGL((k_xb<T, L, false>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                         static_cast<const T*>(c->spec), static_cast<T*>(c->fld[PLL_P]), tw, rtw,
                         nrows, (T)norm);
    }
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T, int L>
bool y_fixed_fits(const pll_context* c) {
  return c->g.nx % (2 * y_seqs<T, L>()) == 0;
}

template <typename T, int L>
int y_fixed(pll_context* c, bool fwd) {
  constexpr int Q = y_seqs<T, L>(), NT = Q * tps<L>();
  const int ntiles = c->g.nx / (2 * Q);
  const unsigned blocks = (unsigned)ntiles * c->g.nz;
  const cpx<T>* tw = static_cast<const cpx<T>*>(c->d_twl_y);
  if (fwd) {
    TimerScope ts(c, K_FFT_Y_FWD);
    hipLaunchKernelGGL((k_yf<T, L>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                       static_cast<T*>(c->spec), tw, ntiles);
  } else {
    TimerScope ts(c, K_FFT_Y_BWD);
    hipLaunchKernelGGL((k_yb<T, L>), dim3(blocks), dim3(NT), 0, c->stream, c->g,
                       static_cast<T*>(c->spec), tw, ntiles);
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

#define PLL_X_SIZES(X) X(4) X(12) X(32) X(48) X(128) X(192) X(384) X(768)
#define PLL_Y_SIZES(X) X(6) X(20) X(48) X(64) X(192) X(256) X(384) X(768)

template <typename T>
bool x_has_fixed(int L) {
#define PLL_CASE(n) if (L == n) return true;
  PLL_X_SIZES(PLL_CASE)
#undef PLL_CASE
  return false;
}

template <typename T>
bool y_has_fixed(const pll_context* c, int L) {
#define PLL_CASE(n) if (L == n) return y_fixed_fits<T, n>(c);
  PLL_Y_SIZES(PLL_CASE)
#undef PLL_CASE
  return false;
}

template <typename T>
int x_t(pll_context* c, bool fwd) {
  if (c->x_fixed) {
    switch (c->g.nx / 2) {
#define PLL_CASE(n) case n: return x_fixed<T, n>(c, fwd);
      PLL_X_SIZES(PLL_CASE)
#undef PLL_CASE
      default: break;
    }
  }
  return fwd ? x_fwd_generic<T>(c) : x_bwd_generic<T>(c);
}

template <typename T>
int y_t(pll_context* c, bool fwd) {
  if (c->y_fixed) {
    switch (c->g.ny) {
#define PLL_CASE(n) case n: return y_fixed<T, n>(c, fwd);
      PLL_Y_SIZES(PLL_CASE)
#undef PLL_CASE
      default: break;
    }
  }
  return y_generic<T>(c, fwd);
}

// W_L^t = exp(-2 pi i t / L), t = 0..L-1, for the fixed-length kernels.
template <typename T>
int unit_table(pll_context* c, int L, void** dst) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  std::vector<std::complex<long double>> t;
  for (int k = 0; k < L; ++k) {
    const long double ang = -two_pi * (long double)k / (long double)L;
    t.emplace_back(cosl(ang), sinl(ang));
  }
  return upload_table<T>(c, t, dst);
}

template <typename T>
int setup_fixed(pll_context* c) {
  const char* env = std::getenv("PLL_FFT_GENERIC");
  const bool force_generic = env && env[0] == '1';
  c->x_fixed = !force_generic && x_has_fixed<T>(c->g.nx / 2);
  c->y_fixed = !force_generic && y_has_fixed<T>(c, c->g.ny);
  int rc;
  if ((rc = unit_table<T>(c, c->g.nx / 2, &c->d_twl_x))) return rc;
  if ((rc = unit_table<T>(c, c->g.ny, &c->d_twl_y))) return rc;
  return PLL_OK;
}

}  // namespace

int fft_setup(pll_context* c) {
  int rc = c->prec == PLL_FP64 ? setup_t<double>(c) : setup_t<float>(c);
  if (rc) return rc;
  return c->prec == PLL_FP64 ? setup_fixed<double>(c) : setup_fixed<float>(c);
}
int launch_fft_x_fwd(pll_context* c) { return c->prec == PLL_FP64 ? x_t<double>(c, true) : x_t<float>(c, true); }

template <typename T>
int x_rhs_t(pll_context* c) {
  switch (c->g.nx / 2) {
#define PLL_CASE(n) case n: return x_fixed<T, n>(c, true, true);
    PLL_X_SIZES(PLL_CASE)
#undef PLL_CASE
    default: break;
  }
  return fail(PLL_ERR_UNSUPPORTED, "no fixed-length x transform for calcRHS fusion");
}

bool fft_x_rhs_available(const pll_context* c) { return c->x_fixed; }

template <typename T>
int x_bwd_halo_t(pll_context* c) {
  switch (c->g.nx / 2) {
#define PLL_CASE(n) case n: return x_fixed<T, n>(c, false, true);
    PLL_X_SIZES(PLL_CASE)
#undef PLL_CASE
    default: break;
  }
  return x_bwd_generic<T>(c);
}

int launch_fft_x_bwd_step(pll_context* c) {
  if (!c->x_fixed) return launch_fft_x_bwd(c);
  return c->prec == PLL_FP64 ? x_bwd_halo_t<double>(c) : x_bwd_halo_t<float>(c);
}

int launch_fft_x_fwd_rhs(pll_context* c) {
  if (!c->x_fixed) return fail(PLL_ERR_UNSUPPORTED, "no fixed-length x transform for calcRHS fusion");
  return c->prec == PLL_FP64 ? x_rhs_t<double>(c) : x_rhs_t<float>(c);
}
int launch_fft_x_bwd(pll_context* c) { return c->prec == PLL_FP64 ? x_t<double>(c, false) : x_t<float>(c, false); }

int launch_poisson_step(pll_context* c, bool rhs) {
  int rc;
  if ((rc = rhs ? launch_fft_x_fwd_rhs(c) : launch_fft_x_fwd(c))) return rc;
  if ((rc = launch_fft_y_fwd(c))) return rc;
  if ((rc = launch_trid_step(c))) return rc;
  if ((rc = launch_fft_y_bwd(c))) return rc;
  return launch_fft_x_bwd_step(c);
}
int launch_fft_y_fwd(pll_context* c) { return c->prec == PLL_FP64 ? y_t<double>(c, true) : y_t<float>(c, true); }
int launch_fft_y_bwd(pll_context* c) { return c->prec == PLL_FP64 ? y_t<double>(c, false) : y_t<float>(c, false); }

}  // namespace pll
