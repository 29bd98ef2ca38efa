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
#include "pll_internal.hpp"

#include <cmath>
#include <complex>

namespace pll {

namespace {

template <typename T>
struct cpx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re + b.re, a.im + b.im};
}

template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re - b.re, a.im - b.im};
}

// s * i * a for s = +1 / -1
template <typename T, int S>
__device__ __forceinline__ cpx<T> rot(cpx<T> a) {
  return S > 0 ? cpx<T>{-a.im, a.re} : cpx<T>{a.im, -a.re};
}

// In-register DFT of R points, X_q = sum_r v_r exp(S*2*pi*i*r*q/R), S = -1
// forward, +1 inverse.  Constants are the exact trigonometric values.
template <typename T, int R, int S>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if (R == 2) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    const cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const cpx<T> s13 = cadd(v[1], v[3]), d13 = rot<T, S>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  } else if (R == 3) {
    const T h3 = T(0.86602540378443864676372317075294);  // sin(pi/3)
    const cpx<T> s = cadd(v[1], v[2]);
    const cpx<T> d = csub(v[1], v[2]);
    const cpx<T> m{v[0].re - T(0.5) * s.re, v[0].im - T(0.5) * s.im};
    const cpx<T> r = rot<T, S>(cpx<T>{h3 * d.re, h3 * d.im});
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if (R == 5) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
    const cpx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const cpx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const cpx<T> m1{v[0].re + c1 * a1.re + c2 * a2.re, v[0].im + c1 * a1.im + c2 * a2.im};
    const cpx<T> m2{v[0].re + c2 * a1.re + c1 * a2.re, v[0].im + c2 * a1.im + c1 * a2.im};
    const cpx<T> n1 = rot<T, S>(cpx<T>{s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im});
    const cpx<T> n2 = rot<T, S>(cpx<T>{s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im});
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

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
int upload_table(const std::vector<std::complex<long double>>& src, void** dst) {
  std::vector<cpx<T>> h(src.size());
  for (size_t t = 0; t < src.size(); ++t) h[t] = cpx<T>{(T)src[t].real(), (T)src[t].imag()};
  PLL_HIP_TRY(hipMalloc(dst, h.size() * sizeof(cpx<T>)));
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
  if ((rc = upload_table<T>(twx, &c->d_tw_x))) return rc;
  if ((rc = upload_table<T>(twy, &c->d_tw_y))) return rc;
  if ((rc = upload_table<T>(rtw, &c->d_rtw_x))) return rc;
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
int x_fwd_t(pll_context* c) {
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
int x_bwd_t(pll_context* c) {
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
int y_t(pll_context* c, bool fwd) {
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

}  // namespace

int fft_setup(pll_context* c) { return c->prec == PLL_FP64 ? setup_t<double>(c) : setup_t<float>(c); }
int launch_fft_x_fwd(pll_context* c) { return c->prec == PLL_FP64 ? x_fwd_t<double>(c) : x_fwd_t<float>(c); }
int launch_fft_x_bwd(pll_context* c) { return c->prec == PLL_FP64 ? x_bwd_t<double>(c) : x_bwd_t<float>(c); }
int launch_fft_y_fwd(pll_context* c) { return c->prec == PLL_FP64 ? y_t<double>(c, true) : y_t<float>(c, true); }
int launch_fft_y_bwd(pll_context* c) { return c->prec == PLL_FP64 ? y_t<double>(c, false) : y_t<float>(c, false); }

}  // namespace pll
