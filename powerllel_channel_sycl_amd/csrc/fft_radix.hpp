// fft_radix.hpp -- in-register DFT butterflies used by the LDS transforms.
//
// dft<T, R, S>(v) overwrites v[0..R) with X_q = sum_r v_r exp(S*2*pi*i*r*q/R),
// S = -1 forward, +1 inverse, for R in {2, 3, 4, 5, 6, 8, 12, 16}.  Radices 2-5
// are closed-form butterflies; the composite ones are a Cooley-Tukey split
// R = A*B evaluated entirely in registers with constant twiddles.  All
// constants are exact trigonometric values (rounded from 21 digits).
// The functions are __host__ __device__ so the CPU self-test can run them.
#pragma once

#include <hip/hip_runtime.h>

namespace pll {

template <typename T>
struct cpx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.re - b.re, a.im - b.im};
}
// S*i*a
template <typename T, int S>
__host__ __device__ __forceinline__ cpx<T> rot(cpx<T> a) {
  return S > 0 ? cpx<T>{-a.im, a.re} : cpx<T>{a.im, -a.re};
}

// cos and sin of 2*pi*t/48, t = 0..11.
__host__ __device__ __forceinline__ double cos48(int t) {
  constexpr double c[12] = {1.0,
                            0.991444861373810411145,
                            0.965925826289068286750,
                            0.923879532511286756128,
                            0.866025403784438646764,
                            0.793353340291235164580,
                            0.707106781186547524401,
                            0.608761429008720639416,
                            0.5,
                            0.382683432365089771728,
                            0.258819045102520762349,
                            0.130526192220051591548};
  return c[t];
}

// exp(S*2*pi*i*m/R) for R dividing 48.
template <typename T, int R, int S>
__host__ __device__ __forceinline__ cpx<T> unit_root(int m) {
  const int t = ((m * (48 / R)) % 48 + 48) % 48;
  const int q = t / 12, r = t % 12;
  const double c = cos48(r), s = cos48((12 - r) % 12) * (r == 0 ? 0.0 : 1.0);
  double cr, sr;
  switch (q) {
    case 0: cr = c; sr = s; break;
    case 1: cr = -s; sr = c; break;
    case 2: cr = -c; sr = -s; break;
    default: cr = s; sr = -c; break;
  }
  return cpx<T>{(T)cr, (T)(S * sr)};
}

template <typename T, int R, int S>
__host__ __device__ __forceinline__ void dft(cpx<T>* v);

template <typename T, int S>
__host__ __device__ __forceinline__ void dft2(cpx<T>* v) {
  const cpx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, int S>
__host__ __device__ __forceinline__ void dft3(cpx<T>* v) {
  const T h3 = T(0.866025403784438646764);  // sin(pi/3)
  const cpx<T> s = cadd(v[1], v[2]);
  const cpx<T> d = csub(v[1], v[2]);
  const cpx<T> m{v[0].re - T(0.5) * s.re, v[0].im - T(0.5) * s.im};
  const cpx<T> r = rot<T, S>(cpx<T>{h3 * d.re, h3 * d.im});
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <typename T, int S>
__host__ __device__ __forceinline__ void dft4(cpx<T>* v) {
  const cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  const cpx<T> s13 = cadd(v[1], v[3]), d13 = rot<T, S>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

template <typename T, int S>
__host__ __device__ __forceinline__ void dft5(cpx<T>* v) {
  const T c1 = T(0.309016994374947424102);   // cos(2pi/5)
  const T c2 = T(-0.809016994374947424102);  // cos(4pi/5)
  const T s1 = T(0.951056516295153572116);   // sin(2pi/5)
  const T s2 = T(0.587785252292473129169);   // sin(4pi/5)
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

// R = A*B: X[k1 + A*k2] = sum_n2 w_R^(n2 k1) w_B^(n2 k2) sum_n1 x[B n1 + n2] w_A^(n1 k1).
template <typename T, int A, int B, int S>
__host__ __device__ __forceinline__ void dft_split(cpx<T>* v) {
  constexpr int R = A * B;
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    cpx<T> a[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
    dft<T, A, S>(a);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      cpx<T> y = a[k1];
      if ((n2 * k1) % R != 0) y = cmul(y, unit_root<T, R, S>(n2 * k1));
      v[B * k1 + n2] = y;
    }
  }
  cpx<T> out[R];
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cpx<T> b[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) b[n2] = v[B * k1 + n2];
    dft<T, B, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) out[k1 + A * k2] = b[k2];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = out[r];
}

template <typename T, int R, int S>
__host__ __device__ __forceinline__ void dft(cpx<T>* v) {
  static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 12 || R == 16,
                "unsupported radix");
  if constexpr (R == 2) dft2<T, S>(v);
  else if constexpr (R == 3) dft3<T, S>(v);
  else if constexpr (R == 4) dft4<T, S>(v);
  else if constexpr (R == 5) dft5<T, S>(v);
  else if constexpr (R == 6) dft_split<T, 2, 3, S>(v);
  else if constexpr (R == 8) dft_split<T, 2, 4, S>(v);
  else if constexpr (R == 12) dft_split<T, 4, 3, S>(v);
  else dft_split<T, 4, 4, S>(v);
}

}  // namespace pll
