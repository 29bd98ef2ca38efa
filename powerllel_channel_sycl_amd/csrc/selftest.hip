// selftest.hip -- host-side entry points that run the in-register DFT
// butterflies of fft_radix.hpp on the CPU, for the CPU test suite.  Built as
// libpll_selftest.so, separate from the product library.
#include "fft_radix.hpp"

namespace {
template <int R, int S>
void run(const double* in, double* out) {
  pll::cpx<double> v[R];
  for (int r = 0; r < R; ++r) v[r] = pll::cpx<double>{in[2 * r], in[2 * r + 1]};
  pll::dft<double, R, S>(v);
  for (int r = 0; r < R; ++r) {
    out[2 * r] = v[r].re;
    out[2 * r + 1] = v[r].im;
  }
}
template <int R>
int run_sign(int sign, const double* in, double* out) {
  if (sign < 0) run<R, -1>(in, out);
  else run<R, 1>(in, out);
  return 0;
}
}  // namespace

extern "C" int pll_selftest_dft(int radix, int sign, const double* in, double* out) {
  switch (radix) {
    case 2: return run_sign<2>(sign, in, out);
    case 3: return run_sign<3>(sign, in, out);
    case 4: return run_sign<4>(sign, in, out);
    case 5: return run_sign<5>(sign, in, out);
    case 6: return run_sign<6>(sign, in, out);
    case 8: return run_sign<8>(sign, in, out);
    case 12: return run_sign<12>(sign, in, out);
    case 16: return run_sign<16>(sign, in, out);
    default: return 1;
  }
}
