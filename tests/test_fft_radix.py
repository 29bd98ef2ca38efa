"""The in-register DFT butterflies of the HIP transforms (fft_radix.hpp), run on
the CPU through libpll_selftest.so and checked against the DFT definition in
extended precision."""
import ctypes as C
import os

import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll

_LIB = os.path.join(os.path.dirname(pll.LIB_PATH), "libpll_selftest.so")


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(_LIB)
    D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.pll_selftest_dft.argtypes = [C.c_int, C.c_int, D, D]
    L.pll_selftest_dft.restype = C.c_int
    return L


@pytest.mark.parametrize("radix", [2, 3, 4, 5, 6, 8, 12, 16])
@pytest.mark.parametrize("sign", [-1, 1])
def test_butterfly_matches_dft_definition(lib, radix, sign):
    rng = np.random.default_rng(radix * 7 + sign)
    z = rng.standard_normal(radix) + 1j * rng.standard_normal(radix)
    inp = np.empty(2 * radix)
    inp[0::2], inp[1::2] = z.real, z.imag
    out = np.empty(2 * radix)
    assert lib.pll_selftest_dft(radix, sign, inp, out) == 0
    got = out[0::2] + 1j * out[1::2]
    n = np.arange(radix, dtype=np.longdouble)
    ang = np.mod(np.outer(n, n), radix) * (sign * 2 * np.pi / np.longdouble(radix))
    zr, zi = z.real.astype(np.longdouble), z.imag.astype(np.longdouble)
    ref = ((np.cos(ang) * zr - np.sin(ang) * zi).sum(1).astype(float)
           + 1j * (np.sin(ang) * zr + np.cos(ang) * zi).sum(1).astype(float))
    assert np.max(np.abs(got - ref)) < 4e-15 * np.max(np.abs(ref))
