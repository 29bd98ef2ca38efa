"""The allocation-free closure gradient (pll_closure_gradient_noalloc, no
counterpart in the reference) against pll_closure_gradient: the same kernels
on a static scratch array instead of the closure's own buffers, so the same
bits.  Two calls in a row, the second with a smaller sample count."""
import ctypes as C

import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll


@pytest.mark.gpu
def test_gradient_noalloc_matches_gradient():
    import torch
    lib = pll.load_library()
    cl = pll.Closure(seed=11)
    cl.set_scales(2.0, 0.5, 1e-4)
    gen = torch.Generator(device="cuda:0").manual_seed(5)
    try:
        for n in (300_001, 41_257):
            print(f"[noalloc] n = {n}", flush=True)
            x = torch.randn((n, 10), generator=gen, device="cuda:0", dtype=torch.float32)
            y = torch.rand((n,), generator=gen, device="cuda:0", dtype=torch.float32)
            torch.cuda.synchronize()
            g_ref, l_ref = cl.gradient(x.data_ptr(), y.data_ptr(), n)
            g = np.zeros(cl.nparam, dtype=np.float32)
            loss = C.c_double()
            rc = lib.pll_closure_gradient_noalloc(cl._h, C.c_void_p(x.data_ptr()),
                                                  C.c_void_p(y.data_ptr()), n,
                                                  g.ctypes.data_as(C.POINTER(C.c_float)),
                                                  C.byref(loss))
            assert rc == 0, lib.pll_last_error()
            np.testing.assert_array_equal(g, g_ref)
            assert loss.value == l_ref
            assert np.isfinite(l_ref) and np.any(g_ref != 0)
    finally:
        cl.close()
