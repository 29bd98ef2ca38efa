"""The production grids on the GPU, checked through size-independent properties
(the CPU oracle is far too slow at these sizes): Re_tau~550 768x384x384 and
Re_tau~1000 1536x768x512 (the single-GPU memory-sizing case), both precisions.

After a few steps from the seeded log profile + noise: no NaN anywhere, the
projection leaves max|div u| at the precision's round-off, the forcing holds the
bulk velocity at vel_force (the init's Mean(nz, ny) quirk leaves it at nz/nx),
and the FP32-mixed velocity stays within FP32 round-off of the FP64 one.
"""
import math

import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll

pytestmark = pytest.mark.gpu

GRIDS = {
    # name: (nx, ny, nz, stretch, re, dt)
    "retau550_768x384x384": (768, 384, 384, 1.8, 10000.0, 5e-4),
    "retau1000_1536x768x512": (1536, 768, 512, 2.0, 20000.0, 2.5e-4),
}


def _channel(name, precision):
    nx, ny, nz, r, re, dt = GRIDS[name]
    dzf = pll.mesh_by_func(1, 0, r, nz, 2.0)
    ch = pll.Channel(nx, ny, nz, 2 * math.pi, math.pi, 2.0, re=re, dt=dt, dzf_global=dzf,
                     precision=precision)
    ch.initFlow("log", noise_intensity=0.1)
    ch.updateBoundVel(0)
    ch.updateBoundP()
    return ch, np.min(dzf[1:-1])


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("name", list(GRIDS))
def test_fullsize_invariants(name, precision):
    print(f"[fullsize] {name} {precision}", flush=True)
    nx, ny, nz = GRIDS[name][:3]
    ch, dzmin = _channel(name, precision)
    with ch:
        assert abs(ch.Mean("u") - nz / nx) < 1e-6  # src/mod_initFlow.cc:92 normalises with Mean(nz, ny)
        ch.step(3)
        for f in ("u", "v", "w", "p"):
            assert ch.checkNaN(f), f
        div, _ = ch.calcMaxDiv()
        cfl, _ = ch.calcMaxCFL()
        # |u| ~ 1, dz_min ~ 3e-3: measured ~1e-13 (FP64) and ~5e-5 (FP32)
        bound = 1e-12 if precision == "fp64" else 2e-4
        print(f"  max div {div:.3e} (bound {bound:.1e}), max cfl {cfl:.3e}", flush=True)
        assert div < bound
        assert 0 < cfl < 1
        tol = 1e-12 if precision == "fp64" else 1e-6
        assert abs(ch.Mean("u") - 1.0) < tol


def test_fullsize_fp32_tracks_fp64():
    name = "retau550_768x384x384"
    print(f"[fullsize] fp32 vs fp64 {name}", flush=True)
    out = {}
    for precision in ("fp64", "fp32_mixed"):
        ch, _ = _channel(name, precision)
        with ch:
            ch.step(5)
            out[precision] = [ch.download(f)[1:-1, 1:-1, 1:-1] for f in ("u", "v", "w")]
    umax = np.max(np.abs(out["fp64"][0]))
    for f, a, b in zip("uvw", out["fp32_mixed"], out["fp64"]):
        err = np.max(np.abs(a - b)) / umax
        print(f"  {f}: max |fp32 - fp64| / max|u| = {err:.2e}", flush=True)
        assert err < 2e-5, f
