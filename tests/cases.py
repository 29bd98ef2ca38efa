"""Problem definitions shared by the CPU and GPU tests.

C1 is the reference's CPU-runnable channel (64^3, 2pi x pi x 2, tanh stretch
1.5, Re_b = 5600 -> re = 2800, dt = 2e-3).  The small cases exercise every FFT
radix (2, 3, 4, 5), non-cubic and ragged shapes.
"""
import math

import numpy as np

CASES = {
    # name: (nx, ny, nz, lx, ly, lz, stretch, re, dt)
    "c1": (64, 64, 64, 2 * math.pi, math.pi, 2.0, 1.5, 2800.0, 2e-3),
    "r235": (24, 20, 12, 2 * math.pi, math.pi, 2.0, 1.5, 2800.0, 2e-3),
    "r3": (96, 48, 32, 2 * math.pi, math.pi, 2.0, 1.8, 10000.0, 5e-4),
    "tiny": (8, 6, 4, 1.3, 0.7, 2.0, 1.5, 100.0, 1e-3),
    # sizes without a fixed-length transform kernel (generic Stockham path)
    "gen": (40, 30, 16, 2 * math.pi, math.pi, 2.0, 1.5, 2800.0, 2e-3),
    # the Re_tau~180 grid and thin slabs of the Re_tau~550 / ~1000 grids: the
    # production transform lengths (x: 128, 384, 768 points; y: 192, 384, 768)
    "c2": (256, 192, 128, 4 * math.pi, 2 * math.pi, 2.0, 1.5, 2800.0, 1e-3),
    "c3slab": (768, 384, 8, 2 * math.pi, math.pi, 2.0, 1.8, 10000.0, 5e-4),
    "c4slab": (1536, 768, 4, 2 * math.pi, math.pi, 2.0, 2.0, 20000.0, 2.5e-4),
    # a tall column (nz = 300)
    "tall": (16, 12, 300, 2 * math.pi, math.pi, 2.0, 1.8, 2800.0, 2e-3),
    # columns deeper than 384 planes: the step's partitioned solve with 32
    # segments of 16 (nz = 512) and of 24 planes (nz = 700)
    "deep512": (16, 12, 512, 2 * math.pi, math.pi, 2.0, 1.8, 2800.0, 2e-3),
    "deep700": (16, 12, 700, 2 * math.pi, math.pi, 2.0, 1.8, 2800.0, 2e-3),
    # the Re_tau~550 grid of the benchmark, whole
    "c3": (768, 384, 384, 2 * math.pi, math.pi, 2.0, 1.8, 10000.0, 5e-4),
    # the Re_tau~1000 grid, whole (604 M points)
    "c4": (1536, 768, 512, 2 * math.pi, math.pi, 2.0, 2.0, 20000.0, 2.5e-4),
    # no dimension a multiple of the closure stage's 32 x 8 x 4 blocks
    "ragged": (40, 30, 10, 2 * math.pi, math.pi, 2.0, 1.5, 2800.0, 2e-3),
    # ragged 64 x 16 tiles (nx = 3*64 + 8, ny = 5*16 + 10) with enough blocks
    # (1056) for the fused FP64 stage 2 to take its 16-row tiles
    "ragtile": (200, 90, 176, 2 * math.pi, math.pi, 2.0, 1.5, 2800.0, 2e-3),
}


def params(name):
    nx, ny, nz, lx, ly, lz, r, re, dt = CASES[name]
    return dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, stretch=r, re=re, dt=dt)


def dzf_of(name):
    from oracle import pll_oracle as po
    p = params(name)
    return po.mesh_by_func(1, 0, p["stretch"], p["nz"], p["lz"])


def make_oracle(name, **kw):
    from oracle import pll_oracle as po
    p = params(name)
    return po.Oracle(p["nx"], p["ny"], p["nz"], p["lx"], p["ly"], p["lz"], p["re"], p["dt"],
                     dzf_of(name), **kw)


def make_channel(name, precision="fp64", **kw):
    import powerllel_channel_sycl_amd as pll
    p = params(name)
    return pll.Channel(p["nx"], p["ny"], p["nz"], p["lx"], p["ly"], p["lz"], p["re"], p["dt"],
                       dzf_of(name), precision=precision, **kw)


def random_field(name, seed, scale=1.0):
    p = params(name)
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal((p["nz"] + 2, p["ny"] + 2, p["nx"] + 2))


def rel_err(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def interior(f):
    return f[1:-1, 1:-1, 1:-1]
