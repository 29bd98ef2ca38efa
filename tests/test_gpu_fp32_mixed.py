"""FP32-mixed path (FP32 fields, stencils and FFTs; FP64 bulk reduction and
tridiagonal arithmetic) against the FP64 oracle on the same seeded state.

Pointwise error after a short horizon, and statistical error (mean velocity
profile, Reynolds shear stress) after a longer one.  Tolerances are stated in
each test; FP32 round-off is ~6e-8 per operation.
"""
import numpy as np
import pytest

from cases import dzf_of, interior, make_channel, make_oracle, random_field, rel_err

pytestmark = pytest.mark.gpu

VEL = ("u", "v", "w")


def _seeded(name):
    o = make_oracle(name)
    f = [o.field() for _ in range(7)]
    o.init_flow(f[0], f[1], f[2], "poi", noise_intensity=0.1)
    o.update_bound_vel(f[0], f[1], f[2])
    o.update_bound_p(f[6])
    return o, f


def _profiles(u, v, w):
    ui, vi, wi = interior(u), interior(v), interior(w)
    um = ui.mean(axis=(1, 2))
    uvm = ((ui - um[:, None, None]) * (0.5 * (wi + np.concatenate([wi[:1] * 0, wi[:-1]])))).mean(
        axis=(1, 2))
    return um, uvm


@pytest.mark.parametrize("name", ["r3", "c1"])
def test_fp32_stages_close_to_fp64(name):
    print("CASE fp32 stages", name, flush=True)
    o = make_oracle(name)
    with make_channel(name, precision="fp32_mixed") as ch:
        u, v, w = (random_field(name, 300 + d) for d in range(3))
        o.update_bound_vel(u, v, w)
        u, v, w = (a.astype(np.float32).astype(np.float64) for a in (u, v, w))
        for fld, a in zip(VEL, (u, v, w)):
            ch.upload(fld, a)
        u1, v1, w1 = (o.field() for _ in range(3))
        ch.timeIntVelRK1()
        o.rk1(u, v, w, u1, v1, w1)
        for fld, a in zip(("u1", "v1", "w1"), (u1, v1, w1)):
            assert rel_err(interior(ch.download(fld)), interior(a)) < 1e-6, fld
        p = o.field()
        ch.calcRHS()
        o.calc_rhs(u, v, w, p)
        assert rel_err(interior(ch.download("p")), interior(p)) < 1e-6
        ch.upload("p", p.astype(np.float32).astype(np.float64))
        ch.execute_poisson_solver()
        o.poisson(p)
        a, b = interior(ch.download("p")), interior(p)
        assert rel_err(a - a.mean(), b - b.mean()) < 1e-4


@pytest.mark.parametrize("name", ["c1", "c2", "tall"])
def test_fp32_trid_matches_oracle(name):
    """The FP32-mixed tridiagonal solve (FP64 arithmetic with reciprocal
    pivots, FP32 work array) against the FP64 oracle on the same random
    right-hand side, for short and tall columns."""
    print("CASE fp32 trid", name, flush=True)
    o = make_oracle(name)
    with make_channel(name, precision="fp32_mixed") as ch:
        rng = np.random.default_rng(5)
        spec = rng.standard_normal((ch.nz, ch.ny, ch.nx)).astype(np.float32).astype(np.float64)
        ch.upload_spec(spec)
        ch.solve_trid()
        got = ch.download_spec()
    o.solve_trid(spec)
    err = np.max(np.abs(got - spec)) / np.max(np.abs(spec))
    print(f"  max error / max |x| = {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("name,steps", [("c1", 10), ("r3", 10), ("c2", 10), ("ragged", 10),
                                        ("deep512", 5), ("deep700", 5)])
def test_fp32_short_horizon_pointwise(name, steps):
    """10 steps (5 on the deep columns, whose step solve is partitioned into
    32 segments): u, v, w within 1e-5 (relative to max|u|) of the FP64 oracle,
    divergence at the FP32 floor, bulk velocity held to 1e-6."""
    print("CASE fp32 short", name, steps, flush=True)
    o, f = _seeded(name)
    with make_channel(name, precision="fp32_mixed") as ch:
        for fld, a in zip(VEL + ("p",), (f[0], f[1], f[2], f[6])):
            ch.upload(fld, a)
        ch.step(steps)
        o.step(steps, *f)
        umax = np.max(np.abs(interior(f[0])))
        for fld, a in zip(VEL, f[:3]):
            err = np.max(np.abs(interior(ch.download(fld)) - interior(a))) / umax
            print(f"  {fld}: max err / max|u| = {err:.2e}")
            assert err < 1e-5, (fld, err)
        div, _ = ch.calcMaxDiv()
        dzmin = np.min(o.mesh()["dzf"])
        assert div < 1e-5 * umax / dzmin, div
        assert abs(ch.Mean("u") - 1.0) < 1e-6


@pytest.mark.parametrize("name", ["c3", "c4"])
def test_fp32_short_horizon_full_grids(name):
    """The whole 768 x 384 x 384 and 1536 x 768 x 512 grids (log profile +
    noise; the second needs ~90 GB of device memory for the two contexts): 10
    FP32-mixed steps against 10 FP64 steps of the GPU path (which matches the
    oracle on the first grid to ~5e-15, test_gpu_parity), u, v, w within 1e-5
    of max|u|, divergence at the FP32 floor, bulk velocity held."""
    steps = 10
    print("CASE fp32 short", name, steps, flush=True)
    with make_channel(name) as ref, make_channel(name, precision="fp32_mixed") as ch:
        for c in (ref, ch):
            c.initFlow("log", noise_intensity=0.1)
            c.step(steps)
        umax = None
        for fld in VEL:
            a = interior(ref.download(fld))
            if umax is None:
                umax = np.max(np.abs(a))
            err = np.max(np.abs(interior(ch.download(fld)) - a)) / umax
            print(f"  {fld}: max err / max|u| = {err:.2e}", flush=True)
            assert err < 1e-5, (fld, err)
        div, _ = ch.calcMaxDiv()
        dzmin = float(np.min(dzf_of(name)[1:-1]))
        print(f"  max div {div:.2e} (bar {1e-5 * umax / dzmin:.2e})", flush=True)
        assert div < 1e-5 * umax / dzmin, div
        assert abs(ch.Mean("u") - 1.0) < 1e-6


def test_fp32_mean_profile_and_stress_long_horizon():
    """200 steps of the 64^3 channel: mean velocity profile within 0.5% and the
    u'w' stress profile within 1% of its peak, against the FP64 oracle."""
    name, steps = "c1", 200
    print("CASE fp32 long", name, steps, flush=True)
    o, f = _seeded(name)
    with make_channel(name, precision="fp32_mixed") as ch:
        for fld, a in zip(VEL + ("p",), (f[0], f[1], f[2], f[6])):
            ch.upload(fld, a)
        ch.step(steps)
        o.step(steps, *f)
        um32, uw32 = _profiles(ch.download("u"), ch.download("v"), ch.download("w"))
    um64, uw64 = _profiles(f[0], f[1], f[2])
    prof_err = np.max(np.abs(um32 - um64)) / np.max(np.abs(um64))
    stress_err = np.max(np.abs(uw32 - uw64)) / max(np.max(np.abs(uw64)), 1e-30)
    print(f"  mean-profile err {prof_err:.2e}, u'w' err {stress_err:.2e}")
    assert prof_err < 5e-3
    assert stress_err < 1e-2
