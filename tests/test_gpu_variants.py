"""Configurations beyond the default run, against the oracle through the C-ABI:
a convecting reference frame (u_crf != 0, the u_crf terms of
src/mod_calcVel.cc:77-125 and the wall value -u_crf of src/mod_updateBound.cc:80),
forcing of v and w as well as u (forceVel's three components,
src/mod_calcVel.cc:280-320, which takes the step's unfused branch), no forcing,
and the unpinned singular mode (p then carries a round-off constant; the
FP32-mixed path pins it anyway, since an FP32 p cannot carry that constant,
and its velocities still match the unpinned FP64 oracle).

Each runs 10 steps from the same seeded state, in FP64 (the 1e-11 bar) and
FP32-mixed (1e-5 of max|u|, the FP32 step bar), and checks transform2CRF."""
import numpy as np
import pytest

from cases import interior, make_channel, make_oracle, rel_err

pytestmark = pytest.mark.gpu

VEL = ("u", "v", "w")
VARIANTS = {
    "crf": dict(u_crf=0.3),
    "forced_uvw": dict(is_forced=(1, 1, 1), vel_force=(1.0, 0.05, 0.0)),
    "unforced": dict(is_forced=(0, 0, 0)),
    "unpinned": dict(pin_singular=0),
}


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("name", ["r235", "r3"])
def test_variant_steps_match_oracle(name, variant, precision):
    print(f"CASE variant {name} {variant} {precision}", flush=True)
    kw = VARIANTS[variant]
    o = make_oracle(name, **kw)
    u, v, w, u1, v1, w1, p = (o.field() for _ in range(7))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    if "u_crf" in kw:
        u -= kw["u_crf"]  # the state in the moving frame (transform2CRF)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    with make_channel(name, precision=precision, **kw) as ch:
        for fld, arr in zip(VEL + ("p",), (u, v, w, p)):
            ch.upload(fld, arr)
        ch.step(10)
        o.step(10, u, v, w, u1, v1, w1, p)
        umax = np.max(np.abs(interior(u)))
        for fld, arr in zip(VEL, (u, v, w)):
            got = interior(ch.download(fld))
            if precision == "fp64":
                e = rel_err(got, interior(arr))
                assert e < 1e-11, (fld, e)
            else:
                e = np.max(np.abs(got - interior(arr))) / umax
                assert e < 1e-5, (fld, e)
        if precision == "fp64":
            a, b = interior(ch.download("p")), interior(p)
            assert rel_err(a - a.mean(), b - b.mean()) < 1e-11
        if kw.get("is_forced", (1, 0, 0))[0]:
            # held at vel_force[0] (a caller in the moving frame lowers it by u_crf)
            assert abs(ch.Mean("u") - 1.0) < (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_transform2crf(precision):
    """transform2CRF (src/mod_calcVel.cc:13-25): u interior -= u_crf; halos untouched."""
    name = "r235"
    print(f"CASE transform2crf {precision}", flush=True)
    o = make_oracle(name)
    u, v, w = (o.field() for _ in range(3))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    with make_channel(name, precision=precision) as ch:
        ch.upload("u", u)
        before = ch.download("u")
        ch.transform2CRF(0.25)
        after = ch.download("u")
    if precision == "fp64":
        np.testing.assert_array_equal(interior(after), interior(before) - 0.25)
    else:  # the difference is rounded to FP32
        np.testing.assert_allclose(interior(after), interior(before) - 0.25, rtol=0, atol=1e-7)
    halo = np.ones_like(after, dtype=bool)
    halo[1:-1, 1:-1, 1:-1] = False
    np.testing.assert_array_equal(after[halo], before[halo])
