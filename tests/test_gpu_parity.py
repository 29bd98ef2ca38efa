"""GPU parity of the HIP path against the CPU FP64 oracle, through the C-ABI.

Each routine gets the same host inputs on both sides.  Stencils, halos and the
tridiagonal solve keep the reference's operation order and are compared bit
for bit; the FFT stages use a different (but exact-in-arithmetic) factorisation
than the oracle and are compared to 1e-13 relative; the full RK2 step is
compared after 10 steps at 1e-11 relative.
"""
import numpy as np
import pytest

from cases import (CASES, interior, make_channel, make_oracle, params, random_field, rel_err)

pytestmark = pytest.mark.gpu

VEL = ("u", "v", "w")
SMALL = ["r235", "r3", "c1"]


def _print_case(*what):
    print("CASE", *what, flush=True)


@pytest.fixture(scope="module")
def ctxs():
    made = {}

    def get(name):
        if name not in made:
            made[name] = (make_channel(name), make_oracle(name))
        return made[name]

    yield get
    for ch, _ in made.values():
        ch.close()


def _bc_ready_fields(o, name, seed):
    """Random velocity fields with the oracle's boundary update applied."""
    u, v, w = (random_field(name, seed + d) for d in range(3))
    o.update_bound_vel(u, v, w)
    return u, v, w


@pytest.mark.parametrize("name", SMALL)
def test_update_bound_vel_bitexact(ctxs, name):
    _print_case("update_bound_vel", name)
    ch, o = ctxs(name)
    f = [random_field(name, 10 + d) for d in range(3)]
    for fld, arr in zip(VEL, f):
        ch.upload(fld, arr)
    ch.updateBoundVel(0)
    o.update_bound_vel(*f)
    for fld, arr in zip(VEL, f):
        np.testing.assert_array_equal(ch.download(fld), arr, err_msg=fld)
    # the (u1, v1, w1) variant
    g = [random_field(name, 20 + d) for d in range(3)]
    for fld, arr in zip(("u1", "v1", "w1"), g):
        ch.upload(fld, arr)
    ch.updateBoundVel(1)
    o.update_bound_vel(*g)
    for fld, arr in zip(("u1", "v1", "w1"), g):
        np.testing.assert_array_equal(ch.download(fld), arr, err_msg=fld)


@pytest.mark.parametrize("name", SMALL)
def test_impose_bc_vel_walls_only_bitexact(ctxs, name):
    # random halos that no exchange has made periodic: the wall rows must read
    # them as they stand, and nothing but the wall rows may change
    _print_case("impose_bc_vel", name)
    ch, o = ctxs(name)
    for which, names in ((0, VEL), (1, ("u1", "v1", "w1"))):
        f = [random_field(name, 30 + 3 * which + d) for d in range(3)]
        before = [a.copy() for a in f]
        for fld, arr in zip(names, f):
            ch.upload(fld, arr)
        ch.imposeBCVel(which)
        o.impose_bc_vel(*f)
        for fld, arr, b in zip(names, f, before):
            got = ch.download(fld)
            np.testing.assert_array_equal(got, arr, err_msg=fld)
            np.testing.assert_array_equal(got[1:-2], b[1:-2], err_msg=fld)


@pytest.mark.parametrize("name", SMALL)
def test_update_bound_p_bitexact(ctxs, name):
    _print_case("update_bound_p", name)
    ch, o = ctxs(name)
    p = random_field(name, 30)
    ch.upload("p", p)
    ch.updateBoundP()
    o.update_bound_p(p)
    np.testing.assert_array_equal(ch.download("p"), p)


@pytest.mark.parametrize("name", SMALL)
def test_rk1_rk2_bitexact(ctxs, name):
    _print_case("rk1/rk2", name)
    ch, o = ctxs(name)
    u, v, w = _bc_ready_fields(o, name, 40)
    u1, v1, w1 = (o.field() for _ in range(3))
    for fld, arr in zip(VEL, (u, v, w)):
        ch.upload(fld, arr)
    ch.timeIntVelRK1()
    o.rk1(u, v, w, u1, v1, w1)
    for fld, arr in zip(("u1", "v1", "w1"), (u1, v1, w1)):
        np.testing.assert_array_equal(interior(ch.download(fld)), interior(arr), err_msg=fld)
    # stage 2 from the oracle's u1 with its halos
    o.update_bound_vel(u1, v1, w1)
    for fld, arr in zip(("u1", "v1", "w1"), (u1, v1, w1)):
        ch.upload(fld, arr)
    ch.timeIntVelRK2()
    o.rk2(u, v, w, u1, v1, w1)
    for fld, arr in zip(VEL, (u, v, w)):
        np.testing.assert_array_equal(interior(ch.download(fld)), interior(arr), err_msg=fld)


@pytest.mark.parametrize("name", SMALL)
def test_calc_rhs_and_correct_bitexact(ctxs, name):
    _print_case("calc_rhs/correct", name)
    ch, o = ctxs(name)
    u, v, w = _bc_ready_fields(o, name, 50)
    for fld, arr in zip(VEL, (u, v, w)):
        ch.upload(fld, arr)
    p = o.field()
    ch.calcRHS()
    o.calc_rhs(u, v, w, p)
    np.testing.assert_array_equal(interior(ch.download("p")), interior(p))
    q = random_field(name, 55)
    o.update_bound_p(q)
    ch.upload("p", q)
    ch.correctVel()
    o.correct_vel(q, u, v, w)
    for fld, arr in zip(VEL, (u, v, w)):
        np.testing.assert_array_equal(interior(ch.download(fld)), interior(arr), err_msg=fld)


@pytest.mark.parametrize("name", SMALL + ["gen", "c3slab", "c4slab", "c2"])
def test_fft_stages(ctxs, name):
    _print_case("fft stages", name)
    ch, o = ctxs(name)
    p = random_field(name, 60)
    ch.upload("p", p)
    spec = o.spec()
    ch.execute_fft_fwd_xpen()
    o.fft_fwd_xpen(p, spec)
    assert rel_err(ch.download_spec(), spec) < 1e-13
    ch.upload_spec(spec)  # continue every stage from identical input
    ch.execute_fft_fwd_ypen()
    o.fft_fwd_ypen(spec)
    assert rel_err(ch.download_spec(), spec) < 1e-13
    ch.upload_spec(spec)
    ch.solve_trid()
    o.solve_trid(spec)
    np.testing.assert_array_equal(ch.download_spec(), spec)
    ch.upload_spec(spec)
    ch.execute_fft_bwd_ypen()
    o.fft_bwd_ypen(spec)
    assert rel_err(ch.download_spec(), spec) < 1e-13
    ch.upload_spec(spec)
    q = o.field()
    ch.upload("p", q)
    ch.execute_fft_bwd_xpen()
    o.fft_bwd_xpen(spec, q)
    assert rel_err(interior(ch.download("p")), interior(q)) < 1e-13


@pytest.mark.parametrize("name", SMALL + ["gen", "c3slab"])
def test_poisson_matches_oracle(ctxs, name):
    _print_case("poisson", name)
    ch, o = ctxs(name)
    u, v, w = _bc_ready_fields(o, name, 70)
    p = o.field()
    o.calc_rhs(u, v, w, p)
    ch.upload("p", p)
    ch.execute_poisson_solver()
    o.poisson(p)
    a, b = interior(ch.download("p")), interior(p)
    assert rel_err(a - a.mean(), b - b.mean()) < 1e-11


@pytest.mark.parametrize("name", SMALL)
def test_mean_force_checks(ctxs, name):
    _print_case("mean/force/checks", name)
    ch, o = ctxs(name)
    u, v, w = _bc_ready_fields(o, name, 80)
    u += 1.0
    for fld, arr in zip(VEL, (u, v, w)):
        ch.upload(fld, arr)
    assert abs(ch.Mean("u") - o.mean(u)) <= 1e-14 * abs(o.mean(u))
    dg, dgi = ch.calcMaxDiv()
    do, doi = o.max_div(u, v, w)
    assert abs(dg - do) <= 1e-14 * do and dgi == doi
    cg, cgi = ch.calcMaxCFL()
    co, coi = o.max_cfl(u, v, w)
    assert abs(cg - co) <= 1e-14 * co and cgi == coi
    m = ch.forceVel(return_mean=True)
    mo = o.force_vel(u, v, w)
    assert abs(m[0] - mo[0]) <= 1e-14 * abs(mo[0])
    assert rel_err(interior(ch.download("u")), interior(u)) < 1e-15
    assert ch.checkNaN("u")
    bad = u.copy()
    bad[3, 2, 1] = np.nan
    ch.upload("p", bad)
    assert not ch.checkNaN("p")


@pytest.mark.parametrize("name", ["c1", "r3"])
def test_init_flow_matches_oracle(ctxs, name):
    _print_case("init_flow", name)
    ch, o = ctxs(name)
    for prof in ("poi", "log"):
        u, v, w = (o.field() for _ in range(3))
        o.init_flow(u, v, w, prof, noise_intensity=0.1)
        ch.initFlow(prof, noise_intensity=0.1)
        for fld, arr in zip(VEL, (u, v, w)):
            assert rel_err(interior(ch.download(fld)), interior(arr)) < 1e-14, (prof, fld)


@pytest.mark.parametrize("name,steps", [("c1", 10), ("r235", 10), ("r3", 5), ("gen", 10), ("c2", 10),
                                        ("ragtile", 5), ("tall", 5), ("deep512", 5), ("deep700", 5)])
def test_full_step_parity(name, steps):
    """10 RK2 steps from the same seeded state: u, v, w and p - <p> to 1e-11."""
    _print_case("full step", name, steps)
    o = make_oracle(name)
    u, v, w, u1, v1, w1, p = (o.field() for _ in range(7))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    with make_channel(name) as ch:
        for fld, arr in zip(VEL + ("p",), (u, v, w, p)):
            ch.upload(fld, arr)
        ch.step(steps)
        o.step(steps, u, v, w, u1, v1, w1, p)
        for fld, arr in zip(VEL, (u, v, w)):
            e = rel_err(interior(ch.download(fld)), interior(arr))
            assert e < 1e-11, (fld, e)
        a, b = interior(ch.download("p")), interior(p)
        assert rel_err(a - a.mean(), b - b.mean()) < 1e-11
        div, _ = ch.calcMaxDiv()
        umax = np.max(np.abs(interior(u)))
        dzmin = np.min(o.mesh()["dzf"])
        assert div < 1e-12 * umax / dzmin, div


@pytest.mark.parametrize("name,steps", [("c3", 10), ("c4", 2)])
def test_full_step_parity_full_grids(name, steps):
    """The whole 768 x 384 x 384 and 1536 x 768 x 512 grids (log profile +
    noise): FP64 RK2 steps against the OpenMP oracle, u, v, w and p - <p> to
    1e-11 (about 20 s of oracle work on 16 cores for each grid; 7 host fields of
    0.92 GB and 4.9 GB)."""
    _print_case("full step", name, steps)
    o = make_oracle(name)
    u, v, w, u1, v1, w1, p = (o.field() for _ in range(7))
    o.init_flow(u, v, w, "log", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    with make_channel(name) as ch:
        for fld, arr in zip(VEL + ("p",), (u, v, w, p)):
            ch.upload(fld, arr)
        ch.step(steps)
        print("  GPU steps done; oracle steps", flush=True)
        o.step(steps, u, v, w, u1, v1, w1, p)
        del u1, v1, w1
        for fld, arr in zip(VEL, (u, v, w)):
            e = rel_err(interior(ch.download(fld)), interior(arr))
            print(f"  {fld}: {e:.2e}", flush=True)
            assert e < 1e-11, (fld, e)
        a, b = interior(ch.download("p")), interior(p)
        e = rel_err(a - a.mean(), b - b.mean())
        print(f"  p - <p>: {e:.2e}", flush=True)
        assert e < 1e-11


@pytest.mark.parametrize("name,precision", [("r3", "fp64"), ("c1", "fp64"), ("r3", "fp32_mixed")])
def test_multistep_folding_matches_single_steps(name, precision):
    """pll_step(n) folds each step's projection, forcing and final boundary
    update into the next step's loads and takes the forcing's bulk mean from
    the predictor (the projection leaves every x-row sum unchanged).  The
    projected values are bit-identical to the separate projection; the bulk
    mean is the same sum of other FP64 roundings, so n fused steps differ from
    n single steps by FP64 round-off (FP32 round-off in FP32-mixed)."""
    _print_case("folding", name, precision)
    o = make_oracle(name)
    u, v, w, p = (o.field() for _ in range(4))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    outs = []
    for mode in ("fused", "single"):
        with make_channel(name, precision=precision) as ch:
            for fld, arr in zip(VEL + ("p",), (u, v, w, p)):
                ch.upload(fld, arr)
            if mode == "fused":
                ch.step(6)
            else:
                for _ in range(6):
                    ch.step(1)
            outs.append([ch.download(f) for f in VEL + ("p",)])
    for a, b, fld in zip(outs[0], outs[1], VEL + ("p",)):
        if precision == "fp64":
            # fields of magnitude ~1: a few FP64 units in the last place
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-13, err_msg=fld)
        else:
            # FP32 fields of magnitude ~1: a few units in the last place
            np.testing.assert_allclose(a, b, rtol=0, atol=2e-6, err_msg=fld)


def test_entry_point_sequence_matches_step():
    """The reference's call sequence through the individual entry points
    (src/main.cc:116-154) matches pll_step to round-off."""
    name = "r3"
    _print_case("entry sequence", name)
    o = make_oracle(name)
    u, v, w, p = (o.field() for _ in range(4))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    res = []
    for mode in ("step", "calls"):
        with make_channel(name) as ch:
            for fld, arr in zip(VEL + ("p",), (u, v, w, p)):
                ch.upload(fld, arr)
            for _ in range(3):
                if mode == "step":
                    ch.step(1)
                else:
                    ch.timeIntVelRK1(); ch.updateBoundVel(1); ch.timeIntVelRK2(); ch.updateBoundVel(0)
                    ch.calcRHS(); ch.execute_poisson_solver(); ch.updateBoundP(); ch.correctVel()
                    ch.forceVel(); ch.updateBoundVel(0)
            res.append([ch.download(f) for f in VEL])
    for a, b in zip(*res):
        assert rel_err(a, b) < 1e-13
