"""The closure's coupling kernels (csrc/k_closure.hip) through the C-ABI
against the CPU restatement (oracle/pll_closure_oracle.py): the box filter,
the training samples (features and labels), and the RK stages with the
eddy-viscosity stress, whose network runs inside the stage kernel."""
import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll
from cases import interior, make_channel, make_oracle, params
from oracle import pll_closure_oracle as mo
from oracle import pll_oracle as po

pytestmark = pytest.mark.gpu

FINE = dict(nx=24, ny=20, nz=12)
FACT = (2, 2, 3)


def _pair(precision_f, precision_c, lz=2.0, stretch=1.5):
    nf = FINE
    nc = dict(nx=nf["nx"] // FACT[0], ny=nf["ny"] // FACT[1], nz=nf["nz"] // FACT[2])
    lx, ly = 2 * np.pi, np.pi
    dzf_f = po.mesh_by_func(1, 0, stretch, nf["nz"], lz)
    dzf_c = po.mesh_by_func(1, 0, stretch, nc["nz"], lz)
    fine = pll.Channel(nf["nx"], nf["ny"], nf["nz"], lx, ly, lz, 2800.0, 1e-3, dzf_f,
                       precision=precision_f)
    coarse = pll.Channel(nc["nx"], nc["ny"], nc["nz"], lx, ly, lz, 2800.0, 4e-3, dzf_c,
                         precision=precision_c)
    of = po.Oracle(nf["nx"], nf["ny"], nf["nz"], lx, ly, lz, 2800.0, 1e-3, dzf_f)
    oc = po.Oracle(nc["nx"], nc["ny"], nc["nz"], lx, ly, lz, 2800.0, 4e-3, dzf_c)
    return fine, coarse, of, oc, dzf_f, (lx / nc["nx"], ly / nc["ny"])


def _random_fine(fine, of, seed):
    rng = np.random.default_rng(seed)
    u, v, w = (rng.standard_normal(fine.field_shape) for _ in range(3))
    u += 1.0
    of.update_bound_vel(u, v, w)
    for f, a in zip("uvw", (u, v, w)):
        fine.upload(f, a)
    # what the fine context holds (FP32 contexts round on upload)
    return fine.download("u"), fine.download("v"), fine.download("w")


@pytest.mark.parametrize("pf,pc", [("fp64", "fp64"), ("fp32_mixed", "fp64"),
                                   ("fp32_mixed", "fp32_mixed")])
def test_filter_matches_oracle(pf, pc):
    print(f"CASE filter {pf}->{pc}", flush=True)
    fine, coarse, of, oc, dzf_f, _ = _pair(pf, pc)
    u, v, w = _random_fine(fine, of, 1)
    coarse.filter_from(fine)
    ref = [oc.field() for _ in range(3)]
    for r, a in zip(ref, mo.box_filter(u, v, w, dzf_f, *FACT)):
        r[1:-1, 1:-1, 1:-1] = a
    oc.update_bound_vel(*ref)
    tol = 1e-13 if pc == "fp64" else 1e-6
    for f, r in zip("uvw", ref):
        got = coarse.download(f)
        assert np.max(np.abs(got - r)) <= tol * np.max(np.abs(r)), f
    fine.close()
    coarse.close()


@pytest.mark.parametrize("pf,pc", [("fp64", "fp64"), ("fp32_mixed", "fp32_mixed")])
def test_samples_match_oracle(pf, pc):
    import torch
    print(f"CASE samples {pf}->{pc}", flush=True)
    fine, coarse, of, oc, dzf_f, (dx, dy) = _pair(pf, pc)
    u, v, w = _random_fine(fine, of, 2)
    cl = pll.Closure(seed=3)
    gs, ys, nus = 0.05, 3.0, 1e-3
    cl.set_scales(gs, ys, nus)
    n = coarse.nx * coarse.ny * coarse.nz
    x = torch.full((n, 10), float("nan"), device="cuda:0")
    y = torch.full((n,), float("nan"), device="cuda:0")
    cl.samples(fine, coarse, x.data_ptr(), y.data_ptr())
    x, y = x.cpu().numpy(), y.cpu().numpy()
    # features of the filtered field the coarse context holds
    U, V, W = (coarse.download(f) for f in "uvw")
    mesh = oc.mesh()
    hz, yw, zs = mo.feat_mesh(mesh, 2.0)
    g = mo.cell_gradients(U, V, W, dx, dy, mesh, hz)
    xf = mo.features(g, gs, ys, yw, zs)
    if pc == "fp64":
        np.testing.assert_array_equal(x, xf)
    else:
        np.testing.assert_allclose(x, xf, rtol=1e-5, atol=1e-6)
    yl = mo.labels(u, v, w, dzf_f, *FACT, g, gs, nus)
    assert 0 < np.count_nonzero(yl) < n  # both clipped and positive labels occur
    np.testing.assert_allclose(y, yl, rtol=1e-5 if pc == "fp64" else 1e-3, atol=1e-6)
    cl.close()
    fine.close()
    coarse.close()


def _closure_with_params(seed, re):
    cl = pll.Closure(seed=seed)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, cl.nparam).astype(np.float32)
    p[mo.O_W1:mo.O_B1].reshape(16, 64)[10:] = 0.0
    cl.set_params(p)
    cl.set_scales(0.05, 2.0, 0.7 / re)
    return cl, p


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("name,crf", [("r235", 0.0), ("gen", 0.0), ("ragged", 0.0), ("tiny", 0.0),
                                      ("r235", 0.3)])
def test_closure_stages_match_oracle(name, crf, precision):
    print(f"CASE closure stages {name} u_crf={crf} {precision}", flush=True)
    prm = params(name)
    kw = dict(u_crf=crf) if crf else {}
    o = make_oracle(name, **kw)
    u, v, w, u1, v1, w1 = (o.field() for _ in range(6))
    o.init_flow(u, v, w, "poi", noise_intensity=0.2)
    u -= crf  # the state in the moving frame
    o.update_bound_vel(u, v, w)
    cl, p = _closure_with_params(11, prm["re"])
    gs, ys, nus = cl.scales()
    mesh = o.mesh()
    hz, yw, zs = mo.feat_mesh(mesh, prm["lz"])
    dx, dy = prm["lx"] / prm["nx"], prm["ly"] / prm["ny"]
    dt = prm["dt"]
    with make_channel(name, precision=precision, **kw) as ch:
        for f, a in zip("uvw", (u, v, w)):
            ch.upload(f, a)
        U, V, W = (ch.download(f) for f in "uvw")
        ch.attach_closure(cl)
        # stage 1
        ch.timeIntVelRK1()
        nut = ch.closure_nut()
        x = mo.features(mo.cell_gradients(U, V, W, dx, dy, mesh, hz), gs, ys, yw, zs)
        want = mo.predict(p, x).reshape(nut.shape) * nus
        err = np.abs(nut - want) / np.max(want)
        assert np.mean(err > 1e-4) < 0.01 and np.max(err) < 1e-2, (np.mean(err > 1e-4), err.max())
        o.rk1(U, V, W, u1, v1, w1)
        su, sv, sw = mo.sgs_divergence(nut, U, V, W, dx, dy, mesh)
        assert np.max(np.abs(su)) > 1e-3 * np.max(np.abs(interior(u1) - interior(U))) / dt
        tol = 1e-12 if precision == "fp64" else 1e-5
        got1 = [ch.download(f) for f in ("u1", "v1", "w1")]
        for g_, r, s in zip(got1, (u1, v1, w1), (su, sv, sw)):
            ref = interior(r) + dt * s
            assert np.max(np.abs(interior(g_) - ref)) <= tol * np.max(np.abs(ref))
        # stage 2 from the context's own predictor
        ch.updateBoundVel(1)
        A1, B1, C1 = (ch.download(f) for f in ("u1", "v1", "w1"))
        ch.timeIntVelRK2()
        nut2 = ch.closure_nut()
        uu, vv, ww = U.copy(), V.copy(), W.copy()
        o.rk2(uu, vv, ww, A1, B1, C1)
        su, sv, sw = mo.sgs_divergence(nut2, A1, B1, C1, dx, dy, mesh)
        for f, r, s in zip("uvw", (uu, vv, ww), (su, sv, sw)):
            ref = interior(r) + (dt * s) * 0.5
            assert np.max(np.abs(interior(ch.download(f)) - ref)) <= tol * np.max(np.abs(ref)), f
        ch.attach_closure(None)
    cl.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_closure_steps_stay_divergence_free(precision):
    name = "r235"
    print(f"CASE closure steps {precision}", flush=True)
    prm = params(name)
    cl, _ = _closure_with_params(12, prm["re"])
    with make_channel(name, precision=precision) as ch:
        ch.initFlow("poi", noise_intensity=0.2)
        ch.attach_closure(cl)
        ch.step(20)
        div, _ = ch.calcMaxDiv()
        assert div < (1e-10 if precision == "fp64" else 1e-3), div
        assert ch.checkNaN("u")
        nut = ch.closure_nut()
        assert np.all(nut >= 0) and np.max(nut) > 0
        assert abs(ch.Mean("u") - 1.0) < 1e-6  # the forcing holds the bulk velocity
        ch.attach_closure(None)
    cl.close()


def _closure_run(name, precision, seed, run):
    """Fields after run(ch) from one seeded state with a closure attached."""
    prm = params(name)
    cl, _ = _closure_with_params(seed, prm["re"])
    o = make_oracle(name)
    u, v, w, p = (o.field() for _ in range(4))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    with make_channel(name, precision=precision) as ch:
        for f, a in zip("uvwp", (u, v, w, p)):
            ch.upload(f, a)
        ch.attach_closure(cl)
        run(ch)
        out = [ch.download(f) for f in "uvwp"]
        ch.attach_closure(None)
    cl.close()
    return out


@pytest.mark.parametrize("name,precision", [("r235", "fp64"), ("r3", "fp64"),
                                            ("r235", "fp32_mixed")])
def test_closure_multistep_folding_matches_single_steps(name, precision):
    """With a closure attached, pll_step(n) applies each step's projection,
    forcing and boundary update in one pass before the next step's stage 1,
    sums the forcing's bulk mean from the predictor in stage 2 and runs calcRHS
    inside the x transform (rebuilding u's boundary on load).
    The projected values are the separate projection's; the bulk mean is the
    same sum of other roundings, so n steps in one call match n calls of one
    step to round-off."""
    print(f"CASE closure folding {name} {precision}", flush=True)

    def fused(ch):
        ch.step(6)

    def single(ch):
        for _ in range(6):
            ch.step(1)
    a = _closure_run(name, precision, 14, fused)
    b = _closure_run(name, precision, 14, single)
    tol = 1e-12 if precision == "fp64" else 2e-6
    for x, y, f in zip(a, b, "uvwp"):
        np.testing.assert_allclose(x, y, rtol=0, atol=tol, err_msg=f)


@pytest.mark.parametrize("name", ["r235", "gen"])
def test_closure_step_matches_entry_point_sequence(name):
    """pll_step with a closure attached (fused sequence where the x transform
    has its calcRHS kernel, r235; the reference's sequence otherwise, gen)
    matches the reference's call sequence through the entry points
    (src/main.cc:116-154) to FP64 round-off."""
    print(f"CASE closure entry sequence {name}", flush=True)

    def step(ch):
        ch.step(1)

    def calls(ch):
        ch.timeIntVelRK1(); ch.updateBoundVel(1); ch.timeIntVelRK2(); ch.updateBoundVel(0)
        ch.calcRHS(); ch.execute_poisson_solver(); ch.updateBoundP(); ch.correctVel()
        ch.forceVel(); ch.updateBoundVel(0)
    a = _closure_run(name, "fp64", 15, step)
    b = _closure_run(name, "fp64", 15, calls)
    for x, y, f in zip(a, b, "uvwp"):
        scale = max(np.max(np.abs(y)), 1e-30)
        assert np.max(np.abs(x - y)) <= 1e-13 * scale, (f, np.max(np.abs(x - y)) / scale)


def test_closure_detached_is_plain_step():
    """A detached closure leaves the plain FP64 step bit for bit."""
    name = "r235"
    cl, _ = _closure_with_params(13, params(name)["re"])
    with make_channel(name) as a, make_channel(name) as b:
        for ch in (a, b):
            ch.initFlow("poi", noise_intensity=0.1)
        b.attach_closure(cl)
        b.attach_closure(None)
        a.step(3)
        b.step(3)
        for f in "uvw":
            np.testing.assert_array_equal(a.download(f), b.download(f))
    cl.close()


def test_closure_api_errors():
    """Bad arguments fail with a status and a message, and leave the objects usable."""
    cl = pll.Closure(seed=1)
    for bad in ((0.0, 1.0, 1.0), (1.0, -1.0, 1.0), (1.0, 1.0, float("nan"))):
        with pytest.raises(pll.PllError):
            cl.set_scales(*bad)
    assert cl.scales() == (1.0, 1.0, 1.0)
    with make_channel("r235") as ch, make_channel("tiny") as other:
        with pytest.raises(pll.PllError, match="no closure attached"):
            ch.closure_nut()
        # grids that do not nest: 8x6x4 is not a coarsening of 24x20x12
        with pytest.raises(pll.PllError, match="integer multiples|boxes differ|unions"):
            other.filter_from(ch)
        with pytest.raises(pll.PllError):
            cl.samples(ch, other, 0, 0)
        ch.attach_closure(cl)
        ch.attach_closure(None)
        ch.step(1)  # the plain step still runs
    cl.close()
