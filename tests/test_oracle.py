"""Pins of the CPU FP64 oracle (CPU only, no GPU).

The reference ships no tests, golden vectors or outputs, so every oracle
function is pinned by an independent check: the DFT definition evaluated in
extended precision, numpy's FFT fixtures, closed forms, conservation and
divergence identities of the discrete operators, dense linear algebra, and the
known answer the C++ standard gives for std::minstd_rand.  A plausible
mistake -- a dropped term, a wrong index or sign, a mispaired eigenvalue --
fails at least one of them.
"""
import math
import os

import numpy as np
import pytest

from cases import interior, make_oracle, params, random_field

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fft_vectors.npz")


@pytest.fixture(scope="module")
def po(oracle_lib):
    return oracle_lib


# ---- FFT ---------------------------------------------------------------------------
def _dft_longdouble(z, sign):
    n = z.size
    j = np.arange(n, dtype=np.longdouble)
    ang = sign * 2 * np.pi * np.outer(j, j) / n
    ang = np.mod(np.outer(j, j), n).astype(np.longdouble) * (sign * 2 * np.longdouble(np.pi) / n)
    zr = z.real.astype(np.longdouble)
    zi = z.imag.astype(np.longdouble)
    re = (np.cos(ang) * zr - np.sin(ang) * zi).sum(axis=1)
    im = (np.sin(ang) * zr + np.cos(ang) * zi).sum(axis=1)
    return re.astype(np.float64) + 1j * im.astype(np.float64)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 20, 30, 48, 64, 96])
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_against_extended_precision_definition(po, n, sign):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = po.dft(z, sign)
    ref = _dft_longdouble(z, sign)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref)) * max(1, math.log2(n))


def test_dft_against_numpy_golden(po):
    g = np.load(GOLDEN)
    sizes = sorted({int(k.split("_")[1]) for k in g.files if k.startswith("x_")})
    for n in sizes:
        z = g[f"z_{n}"]
        np.testing.assert_allclose(po.dft(z, -1), g[f"fft_{n}"], rtol=0, atol=1e-13 * n)
        np.testing.assert_allclose(po.dft(z, +1), g[f"ifft_unscaled_{n}"], rtol=0, atol=1e-13 * n)


def test_rfft_perm_packing_against_numpy_golden(po):
    """perm = [Re X0, Re X(n/2), Re X1, Im X1, ...] (src/mod_fft.cc:79-86)."""
    g = np.load(GOLDEN)
    sizes = sorted({int(k.split("_")[1]) for k in g.files if k.startswith("x_")})
    for n in sizes:
        x, X = g[f"x_{n}"], g[f"rfft_{n}"]
        perm = po.rfft_perm(x)
        want = np.empty(n)
        want[0], want[1] = X[0].real, X[n // 2].real
        want[2::2], want[3::2] = X[1:n // 2].real, X[1:n // 2].imag
        np.testing.assert_allclose(perm, want, rtol=0, atol=1e-13 * n)
        # unscaled inverse: n * x (src/mod_fft.cc:105 + the 1/(nx*ny) of :57)
        np.testing.assert_allclose(po.irfft_perm(perm), n * x, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("n", [6, 8, 12, 20, 64, 96])
def test_eigenvalues_diagonalise_periodic_second_difference(po, n):
    """rfft_perm(D2 x) = lambda_perm * rfft_perm(x): ties the eigenvalue order
    (src/mod_fft.cc:206-218) to the perm packing."""
    lam = po.eigen_values(n)
    m = np.arange(n // 2 + 1)
    closed = 2.0 * (np.cos(2 * np.pi * m / n) - 1.0)
    want = np.empty(n)
    want[0], want[1] = closed[0], closed[n // 2]
    want[2::2], want[3::2] = closed[1:n // 2], closed[1:n // 2]
    np.testing.assert_allclose(lam, want, rtol=0, atol=1e-15)
    x = np.random.default_rng(n).standard_normal(n)
    d2 = np.roll(x, -1) - 2 * x + np.roll(x, 1)
    np.testing.assert_allclose(po.rfft_perm(d2), lam * po.rfft_perm(x), rtol=0, atol=1e-12)


# ---- mesh ----------------------------------------------------------------------------
def test_mesh_tanh_closed_form(po):
    nz, lz, r = 64, 2.0, 1.5
    dzf = po.mesh_by_func(1, 0, r, nz, lz)
    k = np.arange(nz + 1)
    zf = 0.5 * (1 + np.tanh((k / nz - 0.5) * r) / np.tanh(r / 2))
    np.testing.assert_allclose(dzf[1:-1], lz * np.diff(zf), rtol=0, atol=1e-15)
    assert dzf[0] == dzf[1] and dzf[-1] == dzf[-2]
    assert abs(dzf[1:-1].sum() - lz) < 1e-14
    np.testing.assert_allclose(dzf[1:-1], dzf[1:-1][::-1], rtol=1e-13)
    u = po.mesh_by_func(0, 0, r, 10, 3.0)
    np.testing.assert_array_equal(u, np.full(12, 3.0 / 10))


# ---- halos ---------------------------------------------------------------------------
def _closed_form_bc(f, kind, u_crf=0.0):
    """Final halo values at one rank (src/mod_updateBound.cc:11-296)."""
    nzp, nyp, nxp = f.shape
    nx, ny, nz = nxp - 2, nyp - 2, nzp - 2
    wrap = lambda n: np.r_[n, np.arange(1, n + 1), 1]
    g = f.copy()
    core = f[1:nz + 1][:, wrap(ny)][:, :, wrap(nx)]
    g[1:nz + 1] = core
    if kind == "u":
        g[0] = 2 * (0.0 - u_crf) - g[1]
        g[nz + 1] = 2 * (0.0 - u_crf) - g[nz]
    elif kind == "v":
        g[0] = 2 * 0.0 - g[1]
        g[nz + 1] = 2 * 0.0 - g[nz]
    elif kind == "w":
        g[0] = 0.0
        g[nz] = 0.0
        g[nz + 1] = 0.0
    else:
        g[0] = g[1]
        g[nz + 1] = g[nz]
    return g


@pytest.mark.parametrize("name", ["tiny", "r235"])
def test_update_bound_closed_form(po, name):
    o = make_oracle(name)
    u, v, w, p = (random_field(name, s) for s in range(4))
    want = [_closed_form_bc(u, "u"), _closed_form_bc(v, "v"), _closed_form_bc(w, "w"),
            _closed_form_bc(p, "p")]
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    for got, ref in zip((u, v, w, p), want):
        np.testing.assert_array_equal(got, ref)


# ---- stencils ---------------------------------------------------------------------------
def test_rk1_conserves_momentum_up_to_wall_shear(po):
    """Divergence-form convection telescopes to zero with periodic x/y and
    impermeable walls; the viscous term leaves only the wall shear.  Checks the
    whole RK1 operator (src/mod_calcVel.cc:27-129) for dropped or misindexed terms."""
    name = "r235"
    o = make_oracle(name)
    p = params(name)
    m = o.mesh()
    u, v, w = (random_field(name, 90 + d) for d in range(3))
    o.update_bound_vel(u, v, w)
    u1, v1, w1 = (o.field() for _ in range(3))
    o.rk1(u, v, w, u1, v1, w1)
    dzf = m["dzf"][1:-1][:, None, None]
    dzc_inv = m["dzc_inv"]
    re_inv = 1.0 / p["re"]
    for q, q1 in ((u, u1), (v, v1)):
        tend = (interior(q1) - interior(q)) / p["dt"]
        total = (tend * dzf).sum()
        nz = p["nz"]
        top = ((q[nz + 1] - q[nz]) * dzc_inv[nz])[1:-1, 1:-1].sum()
        bot = ((q[1] - q[0]) * dzc_inv[0])[1:-1, 1:-1].sum()
        wall = (top - bot) * re_inv
        scale = np.abs(tend * dzf).sum()
        assert abs(total - wall) <= 1e-12 * scale, (total, wall)


def test_rk1_laminar_profile_only_diffuses(po):
    """u(z) with v = w = 0: convection vanishes and u1 = u + dt*re_inv*Dzz u."""
    name = "r235"
    o = make_oracle(name)
    p = params(name)
    m = o.mesh()
    zc = m["zc_global"]
    u = o.field()
    u[:] = (zc * (2 - zc))[:, None, None]
    v, w = o.field(), o.field()
    o.update_bound_vel(u, v, w)
    u1, v1, w1 = (o.field() for _ in range(3))
    o.rk1(u, v, w, u1, v1, w1)
    col = u[:, 1, 1]
    k = np.arange(1, p["nz"] + 1)
    dzz = ((col[k + 1] - col[k]) * m["dzc_inv"][k] - (col[k] - col[k - 1]) * m["dzc_inv"][k - 1]) \
        * m["dzf_inv"][k]
    want = col[k] + p["dt"] * (dzz / p["re"])
    np.testing.assert_allclose(u1[1:-1, 3, 4], want, rtol=1e-14, atol=1e-15)
    assert np.all(interior(v1) == 0) and np.all(interior(w1) == 0)


@pytest.mark.parametrize("name", ["tiny", "r235", "c1"])
def test_projection_removes_divergence(po, name):
    """calcRHS -> Poisson -> BCp -> correctVel leaves max|div u| at round-off
    (the reference's own runtime check, src/mod_utils.cc:164-205)."""
    o = make_oracle(name)
    u, v, w = (random_field(name, 100 + d) for d in range(3))
    o.update_bound_vel(u, v, w)
    d0, _ = o.max_div(u, v, w)
    p = o.field()
    o.calc_rhs(u, v, w, p)
    o.poisson(p)
    o.update_bound_p(p)
    o.correct_vel(p, u, v, w)
    o.update_bound_vel(u, v, w)
    d1, _ = o.max_div(u, v, w)
    assert d1 < 1e-13 * d0, (d0, d1)


@pytest.mark.parametrize("name", ["tiny", "r235"])
def test_poisson_manufactured_solution(po, name):
    """rhs = -Lap_h(p*) built through correctVel and calcRHS; the solve returns
    -p* up to a constant (pins FFT packing, eigenvalues, pivots and wall terms)."""
    o = make_oracle(name)
    pstar = random_field(name, 7)
    o.update_bound_p(pstar)
    u, v, w = (o.field() for _ in range(3))
    o.correct_vel(pstar, u, v, w)
    o.update_bound_vel(u, v, w)
    p = o.field()
    o.calc_rhs(u, v, w, p)
    o.poisson(p)
    s = interior(p) + interior(pstar)
    assert np.max(np.abs(s - s.mean())) < 1e-11 * np.max(np.abs(interior(pstar)))


def test_thomas_against_dense_solve(po):
    name = "tiny"
    o = make_oracle(name)
    p = params(name)
    t = o.trid()
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    dzf0 = o.mesh()["dzf"]
    a, b0, c = po.trid_coeff(dzf0)
    np.testing.assert_array_equal(a, t["a"])
    np.testing.assert_array_equal(c, t["c"])
    # finite-volume form of the wall-normal operator with Neumann walls
    k = np.arange(nz)
    a_fv = 1.0 / (dzf0[k + 1] * 0.5 * (dzf0[k] + dzf0[k + 1]))
    c_fv = 1.0 / (dzf0[k + 1] * 0.5 * (dzf0[k + 2] + dzf0[k + 1]))
    np.testing.assert_allclose(a[1:], a_fv[1:], rtol=1e-14)
    np.testing.assert_allclose(c[:-1], c_fv[:-1], rtol=1e-14)
    assert a[0] == 0 and c[-1] == 0
    spec = np.random.default_rng(3).standard_normal((nz, ny, nx))
    rhs = spec.copy()
    o.solve_trid(spec)
    for j in range(ny):
        for i in range(nx):
            if i == 0 and j == 0:
                continue
            lam = t["lambdax"][i] + t["lambday"][j]
            A = np.diag(b0 + lam) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
            np.testing.assert_allclose(A @ spec[:, j, i], rhs[:, j, i], rtol=0,
                                       atol=1e-10 * np.abs(rhs[:, j, i]).max())
    # singular kx=ky=0 column: pinned last value, and A x = r up to the null space
    assert spec[-1, 0, 0] == 0.0


def test_singular_pin_does_not_change_velocity(po):
    name = "r235"
    u, v, w = (random_field(name, 200 + d) for d in range(3))
    res = []
    for pin in (1, 0):
        o = make_oracle(name, pin_singular=pin)
        uu, vv, ww = u.copy(), v.copy(), w.copy()
        o.update_bound_vel(uu, vv, ww)
        p = o.field()
        o.calc_rhs(uu, vv, ww, p)
        o.poisson(p)
        o.update_bound_p(p)
        o.correct_vel(p, uu, vv, ww)
        res.append((uu, vv, ww))
    for a, b in zip(*res):
        assert np.max(np.abs(interior(a) - interior(b))) < 1e-9 * np.max(np.abs(interior(b)))


# ---- reductions, checks, init ----------------------------------------------------------
def test_mean_constant_field(po):
    o = make_oracle("r235")
    f = o.field()
    f[:] = 3.25
    assert abs(o.mean(f) - 3.25) < 1e-14
    # src/mod_initFlow.cc:92 quirk: Mean(nz, ny) = (nx/nz) * bulk
    p = params("r235")
    assert abs(o.mean(f, p["nz"], p["ny"]) - 3.25 * p["nx"] / p["nz"]) < 1e-13


def test_checks_find_planted_extrema(po):
    o = make_oracle("r235")
    u, v, w = (o.field() for _ in range(3))
    u[5, 4, 3] = 7.0  # (i,j,k) = (3,4,5); cells i=3 and i=4 tie, the first wins
    dxi = 1.0 / (2 * math.pi / 24)
    val, ijk = o.max_div(u, v, w)
    assert ijk == (3, 4, 5) and abs(val - 7.0 * dxi) < 1e-14 * val
    cval, cijk = o.max_cfl(u, v, w)
    assert cijk == (3, 4, 5) and abs(cval - 3.5 * dxi * params("r235")["dt"]) < 1e-14 * cval
    assert o.check_nan(u)
    u[1, 1, 1] = np.nan
    assert not o.check_nan(u)


def test_minstd_rand_known_answer(po):
    """C++ [rand.predef]: the 10000th draw of a default-seeded minstd_rand is 399268537."""
    x = po.noise_uniform(1, 9999)
    assert round(x * 2147483646.0 + 1) == 399268537


def test_init_flow_bulk_quirk(po):
    """Without noise the Poiseuille init has bulk velocity nz/nx (src/mod_initFlow.cc:92)."""
    o = make_oracle("r235")
    p = params("r235")
    u, v, w = (o.field() for _ in range(3))
    o.init_flow(u, v, w, "poi", with_noise=False)
    assert abs(o.mean(u) - p["nz"] / p["nx"]) < 1e-14
    assert np.all(interior(v) == 0)


def test_full_step_invariants(po):
    """After RK2 steps the flow is divergence-free to round-off and the forcing
    holds the bulk velocity at vel_force (src/mod_calcVel.cc:280-320)."""
    o = make_oracle("r235")
    u, v, w, u1, v1, w1, p = (o.field() for _ in range(7))
    o.init_flow(u, v, w, "poi", noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    o.step(5, u, v, w, u1, v1, w1, p)
    d, _ = o.max_div(u, v, w)
    assert d < 1e-12
    assert abs(o.mean(u) - 1.0) < 1e-13
    assert o.check_nan(u) and o.check_nan(p)
