"""Pins of the oracle's momentum stages and initial profile (CPU only).

Each check comes from the mathematics of the scheme, not from the oracle's own
expressions:

* separable fields reduce every term of the w-momentum tendency
  (src/mod_calcVel.cc:105-125) to a textbook finite difference: centred first
  differences for the convection by a constant or z-interpolated carrier
  velocity, and the stretched-grid second difference built from the face
  coordinates z_f(k) (closed-form tanh mesh) for the viscous term;
* the discrete w momentum is conserved up to the wall terms: divergence-form
  convection and the viscous fluxes telescope over the w control volumes, so
  sum((w1 - w)/dt * dzc) over k = 1..nz-1 equals a closed expression in the
  planes next to the walls;
* RK2 is Heun's method: stage 2 (src/mod_calcVel.cc:194,216,238) equals
  (u + RK1(u1)) / 2, with RK1 the operator already pinned by conservation,
  the laminar diffusion and the closed forms here;
* the "log" initial profile (src/mod_initFlow.cc:23-30) is the law of the
  wall, u+ = y+ up to y+ = 11.6 and u+ = ln(y+)/0.4 + 5.5 above (kappa = 0.4,
  B = 5.5), at y+ = Re_tau * (distance to the nearer wall)/h with Dean's
  correlation Re_tau = 0.09 Re^0.88 (Pope, Turbulent Flows, 2000, sec. 7.1),
  evaluated on the cell centres z_c/l_z; the reference passes its `re` for
  Re (a factor 2^0.88 low for Re_b = 2 re), which is kept.

tests/test_oracle_mutants.py plants mistakes in copies of the oracle and shows
that these pins fail on them.
"""
import numpy as np
import pytest

from cases import interior, make_oracle, params


def _faces(name):
    """Face coordinates z_f(0..nz) of the tanh mesh (src/mod_mesh.cc:95-125,
    mesh_type 1, stretch_func 0), cell widths dzf(1..nz), cell centres
    z_c(1..nz) and the distance between neighbouring centres dzc(1..nz-1)."""
    p = params(name)
    nz, lz, r = p["nz"], p["lz"], p["stretch"]
    k = np.arange(nz + 1)
    zf = lz * 0.5 * (1 + np.tanh((k / nz - 0.5) * r) / np.tanh(r / 2))
    dzf = np.r_[np.nan, np.diff(zf)]              # dzf[k], k = 1..nz
    zc = np.r_[np.nan, 0.5 * (zf[:-1] + zf[1:])]  # zc[k], k = 1..nz
    dzc = np.r_[np.nan, np.diff(zc[1:]), np.nan]  # dzc[k] = zc[k+1] - zc[k], k = 1..nz-1
    return zf, dzf, zc, dzc


def _periodic_xy(f):
    """Fill the x and y halo cells with their periodic images (in place)."""
    f[:, :, 0] = f[:, :, -2]
    f[:, :, -1] = f[:, :, 1]
    f[:, 0, :] = f[:, -2, :]
    f[:, -1, :] = f[:, 1, :]
    return f


def _tendency(o, name, u, v, w):
    """(u1 - u)/dt, (v1 - v)/dt, (w1 - w)/dt of the oracle's RK1 stage."""
    dt = params(name)["dt"]
    u1, v1, w1 = (o.field() for _ in range(3))
    o.rk1(u, v, w, u1, v1, w1)
    return (u1 - u) / dt, (v1 - v) / dt, (w1 - w) / dt


def _stage_rounding(name, f):
    """Scale of the rounding in (w1 - w)/dt: w1 = w + dt*F is rounded at the
    size of w, so the recovered F carries ~1e-16 |w| / dt; the pins' 1e-13
    tolerance is taken relative to this plus the terms' own size."""
    return 0.05 * np.abs(f).max() / params(name)["dt"]


def _w_plane_field(name, prof_ij):
    """w = prof_ij(j, i) on the faces k = 1..nz-1, zero on the wall faces
    k = 0 and k = nz and on the ghost plane nz+1 (the no-penetration state
    src/mod_updateBound.cc:89-141 leaves)."""
    p = params(name)
    nz = p["nz"]
    w = np.zeros((nz + 2, p["ny"] + 2, p["nx"] + 2))
    w[1:nz, 1:-1, 1:-1] = prof_ij
    return _periodic_xy(w)


NAME = "r235"   # 24 x 20 x 12, tanh stretch 1.5: a stretched wall-normal grid


def test_w_convection_and_diffusion_along_x():
    """u = U(z) (cell centres), v = 0, w = f(x): on the planes away from the
    walls the w tendency is -U_face * (f[i+1] - f[i-1])/(2 dx) + (f[i+1] - 2 f[i]
    + f[i-1])/(re dx^2), U_face the linear interpolation of U from the centres
    z_c(k), z_c(k+1) to the face z_f(k)."""
    p = params(NAME)
    o = make_oracle(NAME)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    dx = p["lx"] / nx
    rng = np.random.default_rng(11)
    f = rng.standard_normal(nx)
    zf, dzf, zc, dzc = _faces(NAME)
    U = np.zeros(nz + 2)
    U[1:nz + 1] = 1.0 + np.sin(2.0 * zc[1:])      # any centre profile
    u = _periodic_xy(np.broadcast_to(U[:, None, None], (nz + 2, ny + 2, nx + 2)).copy())
    v = np.zeros_like(u)
    w = _w_plane_field(NAME, np.broadcast_to(f[None, :], (ny, nx)))
    _, _, tw = _tendency(o, NAME, u, v, w)
    fe, fw = np.roll(f, -1), np.roll(f, 1)
    for k in range(2, nz - 1):
        lam = (zf[k] - zc[k]) / (zc[k + 1] - zc[k])
        uface = (1 - lam) * U[k] + lam * U[k + 1]
        want = -uface * (fe - fw) / (2 * dx) + (fe - 2 * f + fw) / (p["re"] * dx * dx)
        scale = np.abs(uface * f / dx).max() + np.abs(f).max() / (p["re"] * dx * dx)
        scale += _stage_rounding(NAME, f)
        for j in (1, ny // 2, ny):
            np.testing.assert_allclose(tw[k, j, 1:-1], want, rtol=0, atol=1e-13 * scale)


def test_w_convection_and_diffusion_along_y():
    """v = V(z), u = 0, w = g(y): the same closed form along y with dy."""
    p = params(NAME)
    o = make_oracle(NAME)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    dy = p["ly"] / ny
    rng = np.random.default_rng(12)
    g = rng.standard_normal(ny)
    zf, dzf, zc, dzc = _faces(NAME)
    V = np.zeros(nz + 2)
    V[1:nz + 1] = 0.7 - np.cos(1.3 * zc[1:])
    v = _periodic_xy(np.broadcast_to(V[:, None, None], (nz + 2, ny + 2, nx + 2)).copy())
    u = np.zeros_like(v)
    w = _w_plane_field(NAME, np.broadcast_to(g[:, None], (ny, nx)))
    _, _, tw = _tendency(o, NAME, u, v, w)
    gn, gs = np.roll(g, -1), np.roll(g, 1)
    for k in range(2, nz - 1):
        lam = (zf[k] - zc[k]) / (zc[k + 1] - zc[k])
        vface = (1 - lam) * V[k] + lam * V[k + 1]
        want = -vface * (gn - gs) / (2 * dy) + (gn - 2 * g + gs) / (p["re"] * dy * dy)
        scale = np.abs(vface * g / dy).max() + np.abs(g).max() / (p["re"] * dy * dy)
        scale += _stage_rounding(NAME, g)
        for i in (1, nx // 3, nx):
            np.testing.assert_allclose(tw[k, 1:-1, i], want, rtol=0, atol=1e-13 * scale)


def test_w_wall_normal_terms_on_stretched_grid():
    """u = v = 0, w = h(z) on the faces, zero at the walls: finite-volume form
    over the control volume [z_c(k), z_c(k+1)] of the face z_f(k),
    tw = -(W_c(k+1)^2 - W_c(k)^2)/dzc(k) + ((h[k+1]-h[k])/dzf(k+1) -
    (h[k]-h[k-1])/dzf(k))/(re dzc(k)), W_c the centre average of h."""
    p = params(NAME)
    o = make_oracle(NAME)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    zf, dzf, zc, dzc = _faces(NAME)
    h = np.zeros(nz + 2)
    h[1:nz] = np.sin(np.pi * zf[1:nz] / p["lz"]) * (1.0 + 0.5 * zf[1:nz])
    w = _periodic_xy(np.broadcast_to(h[:, None, None], (nz + 2, ny + 2, nx + 2)).copy())
    u, v = np.zeros_like(w), np.zeros_like(w)
    _, _, tw = _tendency(o, NAME, u, v, w)
    for k in range(1, nz):
        wc_top, wc_bot = 0.5 * (h[k] + h[k + 1]), 0.5 * (h[k - 1] + h[k])
        conv = (wc_top ** 2 - wc_bot ** 2) / dzc[k]
        visc = ((h[k + 1] - h[k]) / dzf[k + 1] - (h[k] - h[k - 1]) / dzf[k]) / dzc[k] / p["re"]
        scale = abs(h).max() ** 2 / dzc[k] + abs(h).max() / (dzf[1:].min() * dzc[k] * p["re"])
        scale += _stage_rounding(NAME, h)
        np.testing.assert_allclose(tw[k, 1:-1, 1:-1], visc - conv, rtol=0, atol=1e-13 * scale)


def test_w_momentum_conserved_up_to_wall_terms():
    """Random u, v, w after the boundary update: summed over the w control
    volumes (faces k = 1..nz-1, width dzc), convection and the viscous fluxes
    telescope, leaving per column
    -(w[nz-1]^2 - w[1]^2)/4 - (w[nz-1]/dzf(nz) + w[1]/dzf(1))/re."""
    p = params(NAME)
    o = make_oracle(NAME)
    nz = p["nz"]
    zf, dzf, zc, dzc = _faces(NAME)
    rng = np.random.default_rng(13)
    u, v, w = (rng.standard_normal((nz + 2, p["ny"] + 2, p["nx"] + 2)) for _ in range(3))
    o.update_bound_vel(u, v, w)
    _, _, tw = _tendency(o, NAME, u, v, w)
    total = sum((tw[k, 1:-1, 1:-1] * dzc[k]).sum() for k in range(1, nz))
    w1, wn = w[1, 1:-1, 1:-1], w[nz - 1, 1:-1, 1:-1]
    wall = (-0.25 * (wn ** 2 - w1 ** 2) - (wn / dzf[nz] + w1 / dzf[1]) / p["re"]).sum()
    scale = sum((np.abs(tw[k, 1:-1, 1:-1]) * dzc[k]).sum() for k in range(1, nz))
    assert abs(total - wall) <= 1e-12 * scale, (total, wall, scale)


@pytest.mark.parametrize("name", ["tiny", "r235"])
def test_rk2_is_heun_over_rk1(name):
    """Stage 2 of Heun's method: u_new = (u + RK1(u1)) / 2 for every component
    (src/mod_calcVel.cc:194,216,238 against :27-129)."""
    o = make_oracle(name)
    p = params(name)
    rng = np.random.default_rng(14)
    shape = (p["nz"] + 2, p["ny"] + 2, p["nx"] + 2)
    u, v, w, u1, v1, w1 = (rng.standard_normal(shape) for _ in range(6))
    o.update_bound_vel(u, v, w)
    o.update_bound_vel(u1, v1, w1)
    a, b, c = (o.field() for _ in range(3))
    o.rk1(u1, v1, w1, a, b, c)
    un, vn, wn = u.copy(), v.copy(), w.copy()
    o.rk2(un, vn, wn, u1, v1, w1)
    for new, old, star in ((un, u, a), (vn, v, b), (wn, w, c)):
        want = 0.5 * (interior(old) + interior(star))
        np.testing.assert_allclose(interior(new), want, rtol=0,
                                   atol=1e-15 * np.abs(want).max())


@pytest.mark.parametrize("name", ["r235", "r3"])
def test_log_profile_is_the_law_of_the_wall(name):
    """Without noise the "log" initial u is the law of the wall at the cell
    centres, scaled to the init's bulk (src/mod_initFlow.cc:23-30,92)."""
    o = make_oracle(name)
    p = params(name)
    nz = p["nz"]
    zf, dzf, zc, dzc = _faces(name)
    retau = 0.09 * p["re"] ** 0.88                # Dean's correlation
    yplus = np.minimum(zc[1:], p["lz"] - zc[1:]) / (0.5 * p["lz"]) * retau
    kappa, B = 0.4, 5.5
    law = np.where(yplus <= 11.6, yplus, np.log(yplus) / kappa + B)
    u, v, w = (o.field() for _ in range(3))
    o.init_flow(u, v, w, "log", with_noise=False)
    col = u[1:nz + 1, 1, 1]
    np.testing.assert_allclose(col / col.max(), law / law.max(), rtol=1e-12, atol=0)
    assert np.all(u[1:nz + 1, 1:-1, 1:-1] == col[:, None, None])
    # the init's normalisation by Mean(nz, ny) (src/mod_initFlow.cc:92)
    assert abs(o.mean(u) - nz / p["nx"]) < 1e-13
    assert np.all(interior(v) == 0) and np.all(interior(w) == 0)


def test_law_of_the_wall_known_values():
    """Known answers of the closed form used above: the viscous and log laws
    meet at y+ ~ 11.6 (11.63 against 11.6), and for Re = 10000 Dean's
    correlation gives Re_tau ~ 298 with a centreline u+ ~ 19.74."""
    assert abs(np.log(11.6) / 0.4 + 5.5 - 11.6) < 0.05
    retau = 0.09 * 10000.0 ** 0.88
    assert abs(retau - 298.02) < 0.01
    assert abs(np.log(retau) / 0.4 + 5.5 - 19.743) < 0.001
