"""The closure perceptron's CPU restatement (oracle/pll_closure_oracle.py),
pinned without the GPU: BF16 rounding against hand-worked values, and the
back propagation against central finite differences of the loss in FP64."""
import numpy as np

from oracle import pll_closure_oracle as mo


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # BF16 spacing at 1 (8 significant bits)
    cases = {
        1.0: 1.0,
        1.0 + ulp: 1.0 + ulp,
        1.0 + ulp / 2: 1.0,               # tie -> even (mantissa ...0)
        1.0 + 3 * ulp / 2: 1.0 + 2 * ulp,  # tie -> even (mantissa ...10)
        1.0 + ulp / 2 + 2.0 ** -20: 1.0 + ulp,
        -2.5: -2.5,
        3.0e38: 226 / 128 * 2.0 ** 127,  # 3e38 = 1.7633 * 2^127 -> 226/128
    }
    for x, want in cases.items():
        got = float(mo.bf16(np.array([x], dtype=np.float32))[0])
        assert abs(got - want) <= 1e-3 * abs(want), (x, got, want)
    assert np.isinf(mo.bf16(np.array([np.inf], dtype=np.float32))[0])
    assert float(mo.bf16(np.array([one], dtype=np.float32))[0]) == 1.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mo.N_PARAM)
    p[mo.O_W1:mo.O_B1].reshape(16, 64)[10:] = 0.0
    x = rng.normal(0, 1, (37, 10))
    y = np.abs(rng.normal(1, 0.5, 37))
    g, _ = mo.gradient(p, x, y, bf16_round=False)
    # every parameter block, a few entries each
    idx = list(rng.choice(np.arange(mo.O_W1, mo.O_W1 + 10 * 64), 6)) + \
        list(rng.choice(np.arange(mo.O_B1, mo.O_W2), 4)) + \
        list(rng.choice(np.arange(mo.O_W2, mo.O_B2), 6)) + \
        list(rng.choice(np.arange(mo.O_B2, mo.O_W3), 4)) + \
        list(rng.choice(np.arange(mo.O_W3, mo.O_B3), 4)) + [mo.O_B3]
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = mo.loss(q, x, y, bf16_round=False)
        q[i] -= 2 * h
        lm = mo.loss(q, x, y, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # padded input rows carry no gradient
    assert not g[mo.O_W1 + 10 * 64:mo.O_B1].any()


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.3, mo.N_PARAM)
    x = rng.normal(0, 1, (200, 10))
    a = mo.predict(p, x, bf16_round=True)
    b = mo.predict(p, x, bf16_round=False)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(b))


# ---- coupling: stress divergence, gradients, filter, labels -----------------------
import math

from oracle import pll_oracle as po


def _mesh(nz, lz=2.0, stretch=1.5):
    dzf = po.mesh_by_func(1, 0, stretch, nz, lz) if stretch else np.full(nz + 2, lz / nz)
    dzf_inv = 1.0 / dzf
    dzc = np.zeros(nz + 2)
    dzc[:nz + 1] = 0.5 * (dzf[:nz + 1] + dzf[1:])
    dzc_inv = np.zeros(nz + 2)
    dzc_inv[:nz + 1] = 1.0 / dzc[:nz + 1]
    zc = np.zeros(nz + 2)
    zc[0] = -0.5 * dzf[0]
    for k in range(1, nz + 2):
        zc[k] = zc[k - 1] + 0.5 * (dzf[k - 1] + dzf[k])
    return dict(dzf=dzf, dzf_inv=dzf_inv, dzc=dzc, dzc_inv=dzc_inv, zc_global=zc)


def _halo_xy(a):
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]
    a[:, :, 0] = a[:, :, -2]
    a[:, :, -1] = a[:, :, 1]


def _curl_field(nx, ny, nz, dx, dy, mesh, seed):
    """A discretely divergence-free staggered field (the discrete curl of a
    random edge potential), w = 0 on the walls, periodic halos in x and y."""
    rng = np.random.default_rng(seed)
    Ax = rng.standard_normal((nz + 1, ny, nx))  # (i, j+1/2, z face)
    Ay = rng.standard_normal((nz + 1, ny, nx))  # (i+1/2, j, z face)
    Az = rng.standard_normal((nz, ny, nx))      # (i+1/2, j+1/2, k)
    Ax[0] = Ax[nz] = 0.0
    Ay[0] = Ay[nz] = 0.0
    dzf = mesh["dzf"][1:nz + 1][:, None, None]
    u = (Az - np.roll(Az, 1, axis=1)) / dy - (Ay[1:] - Ay[:-1]) / dzf
    v = (Ax[1:] - Ax[:-1]) / dzf - (Az - np.roll(Az, 1, axis=2)) / dx
    w = (Ay - np.roll(Ay, 1, axis=2)) / dx - (Ax - np.roll(Ax, 1, axis=1)) / dy
    U, V, W = (np.zeros((nz + 2, ny + 2, nx + 2)) for _ in range(3))
    U[1:-1, 1:-1, 1:-1], V[1:-1, 1:-1, 1:-1] = u, v
    W[0:nz + 1, 1:-1, 1:-1] = w
    for a in (U, V):
        a[0], a[-1] = -a[1], -a[-2]  # no-slip ghosts
    for a in (U, V, W):
        _halo_xy(a)
    return U, V, W


def _div(u, v, w, dx, dy, mesh):
    dzf_inv = mesh["dzf_inv"][1:-1][:, None, None]
    S = mo._s
    return (S(u, 0, 0, 0) - S(u, -1, 0, 0)) / dx + (S(v, 0, 0, 0) - S(v, 0, -1, 0)) / dy + \
        (S(w, 0, 0, 0) - S(w, 0, 0, -1)) * dzf_inv


def test_curl_field_is_divergence_free():
    nx, ny, nz = 12, 10, 9
    mesh = _mesh(nz)
    u, v, w = _curl_field(nx, ny, nz, 0.5, 0.3, mesh, 1)
    assert np.max(np.abs(_div(u, v, w, 0.5, 0.3, mesh))) < 1e-12 * np.max(np.abs(u)) / 0.3


def test_stress_divergence_constant_nu_is_laplacian():
    """2 div(c S) = c lap(u) + c grad(div u): for constant nu and a divergence-
    free field the stress divergence is c times the staggered Laplacian
    (faces not touching a wall edge)."""
    nx, ny, nz = 12, 10, 11
    dx, dy, c = 0.5, 0.3, 0.37
    mesh = _mesh(nz)
    u, v, w = _curl_field(nx, ny, nz, dx, dy, mesh, 2)
    su, sv, sw = mo.sgs_divergence(np.full((nz, ny, nx), c), u, v, w, dx, dy, mesh)
    S = mo._s
    dzf, dzc = mesh["dzf"], mesh["dzc"]

    def lap_uv(a):
        k = np.arange(1, nz + 1)[:, None, None]
        return (S(a, 1, 0, 0) - 2 * S(a, 0, 0, 0) + S(a, -1, 0, 0)) / dx**2 + \
            (S(a, 0, 1, 0) - 2 * S(a, 0, 0, 0) + S(a, 0, -1, 0)) / dy**2 + \
            ((S(a, 0, 0, 1) - S(a, 0, 0, 0)) / dzc[k] - (S(a, 0, 0, 0) - S(a, 0, 0, -1)) / dzc[k - 1]) / dzf[k]
    k = np.arange(1, nz + 1)[:, None, None]
    lw = (S(w, 1, 0, 0) - 2 * S(w, 0, 0, 0) + S(w, -1, 0, 0)) / dx**2 + \
        (S(w, 0, 1, 0) - 2 * S(w, 0, 0, 0) + S(w, 0, -1, 0)) / dy**2 + \
        ((S(w, 0, 0, 1) - S(w, 0, 0, 0)) / dzf[k + 1] - (S(w, 0, 0, 0) - S(w, 0, 0, -1)) / dzf[k]) / dzc[k]
    for s_, lap, ks in ((su, lap_uv(u), slice(1, nz - 1)), (sv, lap_uv(v), slice(1, nz - 1)),
                        (sw, lw, slice(0, nz - 1))):
        ref = c * lap[ks]
        assert np.max(np.abs(s_[ks] - ref)) < 1e-11 * np.max(np.abs(ref))


def test_stress_divergence_dissipation_identity():
    """Summation by parts, exact for the staggered operators: the work of the
    stress divergence on the velocity is minus the dissipation 2 nu S:S summed
    over cells (normal strains) and edges (shear strains, no wall edges)."""
    nx, ny, nz = 10, 8, 7
    dx, dy = 0.4, 0.35
    mesh = _mesh(nz)
    rng = np.random.default_rng(5)
    u, v, w = (rng.standard_normal((nz + 2, ny + 2, nx + 2)) for _ in range(3))
    w[0] = 0.0
    w[nz] = 0.0
    for a in (u, v, w):
        _halo_xy(a)
    nu = rng.uniform(0.0, 2.0, (nz, ny, nx))
    su, sv, sw = mo.sgs_divergence(nu, u, v, w, dx, dy, mesh)
    dzf, dzc = mesh["dzf"][1:nz + 1], mesh["dzc"][1:nz + 1]
    S = mo._s
    work = dx * dy * (np.sum((S(u, 0, 0, 0) * su + S(v, 0, 0, 0) * sv) * dzf[:, None, None]) +
                      np.sum((S(w, 0, 0, 0) * sw)[:nz - 1] * dzc[:nz - 1, None, None]))
    # cells
    exx = (S(u, 0, 0, 0) - S(u, -1, 0, 0)) / dx
    eyy = (S(v, 0, 0, 0) - S(v, 0, -1, 0)) / dy
    ezz = (S(w, 0, 0, 0) - S(w, 0, 0, -1)) / dzf[:, None, None]
    diss = np.sum(2 * nu * (exx**2 + eyy**2 + ezz**2) * dzf[:, None, None]) * dx * dy
    # xy edges (i+1/2, j+1/2, k)
    ne = 0.25 * (nu + np.roll(nu, -1, 2) + np.roll(nu, -1, 1) + np.roll(np.roll(nu, -1, 1), -1, 2))
    gxy = (S(u, 0, 1, 0) - S(u, 0, 0, 0)) / dy + (S(v, 1, 0, 0) - S(v, 0, 0, 0)) / dx
    diss += np.sum(ne * gxy**2 * dzf[:, None, None]) * dx * dy
    # xz edges (i+1/2, j, k+1/2) and yz edges (i, j+1/2, k+1/2), k = 1..nz-1
    dzc_in = mesh["dzc"][1:nz][:, None, None]
    nxz = 0.25 * (nu[:-1] + np.roll(nu[:-1], -1, 2) + nu[1:] + np.roll(nu[1:], -1, 2))
    gxz = ((S(u, 0, 0, 1) - S(u, 0, 0, 0))[:-1] / dzc_in + (S(w, 1, 0, 0) - S(w, 0, 0, 0))[:-1] / dx)
    nyz = 0.25 * (nu[:-1] + np.roll(nu[:-1], -1, 1) + nu[1:] + np.roll(nu[1:], -1, 1))
    gyz = ((S(v, 0, 0, 1) - S(v, 0, 0, 0))[:-1] / dzc_in + (S(w, 0, 1, 0) - S(w, 0, 0, 0))[:-1] / dy)
    diss += np.sum((nxz * gxz**2 + nyz * gyz**2) * dzc_in) * dx * dy
    assert diss > 0
    assert abs(work + diss) < 1e-12 * diss


def _analytic(nx, ny, nz, lx, ly, mesh):
    """Smooth fields sampled at their staggered points (halos included) and
    their nine exact gradients at the cell centres."""
    dx, dy = lx / nx, ly / ny
    kx, ky = 2 * math.pi / lx, 2 * math.pi / ly
    xc = (np.arange(nx + 2) - 0.5) * dx
    yc = (np.arange(ny + 2) - 0.5) * dy
    xf, yf = xc + 0.5 * dx, yc + 0.5 * dy
    zc = mesh["zc_global"]
    zf = np.concatenate([[0.0], np.cumsum(mesh["dzf"][1:nz + 1]), [0.0]])
    zf[nz + 1] = zf[nz] + mesh["dzf"][nz + 1]
    Z, Y, X = np.meshgrid(zc, yc, xf, indexing="ij")
    u = np.sin(kx * X + 0.3) * np.cos(ky * Y) * Z * (2 - Z)
    Z, Y, X = np.meshgrid(zc, yf, xc, indexing="ij")
    v = np.cos(kx * X) * np.sin(ky * Y + 0.5) * (1 + Z**2)
    Z, Y, X = np.meshgrid(zf, yc, xc, indexing="ij")
    w = np.sin(kx * X + 1) * np.sin(ky * Y + 0.2) * np.sin(0.5 * math.pi * Z)
    Z, Y, X = np.meshgrid(zc[1:-1], yc[1:-1], xc[1:-1], indexing="ij")
    a, b = kx * X, ky * Y
    g = np.stack([
        kx * np.cos(a + 0.3) * np.cos(b) * Z * (2 - Z),
        -ky * np.sin(a + 0.3) * np.sin(b) * Z * (2 - Z),
        np.sin(a + 0.3) * np.cos(b) * (2 - 2 * Z),
        -kx * np.sin(a) * np.sin(b + 0.5) * (1 + Z**2),
        ky * np.cos(a) * np.cos(b + 0.5) * (1 + Z**2),
        np.cos(a) * np.sin(b + 0.5) * 2 * Z,
        kx * np.cos(a + 1) * np.sin(b + 0.2) * np.sin(0.5 * math.pi * Z),
        ky * np.sin(a + 1) * np.cos(b + 0.2) * np.sin(0.5 * math.pi * Z),
        np.sin(a + 1) * np.sin(b + 0.2) * 0.5 * math.pi * np.cos(0.5 * math.pi * Z),
    ])
    return u, v, w, g


def test_cell_gradients_second_order():
    errs = []
    for n in (16, 32):
        nx, ny, nz = n, n, n
        lx, ly = 2 * math.pi, math.pi
        mesh = _mesh(nz, stretch=0.0)
        u, v, w, g = _analytic(nx, ny, nz, lx, ly, mesh)
        hz, _, _ = mo.feat_mesh(mesh, 2.0)
        got = mo.cell_gradients(u, v, w, lx / nx, ly / ny, mesh, hz)
        errs.append([np.max(np.abs(got[c] - g[c])) / np.max(np.abs(g[c])) for c in range(9)])
    errs = np.array(errs)
    assert np.all(errs[0] < 0.05), errs[0]
    assert np.all(errs[0] / errs[1] > 3.0), errs[0] / errs[1]  # second order


def test_features_mirror_symmetric():
    """The reflection z -> lz - z (w -> -w) maps the features of cell k to
    those of cell nz+1-k: the sign-odd gradients are flipped in the upper half."""
    nx, ny, nz = 8, 6, 10
    mesh = _mesh(nz)
    rng = np.random.default_rng(9)
    u, v, w = (rng.standard_normal((nz + 2, ny + 2, nx + 2)) for _ in range(3))
    w[0] = w[nz] = 0.0
    for a in (u, v, w):
        _halo_xy(a)
    um, vm = u[::-1].copy(), v[::-1].copy()
    wm = np.zeros_like(w)
    wm[0:nz + 1] = -w[nz::-1]
    hz, yw, zs = mo.feat_mesh(mesh, 2.0)
    f = mo.features(mo.cell_gradients(u, v, w, 0.3, 0.2, mesh, hz), 0.1, 2.0, yw, zs)
    fm = mo.features(mo.cell_gradients(um, vm, wm, 0.3, 0.2, mesh, hz), 0.1, 2.0, yw, zs)
    f = f.reshape(nz, ny, nx, 10)
    fm = fm.reshape(nz, ny, nx, 10)[::-1]
    np.testing.assert_allclose(f, fm, rtol=1e-5, atol=1e-6)


def test_box_filter_keeps_divergence_free():
    fx, fy, fz = 2, 3, 3
    nxc, nyc, nzc = 5, 4, 6
    mesh_f, mesh_c = _mesh(nzc * fz), _mesh(nzc)
    dxc, dyc = 0.4, 0.3
    u, v, w = _curl_field(nxc * fx, nyc * fy, nzc * fz, dxc / fx, dyc / fy, mesh_f, 3)
    uc, vc, wc = mo.box_filter(u, v, w, mesh_f["dzf"], fx, fy, fz)
    U, V, W = (np.zeros((nzc + 2, nyc + 2, nxc + 2)) for _ in range(3))
    U[1:-1, 1:-1, 1:-1], V[1:-1, 1:-1, 1:-1], W[1:-1, 1:-1, 1:-1] = uc, vc, wc
    for a in (U, V, W):
        _halo_xy(a)
    np.testing.assert_allclose(mesh_c["dzf"][1:-1], mesh_f["dzf"][1:-1].reshape(nzc, fz).sum(1),
                               rtol=1e-13)
    assert np.max(np.abs(wc[-1])) < 1e-12  # the top wall
    d = _div(U, V, W, dxc, dyc, mesh_c)
    assert np.max(np.abs(d)) < 1e-12 * np.max(np.abs(uc)) / dyc


def test_box_stress_is_weighted_covariance():
    fx, fy, fz = 2, 2, 3
    nxc, nyc, nzc = 3, 2, 2
    nxf, nyf, nzf = nxc * fx, nyc * fy, nzc * fz
    mesh = _mesh(nzf)
    rng = np.random.default_rng(4)
    u, v, w = (rng.standard_normal((nzf + 2, nyf + 2, nxf + 2)) for _ in range(3))
    t = mo.box_stress(u, v, w, mesh["dzf"], fx, fy, fz)
    cu = 0.5 * (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 1:-1])
    cv = 0.5 * (v[1:-1, :-2, 1:-1] + v[1:-1, 1:-1, 1:-1])
    cw = 0.5 * (w[:-2, 1:-1, 1:-1] + w[1:-1, 1:-1, 1:-1])
    for K in range(nzc):
        for J in range(nyc):
            for I in range(nxc):
                sl = (slice(K * fz, (K + 1) * fz), slice(J * fy, (J + 1) * fy), slice(I * fx, (I + 1) * fx))
                wts = np.broadcast_to(mesh["dzf"][1 + K * fz:1 + (K + 1) * fz, None, None],
                                      (fz, fy, fx)).ravel()
                cov = np.cov(np.stack([cu[sl].ravel(), cv[sl].ravel(), cw[sl].ravel()]),
                             aweights=wts, bias=True)
                for name, (a, b) in {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1),
                                     "xz": (0, 2), "yz": (1, 2)}.items():
                    assert abs(t[name][K, J, I] - cov[a, b]) < 1e-12 * (1 + abs(cov[a, b]))


def test_nu_star_of_constructed_stresses():
    rng = np.random.default_rng(8)
    g = rng.standard_normal((9, 50))
    tr = (g[0] + g[4] + g[8]) / 3
    g[0] -= tr
    g[4] -= tr
    g[8] -= tr
    S = {"xx": g[0], "yy": g[4], "zz": g[8], "xy": 0.5 * (g[1] + g[3]), "xz": 0.5 * (g[2] + g[6]),
         "yz": 0.5 * (g[5] + g[7])}
    ss = sum(S[k]**2 for k in ("xx", "yy", "zz")) + 2 * sum(S[k]**2 for k in ("xy", "xz", "yz"))
    nu0, p, gs, nus = 0.013, 0.7, 2.0, 0.001
    eps = 1e-4 / gs**2
    t = {k: -2 * nu0 * S[k] + (p if k in ("xx", "yy", "zz") else 0.0) for k in S}
    got = mo.nu_star(t, g, gs, nus, ymax=1e9)
    np.testing.assert_allclose(got, nu0 * 2 * ss / (2 * ss + eps) / nus, rtol=1e-12)
    t = {k: 2 * nu0 * S[k] for k in S}  # backscatter is clipped
    assert not mo.nu_star(t, g, gs, nus).any()
    assert np.all(mo.nu_star({k: -1e3 * S[k] for k in S}, g, gs, nus) == 50.0)  # the cap
