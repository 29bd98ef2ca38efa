"""CPU restatement of the learned closure (csrc/k_closure.hip): the
perceptron, its back propagation, and its coupling to the channel (features,
box filter, training labels, eddy-viscosity stress divergence).  The reference
has no closure; these functions restate this library's own definitions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels, never by the product path.

Pins (tests/test_closure_oracle.py): BF16 rounding against hand-worked values;
back propagation against finite differences; the stress divergence against the
Laplacian identity (constant nu, divergence-free field) and the exact
dissipation identity sum(u . div tau) = -sum(2 nu S:S) over cells and edges;
gradients against analytic fields at second order; the box filter against
exact divergence preservation; box stresses against numpy's weighted
covariance; labels against constructed stress-strain pairs.

10 -> 64 -> 64 -> 1, ReLU hidden layers, softplus output, loss = mean of
(softplus(o) - y)^2.  With bf16=True the inputs, weights, activations h1 and h2
and the back-propagated deltas are rounded to BF16 (nearest, ties to even) at
the points where the GPU rounds them before an MFMA product; sums are FP64.
With bf16=False everything is FP64: the form the finite-difference check pins.
"""
import numpy as np

N_IN, N_IN_PAD, N_H = 10, 16, 64
O_W1 = 0
O_B1 = O_W1 + N_IN_PAD * N_H
O_W2 = O_B1 + N_H
O_B2 = O_W2 + N_H * N_H
O_W3 = O_B2 + N_H
O_B3 = O_W3 + N_H
N_PARAM = O_B3 + 1


def bf16(a):
    """Round float32 values to BF16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[O_W1:O_B1].reshape(N_IN_PAD, N_H).copy()
    W1[N_IN:] = 0.0
    return dict(W1=W1, b1=p[O_B1:O_W2], W2=p[O_W2:O_B2].reshape(N_H, N_H), b2=p[O_B2:O_W3],
                W3=p[O_W3:O_B3], b3=p[O_B3])


def _r(a, on):
    return bf16(a).astype(np.float64) if on else np.asarray(a, dtype=np.float64)


def softplus(o):
    return np.where(o > 20, o, np.log1p(np.exp(np.minimum(o, 20))))


def forward(p, x, bf16_round=True):
    w = unpack(p)
    xb = np.zeros((x.shape[0], N_IN_PAD))
    xb[:, :N_IN] = _r(x, bf16_round)
    W1, W2, W3 = (_r(w[k], bf16_round) for k in ("W1", "W2", "W3"))
    h1 = np.maximum(xb @ W1 + w["b1"], 0.0)
    if bf16_round:
        h1 = h1.astype(np.float32).astype(np.float64)
    h1b = _r(h1, bf16_round)
    h2 = np.maximum(h1b @ W2 + w["b2"], 0.0)
    if bf16_round:
        h2 = h2.astype(np.float32).astype(np.float64)
    h2b = _r(h2, bf16_round)
    o = h2b @ W3 + w["b3"]
    return dict(xb=xb, W1=W1, W2=W2, W3=W3, h1=h1, h1b=h1b, h2=h2, h2b=h2b, o=o)


def predict(p, x, bf16_round=True):
    return softplus(forward(p, x, bf16_round)["o"])


def loss(p, x, y, bf16_round=True):
    f = forward(p, x, bf16_round)
    return float(np.mean((softplus(f["o"]) - y) ** 2))


def gradient(p, x, y, bf16_round=True):
    """(gradient over the parameter vector, loss), the back propagation of
    csrc/k_closure.hip step by step."""
    f = forward(p, x, bf16_round)
    n = x.shape[0]
    o = f["o"]
    d = softplus(o) - y
    g = 2.0 * d / n / (1.0 + np.exp(-o))
    grad = np.zeros(N_PARAM)
    grad[O_W3:O_B3] = f["h2b"].T @ g
    grad[O_B3] = g.sum()
    d2 = np.where(f["h2"] > 0, g[:, None] * f["W3"][None, :], 0.0)
    grad[O_B2:O_W3] = d2.sum(axis=0)
    d2b = _r(d2, bf16_round)
    grad[O_W2:O_B2] = (f["h1b"].T @ d2b).ravel()
    d1 = np.where(f["h1"] > 0, d2b @ f["W2"].T, 0.0)
    grad[O_B1:O_W2] = d1.sum(axis=0)
    d1b = _r(d1, bf16_round)
    gW1 = f["xb"].T @ d1b
    gW1[N_IN:] = 0.0
    grad[O_W1:O_B1] = gW1.ravel()
    return grad, float(np.mean(d ** 2))


# ---------------------------------------------------------------------------
# Coupling of the closure to the channel (csrc/k_closure.hip): features of a
# staggered field, the box filter, the training labels and the divergence of
# the eddy-viscosity stress.  Fields are halo-padded (nz+2, ny+2, nx+2) float64
# arrays in the reference layout with their halos already set.  Operation order
# follows the kernels term by term, so FP64 results agree to round-off.
# ---------------------------------------------------------------------------

def _s(a, di, dj, dk):
    """Interior view of a halo-padded array shifted by (di, dj, dk)."""
    nz, ny, nx = a.shape[0] - 2, a.shape[1] - 2, a.shape[2] - 2
    return a[1 + dk:nz + 1 + dk, 1 + dj:ny + 1 + dj, 1 + di:nx + 1 + di]


def feat_mesh(mesh, lz):
    """Per-plane hz = 1/(2 (dzc(k) + dzc(k-1))), wall distance and reflection
    sign of the cell centres k = 1..nz (mesh: Oracle.mesh())."""
    dzc, zc = mesh["dzc"], mesh["zc_global"]
    nz = len(dzc) - 2
    k = np.arange(1, nz + 1)
    hz = 0.5 / (dzc[k] + dzc[k - 1])
    yw = np.minimum(zc[k], lz - zc[k])
    zs = np.where(zc[k] < 0.5 * lz, 1.0, -1.0)
    return hz, yw, zs


def cell_gradients(u, v, w, dx, dy, mesh, hz):
    """The nine velocity gradients at the interior cell centres, (9, nz, ny, nx):
    du/dx, du/dy, du/dz, dv/dx, dv/dy, dv/dz, dw/dx, dw/dy, dw/dz."""
    dxi, dyi = 1.0 / dx, 1.0 / dy
    qdxi, qdyi = 0.25 * dxi, 0.25 * dyi
    H = hz[:, None, None]
    dfi = mesh["dzf_inv"][1:-1][:, None, None]
    S = _s
    return np.stack([
        (S(u, 0, 0, 0) - S(u, -1, 0, 0)) * dxi,
        ((S(u, 0, 1, 0) + S(u, -1, 1, 0)) - (S(u, 0, -1, 0) + S(u, -1, -1, 0))) * qdyi,
        ((S(u, 0, 0, 1) + S(u, -1, 0, 1)) - (S(u, 0, 0, -1) + S(u, -1, 0, -1))) * H,
        ((S(v, 1, 0, 0) + S(v, 1, -1, 0)) - (S(v, -1, 0, 0) + S(v, -1, -1, 0))) * qdxi,
        (S(v, 0, 0, 0) - S(v, 0, -1, 0)) * dyi,
        ((S(v, 0, 0, 1) + S(v, 0, -1, 1)) - (S(v, 0, 0, -1) + S(v, 0, -1, -1))) * H,
        ((S(w, 1, 0, 0) + S(w, 1, 0, -1)) - (S(w, -1, 0, 0) + S(w, -1, 0, -1))) * qdxi,
        ((S(w, 0, 1, 0) + S(w, 0, 1, -1)) - (S(w, 0, -1, 0) + S(w, 0, -1, -1))) * qdyi,
        (S(w, 0, 0, 0) - S(w, 0, 0, -1)) * dfi,
    ])


def features(g, gs, ys, yw, zs):
    """(nz*ny*nx, 10) float32 network inputs, cells in x-fastest order."""
    q = g * gs
    for c in (2, 5, 6, 7):
        q[c] = q[c] * zs[:, None, None]
    f = np.empty(g.shape[1:] + (N_IN,), dtype=np.float32)
    for c in range(9):
        f[..., c] = q[c].astype(np.float32)
    f[..., 9] = np.broadcast_to((yw * ys)[:, None, None], g.shape[1:]).astype(np.float32)
    return f.reshape(-1, N_IN)


def box_filter(uf, vf, wf, dzf_f, fx, fy, fz):
    """Interior (nzc, nyc, nxc) coarse u, v, w: area averages of the fine face
    values on each coarse face (dz-weighted in z)."""
    nzf, nyf, nxf = uf.shape[0] - 2, uf.shape[1] - 2, uf.shape[2] - 2
    nxc, nyc, nzc = nxf // fx, nyf // fy, nzf // fz
    dz = dzf_f[1:nzf + 1].reshape(nzc, fz)
    # u on the fine faces i = I*fx, v on j = J*fy, w on k = K*fz
    ui = uf[1:-1, 1:-1, fx:nxf + 1:fx].reshape(nzc, fz, nyc, fy, nxc)
    vi = vf[1:-1, fy:nyf + 1:fy, 1:-1].reshape(nzc, fz, nyc, nxc, fx)
    wi = wf[fz:nzf + 1:fz, 1:-1, 1:-1].reshape(nzc, nyc, fy, nxc, fx)
    dzs = dz.sum(axis=1)
    uc = (ui.sum(axis=3) * dz[:, :, None, None]).sum(axis=1) / (fy * dzs[:, None, None])
    vc = (vi.sum(axis=4) * dz[:, :, None, None]).sum(axis=1) / (fx * dzs[:, None, None])
    wc = wi.sum(axis=(2, 4)) / (fx * fy)
    return uc, vc, wc


def box_stress(uf, vf, wf, dzf_f, fx, fy, fz):
    """Subgrid stress of each coarse box, tau_ij = <c_i c_j> - <c_i><c_j>, with
    <.> the dz-weighted average over the box's fine cells of the fine
    cell-centre velocities c; dict of (nzc, nyc, nxc) arrays xx yy zz xy xz yz."""
    nzf, nyf, nxf = uf.shape[0] - 2, uf.shape[1] - 2, uf.shape[2] - 2
    nxc, nyc, nzc = nxf // fx, nyf // fy, nzf // fz
    cu = 0.5 * (uf[1:-1, 1:-1, :-2] + uf[1:-1, 1:-1, 1:-1])
    cv = 0.5 * (vf[1:-1, :-2, 1:-1] + vf[1:-1, 1:-1, 1:-1])
    cw = 0.5 * (wf[:-2, 1:-1, 1:-1] + wf[1:-1, 1:-1, 1:-1])
    wz = np.broadcast_to(dzf_f[1:nzf + 1][:, None, None], cu.shape)

    def box(a):
        return (a * wz).reshape(nzc, fz, nyc, fy, nxc, fx).sum(axis=(1, 3, 5))
    W = box(np.ones_like(cu))
    c = (cu, cv, cw)
    m = [box(a) / W for a in c]
    pairs = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}
    return {k: box(c[a] * c[b]) / W - m[a] * m[b] for k, (a, b) in pairs.items()}


def nu_star(t, g, gs, nus, eps_scaled=1e-4, ymax=50.0):
    """Label: max(0, -tau^d_ij S_ij) / (2 S_ij S_ij + eps_scaled/gs^2) / nus,
    at most ymax; S the symmetric part of the gradients g (9, ...)."""
    sxx, syy, szz = g[0], g[4], g[8]
    sxy, sxz, syz = 0.5 * (g[1] + g[3]), 0.5 * (g[2] + g[6]), 0.5 * (g[5] + g[7])
    ss = ((sxx * sxx + syy * syy) + szz * szz) + 2.0 * ((sxy * sxy + sxz * sxz) + syz * syz)
    ts = ((t["xx"] * sxx + t["yy"] * syy) + t["zz"] * szz) + \
        2.0 * ((t["xy"] * sxy + t["xz"] * sxz) + t["yz"] * syz) - \
        ((t["xx"] + t["yy"]) + t["zz"]) / 3.0 * ((sxx + syy) + szz)
    nu = np.maximum(0.0, -ts) / (2.0 * ss + eps_scaled / (gs * gs))
    return np.minimum(nu / nus, ymax)


def labels(uf, vf, wf, dzf_f, fx, fy, fz, g, gs, nus):
    """Training labels of the coarse cells, x-fastest float32."""
    return nu_star(box_stress(uf, vf, wf, dzf_f, fx, fy, fz), g, gs, nus).reshape(-1).astype(
        np.float32)


def sgs_divergence(nu, u, v, w, dx, dy, mesh):
    """Divergence of 2 nu S_ij at the u, v and w faces of the interior cells
    (each (nz, ny, nx)); nu: (nz, ny, nx) eddy viscosity at the cell centres.
    Shear stresses live on cell edges with nu averaged from the four cells
    around the edge; wall edges carry none; s_w is 0 on the top wall face."""
    nz, ny, nx = nu.shape
    dxi, dyi = 1.0 / dx, 1.0 / dy
    N0 = np.zeros((nz + 2, ny + 2, nx + 2))
    N0[1:-1] = np.pad(nu, ((0, 0), (1, 1), (1, 1)), mode="wrap")
    N = lambda di, dj, dk: _s(N0, di, dj, dk)
    U = lambda di, dj, dk: _s(u, di, dj, dk)
    V = lambda di, dj, dk: _s(v, di, dj, dk)
    W = lambda di, dj, dk: _s(w, di, dj, dk)
    dci = mesh["dzc_inv"][1:-1][:, None, None]
    dcm = mesh["dzc_inv"][:-2][:, None, None]
    dfi = mesh["dzf_inv"][1:-1][:, None, None]
    dfp = mesh["dzf_inv"][2:][:, None, None]
    kk = np.arange(1, nz + 1)[:, None, None]
    top, bot = kk == nz, kk == 1
    q = 0.25
    # u face
    sE = (U(1, 0, 0) - U(0, 0, 0)) * dxi
    sW = (U(0, 0, 0) - U(-1, 0, 0)) * dxi
    fxx = ((N(1, 0, 0) * sE - N(0, 0, 0) * sW) * 2.0) * dxi
    nN = q * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 1, 0)) + N(1, 1, 0))
    nS = q * (((N(0, -1, 0) + N(1, -1, 0)) + N(0, 0, 0)) + N(1, 0, 0))
    gN = (U(0, 1, 0) - U(0, 0, 0)) * dyi + (V(1, 0, 0) - V(0, 0, 0)) * dxi
    gS = (U(0, 0, 0) - U(0, -1, 0)) * dyi + (V(1, -1, 0) - V(0, -1, 0)) * dxi
    fxy = (nN * gN - nS * gS) * dyi
    nT = np.where(top, 0.0, q * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 0, 1)) + N(1, 0, 1)))
    nB = np.where(bot, 0.0, q * (((N(0, 0, -1) + N(1, 0, -1)) + N(0, 0, 0)) + N(1, 0, 0)))
    gT = (U(0, 0, 1) - U(0, 0, 0)) * dci + (W(1, 0, 0) - W(0, 0, 0)) * dxi
    gB = (U(0, 0, 0) - U(0, 0, -1)) * dcm + (W(1, 0, -1) - W(0, 0, -1)) * dxi
    fxz = (nT * gT - nB * gB) * dfi
    s_u = (fxx + fxy) + fxz
    # v face
    nE = q * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 1, 0)) + N(1, 1, 0))
    nW = q * (((N(-1, 0, 0) + N(0, 0, 0)) + N(-1, 1, 0)) + N(0, 1, 0))
    gE = (U(0, 1, 0) - U(0, 0, 0)) * dyi + (V(1, 0, 0) - V(0, 0, 0)) * dxi
    gW = (U(-1, 1, 0) - U(-1, 0, 0)) * dyi + (V(0, 0, 0) - V(-1, 0, 0)) * dxi
    fyx = (nE * gE - nW * gW) * dxi
    sN = (V(0, 1, 0) - V(0, 0, 0)) * dyi
    sS = (V(0, 0, 0) - V(0, -1, 0)) * dyi
    fyy = ((N(0, 1, 0) * sN - N(0, 0, 0) * sS) * 2.0) * dyi
    nT = np.where(top, 0.0, q * (((N(0, 0, 0) + N(0, 1, 0)) + N(0, 0, 1)) + N(0, 1, 1)))
    nB = np.where(bot, 0.0, q * (((N(0, 0, -1) + N(0, 1, -1)) + N(0, 0, 0)) + N(0, 1, 0)))
    gT = (V(0, 0, 1) - V(0, 0, 0)) * dci + (W(0, 1, 0) - W(0, 0, 0)) * dyi
    gB = (V(0, 0, 0) - V(0, 0, -1)) * dcm + (W(0, 1, -1) - W(0, 0, -1)) * dyi
    fyz = (nT * gT - nB * gB) * dfi
    s_v = (fyx + fyy) + fyz
    # w face
    nE = q * (((N(0, 0, 0) + N(1, 0, 0)) + N(0, 0, 1)) + N(1, 0, 1))
    nW = q * (((N(-1, 0, 0) + N(0, 0, 0)) + N(-1, 0, 1)) + N(0, 0, 1))
    gE = (U(0, 0, 1) - U(0, 0, 0)) * dci + (W(1, 0, 0) - W(0, 0, 0)) * dxi
    gW = (U(-1, 0, 1) - U(-1, 0, 0)) * dci + (W(0, 0, 0) - W(-1, 0, 0)) * dxi
    fzx = (nE * gE - nW * gW) * dxi
    nN = q * (((N(0, 0, 0) + N(0, 1, 0)) + N(0, 0, 1)) + N(0, 1, 1))
    nS = q * (((N(0, -1, 0) + N(0, 0, 0)) + N(0, -1, 1)) + N(0, 0, 1))
    gN = (V(0, 0, 1) - V(0, 0, 0)) * dci + (W(0, 1, 0) - W(0, 0, 0)) * dyi
    gS = (V(0, -1, 1) - V(0, -1, 0)) * dci + (W(0, 0, 0) - W(0, -1, 0)) * dyi
    fzy = (nN * gN - nS * gS) * dyi
    sT = (W(0, 0, 1) - W(0, 0, 0)) * dfp
    sB = (W(0, 0, 0) - W(0, 0, -1)) * dfi
    fzz = ((N(0, 0, 1) * sT - N(0, 0, 0) * sB) * 2.0) * dci
    s_w = np.where(top, 0.0, (fzx + fzy) + fzz)
    return s_u, s_v, s_w
