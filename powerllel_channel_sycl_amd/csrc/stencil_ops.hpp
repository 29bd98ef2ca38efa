// stencil_ops.hpp -- the momentum tendency of the reference's RK stages as a
// device function over any stencil accessor, shared by the plain stage kernels
// (k_stencil.hip) and the closure stage kernel (k_closure.hip).
#pragma once

#include "pll_internal.hpp"

namespace pll {

// ---------------------------------------------------------------------------
// Momentum tendency F = visc - conv of the three staggered components at cell
// (i,j,k), src/mod_calcVel.cc:61-125 (stage 1) and :174-238 (stage 2).  Faces:
// u at (i+1/2, j, k), v at (i, j+1/2, k), w at (i, j, k+1/2).
// ---------------------------------------------------------------------------
template <typename T, class A>
__device__ __forceinline__ void tendency(const A& U, const A& V, const A& W, const MeshT<T>& m,
                                         int k, T dxi, T dyi, T rei, T ucrf, T& tu, T& tv,
                                         T& tw) {
  const T quarter = T(0.25), half = T(0.5);
  const T dzf_k = m.dzf[k], dzf_p = m.dzf[k + 1], dzf_m = m.dzf[k - 1];
  const T dci_k = m.dzc_inv[k], dci_m = m.dzc_inv[k - 1];
  const T dfi_k = m.dzf_inv[k], dfi_p = m.dzf_inv[k + 1];

  // ---- x-momentum at the u face ----
  {
    const T q0 = U(0, 0, 0), qe = U(1, 0, 0), qw = U(-1, 0, 0), qn = U(0, 1, 0);
    const T qs = U(0, -1, 0), qt = U(0, 0, 1), qb = U(0, 0, -1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = (q0 * dzf_p + qt * dzf_k) * dci_k;
    const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
    const T vn = V(0, 0, 0) + V(1, 0, 0);
    const T vs = V(0, -1, 0) + V(1, -1, 0);
    const T wt = W(0, 0, 0) + W(1, 0, 0);
    const T wb = W(0, 0, -1) + W(1, 0, -1);
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dci_k, gb = (q0 - qb) * dci_m;
    T conv = quarter * ((ae * ae - aw * aw) * dxi + (an * vn - as * vs) * dyi +
                        (at * wt - ab * wb) * dfi_k);
    conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dfi_k) * rei;
    tu = visc - conv;
  }
  // ---- y-momentum at the v face ----
  {
    const T q0 = V(0, 0, 0), qe = V(1, 0, 0), qw = V(-1, 0, 0), qn = V(0, 1, 0);
    const T qs = V(0, -1, 0), qt = V(0, 0, 1), qb = V(0, 0, -1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = (q0 * dzf_p + qt * dzf_k) * dci_k;
    const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
    const T ue = U(0, 0, 0) + U(0, 1, 0);
    const T uw = U(-1, 0, 0) + U(-1, 1, 0);
    const T wt = W(0, 0, 0) + W(0, 1, 0);
    const T wb = W(0, 0, -1) + W(0, 1, -1);
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dci_k, gb = (q0 - qb) * dci_m;
    T conv = quarter * ((ue * ae - uw * aw) * dxi + (an * an - as * as) * dyi +
                        (wt * at - wb * ab) * dfi_k);
    conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dfi_k) * rei;
    tv = visc - conv;
  }
  // ---- z-momentum at the w face ----
  {
    const T q0 = W(0, 0, 0), qe = W(1, 0, 0), qw = W(-1, 0, 0), qn = W(0, 1, 0);
    const T qs = W(0, -1, 0), qt = W(0, 0, 1), qb = W(0, 0, -1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = q0 + qt, ab = q0 + qb;
    const T ue = (U(0, 0, 0) * dzf_p + U(0, 0, 1) * dzf_k) * dci_k;
    const T uw = (U(-1, 0, 0) * dzf_p + U(-1, 0, 1) * dzf_k) * dci_k;
    const T vn = (V(0, 0, 0) * dzf_p + V(0, 0, 1) * dzf_k) * dci_k;
    const T vs = (V(0, -1, 0) * dzf_p + V(0, -1, 1) * dzf_k) * dci_k;
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dfi_p, gb = (q0 - qb) * dfi_k;
    T conv = quarter * ((ue * ae - uw * aw) * dxi + (vn * an - vs * as) * dyi +
                        (at * at - ab * ab) * dci_k);
    conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dci_k) * rei;
    tw = visc - conv;
  }
}

}  // namespace pll
