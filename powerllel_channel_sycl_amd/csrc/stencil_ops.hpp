// stencil_ops.hpp -- the momentum tendency of the reference's RK stages as a
// device function over any stencil accessor, shared by the plain stage kernels
// (k_stencil.hip) and the closure stage kernel (k_closure.hip).
#pragma once

#include "pll_internal.hpp"

namespace pll {

// The terms of the three components on the lower z face of a cell: the
// convective product and the z gradient of u (ab*wb, gb), of v (wb*ab, gb)
// and of w (ab*ab, gb).  They are the upper-face terms of the cell below,
// bit for bit (the same operands; the face sums differ only by the order of a
// commutative addition), so a thread marching up a column carries them from
// one plane to the next instead of computing them twice.
template <typename T>
struct ZFace {
  T cu, gu, cv, gv, cw, gw;
};

// ---------------------------------------------------------------------------
// Momentum tendency F = visc - conv of the three staggered components at cell
// (i,j,k), src/mod_calcVel.cc:61-125 (stage 1) and :174-238 (stage 2).  Faces:
// u at (i+1/2, j, k), v at (i, j+1/2, k), w at (i, j, k+1/2).  CRF adds the
// frame-shift term 0.5*u_crf*(ae - aw)*dxi (at u_crf = 0 it adds a zero and
// is compiled out).  With CARRY the lower-face terms come from zf, which then
// receives the upper-face terms; without it they are computed here.
// ---------------------------------------------------------------------------
template <typename T, bool CRF = true, bool CARRY = false, class A>
__device__ __forceinline__ void tendency_z(const A& U, const A& V, const A& W, const MeshT<T>& m,
                                           int k, T dxi, T dyi, T rei, T ucrf, T& tu, T& tv,
                                           T& tw, ZFace<T>& zf) {
  const T quarter = T(0.25), half = T(0.5);
  const T dzf_k = m.dzf[k], dzf_p = m.dzf[k + 1], dzf_m = m.dzf[k - 1];
  const T dci_k = m.dzc_inv[k], dci_m = m.dzc_inv[k - 1];
  const T dfi_k = m.dzf_inv[k], dfi_p = m.dzf_inv[k + 1];

  // ---- x-momentum at the u face ----
  {
    const T q0 = U(0, 0, 0), qe = U(1, 0, 0), qw = U(-1, 0, 0), qn = U(0, 1, 0);
    const T qs = U(0, -1, 0), qt = U(0, 0, 1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = (q0 * dzf_p + qt * dzf_k) * dci_k;
    const T vn = V(0, 0, 0) + V(1, 0, 0);
    const T vs = V(0, -1, 0) + V(1, -1, 0);
    const T wt = W(0, 0, 0) + W(1, 0, 0);
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dci_k;
    const T ct = at * wt;
    T cb, gb;
    if (CARRY) {
      cb = zf.cu;
      gb = zf.gu;
    } else {
      const T qb = U(0, 0, -1);
      const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
      const T wb = W(0, 0, -1) + W(1, 0, -1);
      cb = ab * wb;
      gb = (q0 - qb) * dci_m;
    }
    T conv = quarter * ((ae * ae - aw * aw) * dxi + (an * vn - as * vs) * dyi + (ct - cb) * dfi_k);
    if (CRF) conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dfi_k) * rei;
    tu = visc - conv;
    zf.cu = ct;
    zf.gu = gt;
  }
  // ---- y-momentum at the v face ----
  {
    const T q0 = V(0, 0, 0), qe = V(1, 0, 0), qw = V(-1, 0, 0), qn = V(0, 1, 0);
    const T qs = V(0, -1, 0), qt = V(0, 0, 1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = (q0 * dzf_p + qt * dzf_k) * dci_k;
    const T ue = U(0, 0, 0) + U(0, 1, 0);
    const T uw = U(-1, 0, 0) + U(-1, 1, 0);
    const T wt = W(0, 0, 0) + W(0, 1, 0);
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dci_k;
    const T ct = wt * at;
    T cb, gb;
    if (CARRY) {
      cb = zf.cv;
      gb = zf.gv;
    } else {
      const T qb = V(0, 0, -1);
      const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
      const T wb = W(0, 0, -1) + W(0, 1, -1);
      cb = wb * ab;
      gb = (q0 - qb) * dci_m;
    }
    T conv = quarter * ((ue * ae - uw * aw) * dxi + (an * an - as * as) * dyi + (ct - cb) * dfi_k);
    if (CRF) conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dfi_k) * rei;
    tv = visc - conv;
    zf.cv = ct;
    zf.gv = gt;
  }
  // ---- z-momentum at the w face ----
  {
    const T q0 = W(0, 0, 0), qe = W(1, 0, 0), qw = W(-1, 0, 0), qn = W(0, 1, 0);
    const T qs = W(0, -1, 0), qt = W(0, 0, 1);
    const T ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const T at = q0 + qt;
    const T ue = (U(0, 0, 0) * dzf_p + U(0, 0, 1) * dzf_k) * dci_k;
    const T uw = (U(-1, 0, 0) * dzf_p + U(-1, 0, 1) * dzf_k) * dci_k;
    const T vn = (V(0, 0, 0) * dzf_p + V(0, 0, 1) * dzf_k) * dci_k;
    const T vs = (V(0, -1, 0) * dzf_p + V(0, -1, 1) * dzf_k) * dci_k;
    const T ge = (qe - q0) * dxi, gw = (q0 - qw) * dxi;
    const T gn = (qn - q0) * dyi, gs = (q0 - qs) * dyi;
    const T gt = (qt - q0) * dfi_p;
    const T ct = at * at;
    T cb, gb;
    if (CARRY) {
      cb = zf.cw;
      gb = zf.gw;
    } else {
      const T qb = W(0, 0, -1);
      const T ab = q0 + qb;
      cb = ab * ab;
      gb = (q0 - qb) * dfi_k;
    }
    T conv = quarter * ((ue * ae - uw * aw) * dxi + (vn * an - vs * as) * dyi + (ct - cb) * dci_k);
    if (CRF) conv = conv + half * ucrf * (ae - aw) * dxi;
    const T visc = ((ge - gw) * dxi + (gn - gs) * dyi + (gt - gb) * dci_k) * rei;
    tw = visc - conv;
    zf.cw = ct;
    zf.gw = gt;
  }
}

// The tendency of one cell, every term computed here (the reference form).
template <typename T, class A>
__device__ __forceinline__ void tendency(const A& U, const A& V, const A& W, const MeshT<T>& m,
                                         int k, T dxi, T dyi, T rei, T ucrf, T& tu, T& tv,
                                         T& tw) {
  ZFace<T> zf;
  tendency_z<T, true, false>(U, V, W, m, k, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
}

// The lower-face terms of cell k alone (to start a carried march at plane k),
// the expressions tendency_z uses without CARRY.
template <typename T, class A>
__device__ __forceinline__ ZFace<T> zface_lower(const A& U, const A& V, const A& W,
                                                const MeshT<T>& m, int k) {
  const T dzf_k = m.dzf[k], dzf_m = m.dzf[k - 1];
  const T dci_m = m.dzc_inv[k - 1], dfi_k = m.dzf_inv[k];
  ZFace<T> zf;
  {
    const T q0 = U(0, 0, 0), qb = U(0, 0, -1);
    const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
    const T wb = W(0, 0, -1) + W(1, 0, -1);
    zf.cu = ab * wb;
    zf.gu = (q0 - qb) * dci_m;
  }
  {
    const T q0 = V(0, 0, 0), qb = V(0, 0, -1);
    const T ab = (q0 * dzf_m + qb * dzf_k) * dci_m;
    const T wb = W(0, 0, -1) + W(0, 1, -1);
    zf.cv = wb * ab;
    zf.gv = (q0 - qb) * dci_m;
  }
  {
    const T q0 = W(0, 0, 0), qb = W(0, 0, -1);
    const T ab = q0 + qb;
    zf.cw = ab * ab;
    zf.gw = (q0 - qb) * dfi_k;
  }
  return zf;
}

}  // namespace pll
