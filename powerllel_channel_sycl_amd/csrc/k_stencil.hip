// k_stencil.hip -- momentum RK stages, Poisson RHS, projection, forcing, the
// boundary/halo update, the bulk-velocity reduction, the initial condition and
// the runtime checks.  All kernels are templated on the field type T (double
// for the FP64 path, float for the FP32-mixed path).
//
// FP64 parity: the library is compiled with -ffp-contract=off and every
// expression below keeps the association of the reference formula it cites,
// so an FP64 launch reproduces the reference arithmetic bit for bit.
#include "pll_internal.hpp"
#include "stencil_ops.hpp"

#include <cmath>
#include <type_traits>
#include <utility>
#include <vector>

namespace pll {

namespace {

constexpr int kTx = 64;  // threads along x (one wavefront)
constexpr int kTy = 4;   // rows along y per block
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// Stencil value (di, dj, dk) around the thread's cell, read from an LDS ring
// of three tile planes (k-1, k, k+1) of width SX.
template <typename T, int SX>
struct PlaneAcc {
  const T* m;
  const T* z;
  const T* p;
  int c;
  __device__ __forceinline__ T operator()(int di, int dj, int dk) const {
    const T* pl = dk < 0 ? m : (dk > 0 ? p : z);
    return pl[c + di + dj * SX];
  }
};

// FP32-mixed form of the same operator, scaled by S (dt in stage 1, dt/2 in
// stage 2): algebraically equal to S*(visc - conv) above, with the mesh factors
// gathered into per-plane coefficients (RkCoef) and the viscous terms written
// as second differences.  It is not bit-identical to the reference order and
// is used only by the FP32-mixed path, whose error is measured against the
// FP64 oracle.
struct RkCoef {
  float a1, a2, b1, b2;  // (u,v) face interpolation to k+1/2 and k-1/2
  float zuv, cp, cm;     // S*0.25*dzf_inv(k); S*re_inv*dzc_inv(k|k-1)*dzf_inv(k)
  float zw, dp, dm;      // S*0.25*dzc_inv(k); S*re_inv*dzf_inv(k+1|k)*dzc_inv(k)
  float cdu, cdw;        // diagonal of the viscous operator: 2 vx + 2 vy + cp + cm (u, v), + dp + dm (w)
};
struct RkUniform {
  float cx, cy, vx, vy, crf;  // S*0.25*dxi, S*0.25*dyi, S*re_inv*dxi^2, S*re_inv*dyi^2, S*0.5*u_crf*dxi
};

// The lower z-face convective products of the three components (ab*wb,
// wb*ab, ab*ab): the upper-face products of the cell below up to FP32
// rounding (the face interpolation coefficients of plane k-1's upper face are
// plane k's lower ones), carried up a column like the FP64 ZFace.
struct ZConv {
  float u, v, w;
};

template <class A>
__device__ __forceinline__ ZConv zconv_lower(const A& U, const A& V, const A& W, const RkCoef& q) {
#pragma clang fp contract(fast)
  ZConv z;
  const float ub = U(0, 0, 0) * q.b1 + U(0, 0, -1) * q.b2;
  z.u = ub * (W(0, 0, -1) + W(1, 0, -1));
  const float vb = V(0, 0, 0) * q.b1 + V(0, 0, -1) * q.b2;
  z.v = (W(0, 0, -1) + W(0, 1, -1)) * vb;
  const float wb = W(0, 0, 0) + W(0, 0, -1);
  z.w = wb * wb;
  return z;
}

// The upper z-face convective products of the cell at the accessor's plane
// (at*wt, wt*at, at*at), the expressions of tendency_fast: what its CARRY form
// hands to the plane above.
template <class A>
__device__ __forceinline__ ZConv zconv_upper(const A& U, const A& V, const A& W, const RkCoef& q) {
#pragma clang fp contract(fast)
  ZConv z;
  {
    const float at = U(0, 0, 0) * q.a1 + U(0, 0, 1) * q.a2;
    const float wt = W(0, 0, 0) + W(1, 0, 0);
    z.u = at * wt;
  }
  {
    const float at = V(0, 0, 0) * q.a1 + V(0, 0, 1) * q.a2;
    const float wt = W(0, 0, 0) + W(0, 1, 0);
    z.v = wt * at;
  }
  {
    const float at = W(0, 0, 0) + W(0, 0, 1);
    z.w = at * at;
  }
  return z;
}

// CARRY: the lower-face products come from zc, which then receives the upper ones.
template <bool CRF, bool CARRY = false, class A>
__device__ __forceinline__ void tendency_fast(const A& U, const A& V, const A& W, const RkCoef& q,
                                              const RkUniform& cu, float& iu, float& iv,
                                              float& iw, ZConv* zc = nullptr) {
  // the FP32 form is not held to the reference's rounding: let products fuse into FMAs
#pragma clang fp contract(fast)
  {
    const float q0 = U(0, 0, 0), qe = U(1, 0, 0), qw = U(-1, 0, 0), qn = U(0, 1, 0);
    const float qs = U(0, -1, 0), qt = U(0, 0, 1), qb = U(0, 0, -1);
    const float ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const float at = q0 * q.a1 + qt * q.a2;
    const float vn = V(0, 0, 0) + V(1, 0, 0), vs = V(0, -1, 0) + V(1, -1, 0);
    const float wt = W(0, 0, 0) + W(1, 0, 0);
    const float ct = at * wt;
    const float cb = CARRY ? zc->u : (q0 * q.b1 + qb * q.b2) * (W(0, 0, -1) + W(1, 0, -1));
    if (CARRY) zc->u = ct;
    float conv = cu.cx * (ae * ae - aw * aw) + cu.cy * (an * vn - as * vs) + q.zuv * (ct - cb);
    if (CRF) conv += cu.crf * (ae - aw);
    // viscous part as neighbour sums minus the diagonal: fewer operations
    iu = cu.vx * (qe + qw) + cu.vy * (qn + qs) + q.cp * qt + q.cm * qb - q.cdu * q0 - conv;
  }
  {
    const float q0 = V(0, 0, 0), qe = V(1, 0, 0), qw = V(-1, 0, 0), qn = V(0, 1, 0);
    const float qs = V(0, -1, 0), qt = V(0, 0, 1), qb = V(0, 0, -1);
    const float ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const float at = q0 * q.a1 + qt * q.a2;
    const float ue = U(0, 0, 0) + U(0, 1, 0), uw = U(-1, 0, 0) + U(-1, 1, 0);
    const float wt = W(0, 0, 0) + W(0, 1, 0);
    const float ct = wt * at;
    const float cb = CARRY ? zc->v : (W(0, 0, -1) + W(0, 1, -1)) * (q0 * q.b1 + qb * q.b2);
    if (CARRY) zc->v = ct;
    float conv = cu.cx * (ue * ae - uw * aw) + cu.cy * (an * an - as * as) + q.zuv * (ct - cb);
    if (CRF) conv += cu.crf * (ae - aw);
    iv = cu.vx * (qe + qw) + cu.vy * (qn + qs) + q.cp * qt + q.cm * qb - q.cdu * q0 - conv;
  }
  {
    const float q0 = W(0, 0, 0), qe = W(1, 0, 0), qw = W(-1, 0, 0), qn = W(0, 1, 0);
    const float qs = W(0, -1, 0), qt = W(0, 0, 1), qb = W(0, 0, -1);
    const float ae = q0 + qe, aw = qw + q0, an = q0 + qn, as = qs + q0;
    const float at = q0 + qt;
    const float ue = U(0, 0, 0) * q.a1 + U(0, 0, 1) * q.a2;
    const float uw = U(-1, 0, 0) * q.a1 + U(-1, 0, 1) * q.a2;
    const float vn = V(0, 0, 0) * q.a1 + V(0, 0, 1) * q.a2;
    const float vs = V(0, -1, 0) * q.a1 + V(0, -1, 1) * q.a2;
    const float ct = at * at;
    float cb;
    if (CARRY) {
      cb = zc->w;
      zc->w = ct;
    } else {
      const float ab = q0 + qb;
      cb = ab * ab;
    }
    float conv = cu.cx * (ue * ae - uw * aw) + cu.cy * (vn * an - vs * as) + q.zw * (ct - cb);
    if (CRF) conv += cu.crf * (ae - aw);
    iw = cu.vx * (qe + qw) + cu.vy * (qn + qs) + q.dp * qt + q.dm * qb - q.cdw * q0 - conv;
  }
}


constexpr int kRkTx = 64;  // tile width along x (one wavefront per row)
constexpr int kRkTy = 8;   // tile rows along y

// Tile rows of the stage and projection kernels: 4 for the FP64 projection
// (4-7 % faster than 8 at 768x384x384 on MI355X), 8 elsewhere (the FP32 stages
// lose 10-28 % with 4 and 28 % with 16; the FP64 stage 1, which applies the
// deferred projection on load, is 4 % faster with 8 than with 4).  The fused FP64 stage 2 +
// calcRHS takes 12 rows where that still gives kRhsTallBlocks blocks (fewer
// y-halo rows per output row than 8) and 8 on smaller grids, where taller tiles
// would leave too few blocks (4 % slower at 96x64x64).  12 rows (768 threads,
// up to 168 VGPRs, one block per CU) rather than 16: at 1024 threads its
// deferred-projection variant is held to 128 VGPRs and spills.
constexpr int kRhsTallTy = 12;
constexpr long kRhsTallBlocks = 1024;
template <typename T>
constexpr int correct_ty() {
  return std::is_same<T, double>::value ? 4 : kRkTy;
}

// Stage 1: d = s + dt*F(s).  Stage 2: d = (s + dt*F(s) + d)/2.
// A block owns a 64 x 8 (i,j) tile and marches a chunk of k planes.  The three
// source fields of planes k-1, k, k+1 (tile plus a one-cell halo) sit in an
// LDS ring; plane k+2 is fetched into registers while plane k is computed and
// replaces plane k-1 afterwards, so every source value leaves HBM once per block.
//
// Inside a multi-step run the previous step's forcing and final boundary
// update are folded into the loads (`fold`): `force` (device scalars, one per
// forced component) is added to every source value, and every halo cell is
// rebuilt from interior cells by the reference's closed-form rules -- x/y
// halos are periodic images, the z ghost planes are 2*bc - (u + f) for u,
// -(v + f) for v and 0 for w, and w on plane nz is 0
// (src/mod_updateBound.cc:11-144) -- so the result is bit-identical to forcing,
// updating the boundary and then running the stage.  Stage 2 adds the
// deferred forcing to the point value of u it reads.
//
// With `corr` the previous step's projection is deferred as well: the stored
// velocity is the predictor u* and every value read becomes
// u* - dt*grad(p) (src/mod_calcVel.cc:272-274) before the force and the
// boundary rules apply; p carries valid halos (updateBoundP ran).  With `acc`
// stage 2 also sums its new u times dzflzi(k) in FP64 into partial[block]: the
// bulk mean of the next projected velocity, since the projection changes every
// x-row's sum only by p(nx+1) - p(1) = 0.
// k_rk is the FP32-mixed form (tendency_fast, per-plane coefficients);
// k_rk_late below is the FP64 form in the reference's arithmetic.  PROJ
// compiles the deferred projection in; stage 1 launches the PROJ variant when
// the projection is deferred.
template <typename T, int STAGE, bool PROJ = true>
__global__ __launch_bounds__(kRkTx * kRkTy) void k_rk(Geom g, const T* __restrict__ su,
                                                     const T* __restrict__ sv,
                                                     const T* __restrict__ sw, T* __restrict__ du,
                                                     T* __restrict__ dv, T* __restrict__ dw,
                                                     MeshT<T> m, T dxi, T dyi, T rei, T dt,
                                                     T ucrf, int kchunk, int ntx, int nty,
                                                     int fold, int defer, T bc_u,
                                                     const double* __restrict__ force,
                                                     int forced_mask,
                                                     const RkCoef* __restrict__ coef,
                                                     RkUniform cu, int crf,
                                                     const T* __restrict__ pp, int corr,
                                                     T dtdxi, T dtdyi, int acc,
                                                     const double* __restrict__ dzflzi,
                                                     double* __restrict__ partial, int bcload) {
  constexpr int TX = kRkTx, TY = kRkTy, SX = TX + 2, SY = TY + 2, PL = SX * SY,
                NT = TX * TY;
  constexpr int NPF = (PL + NT - 1) / NT;
  static_assert(std::is_same<T, float>::value, "the FP64 stages run k_rk_late");
  corr = PROJ ? corr : 0;
  // halo cells from the boundary rules on load: with the folded forcing, or
  // alone (bcload) for a source whose boundary update was not run
  bcload = (fold || bcload) ? 1 : 0;
  __shared__ T sh[3][3][PL];  // [field][ring slot][tile cell]
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  const int i0 = ti * TX, j0 = tj * TY;  // tile cell (0,0) is the halo corner (i0, j0)
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool active = i <= g.nx && j <= g.ny;
  // deferred forcing of the three components (applied only where forced):
  // to every loaded value when folding, to the point value of stage 2 when deferring
  const bool any = fold || defer;
  const bool fu = any && (forced_mask & 1), fv = any && (forced_mask & 2),
             fw = any && (forced_mask & 4);
  const T f_u = fu ? (T)force[0] : T(0), f_v = fv ? (T)force[1] : T(0),
          f_w = fw ? (T)force[2] : T(0);

  long gaddr[NPF];  // plane-relative source address of this thread's tile cells (-1: outside)
  int cell[NPF];
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int e = threadIdx.x + n * NT;
    int gx = i0 + e % SX, gy = j0 + e / SX;
    cell[n] = e;
    const bool inside = e < PL && gx <= g.nx + 1 && gy <= g.ny + 1;
    if (bcload) {  // halo cells read their periodic image
      gx = gx == 0 ? g.nx : (gx == g.nx + 1 ? 1 : gx);
      gy = gy == 0 ? g.ny : (gy == g.ny + 1 ? 1 : gy);
    }
    gaddr[n] = inside ? gx + g.sx * gy : -1;
  }
  T pre[3][NPF];
  T pabove[NPF];      // p of plane ks+1 from the previous fetch (corr)
  int pabove_k = -1;  // the plane pabove belongs to
  auto fetch = [&](int kk) {
    const bool ghost = bcload && (kk == 0 || kk == g.nz + 1);
    const int ks = ghost ? (kk == 0 ? 1 : g.nz) : kk;
    const bool wzero = bcload && (kk == 0 || kk >= g.nz);
    // the stage-1 source is the stored velocity; stage 2's source u1 is not projected
    const bool pj = STAGE == 1 && corr;
    const T dzci = pj ? m.dzc_inv[ks] : T(0);
    const bool reuse = pabove_k == ks;
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        const long a = gaddr[n] + g.sxy * ks;
        T qu = su[a], qv = sv[a], qw = wzero ? T(0) : sw[a];
        if (pj) {
          const T p0 = reuse ? pabove[n] : pp[a];
          pabove[n] = pp[a + g.sxy];
          qu = qu - (pp[a + 1] - p0) * dtdxi;
          qv = qv - (pp[a + g.sx] - p0) * dtdyi;
          // same association as k_correct: ((p(k+1) - p(k)) * dt) * dzc_inv(k)
          if (!wzero) qw = qw - (pabove[n] - p0) * dt * dzci;
        }
        if (fold) {
          if (fu) qu = qu + f_u;
          if (fv) qv = qv + f_v;
          if (fw && !wzero) qw = qw + f_w;
        }
        if (ghost) {
          qu = T(2) * bc_u - qu;  // src/mod_updateBound.cc:92,123
          qv = T(2) * T(0) - qv;
        }
        pre[0][n] = qu;
        pre[1][n] = qv;
        pre[2][n] = qw;
      }
    pabove_k = pj ? ks + 1 : -1;
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        sh[0][slot][cell[n]] = pre[0][n];
        sh[1][slot][cell[n]] = pre[1][n];
        sh[2][slot][cell[n]] = pre[2][n];
      }
  };
  if (k0 > k1) {
    if (acc && threadIdx.x == 0) partial[b] = 0.0;
    return;
  }
  double msum = 0.0;  // FP64 partial of the bulk mean (acc)
  fetch(k0 - 1);
  commit(0);
  fetch(k0);
  commit(1);
  fetch(k0 + 1);
  commit(2);
  __syncthreads();
  int sm = 0, s0 = 1, sp = 2;
  const int c = (ty + 1) * SX + tx + 1;
  // stage 2's point values and, with corr, the pressures around the point
  // (fetching them a plane ahead costs more in occupancy than it saves)
  struct Point {
    T u, v, w, p0, pe, pn, pt;
  };
  auto load_point = [&](int kk, Point& q) {
    const long a = g.idx(i, j, kk);
    q.u = du[a];
    q.v = dv[a];
    q.w = dw[a];
    if (corr) {
      q.p0 = pp[a];
      q.pe = pp[a + 1];
      q.pn = pp[a + g.sx];
      q.pt = pp[a + g.sxy];
    }
  };
  Point qc{};
  ZConv zc{};  // the carried lower-face convective products of the thread's column
  if (active) {
    const PlaneAcc<T, SX> U{sh[0][sm], sh[0][s0], sh[0][sp], c};
    const PlaneAcc<T, SX> V{sh[1][sm], sh[1][s0], sh[1][sp], c};
    const PlaneAcc<T, SX> W{sh[2][sm], sh[2][s0], sh[2][sp], c};
    zc = zconv_lower(U, V, W, coef[k0]);
  }
  for (int k = k0; k <= k1; ++k) {
    const bool more = k + 2 <= k1 + 1;
    if (more) fetch(k + 2);
    if (STAGE == 2 && active) load_point(k, qc);
    if (active) {
      const PlaneAcc<T, SX> U{sh[0][sm], sh[0][s0], sh[0][sp], c};
      const PlaneAcc<T, SX> V{sh[1][sm], sh[1][s0], sh[1][sp], c};
      const PlaneAcc<T, SX> W{sh[2][sm], sh[2][s0], sh[2][sp], c};
      const long a = g.idx(i, j, k);
      T ou = T(0), ov = T(0), ow = T(0);
      if (STAGE == 2) {
        ou = qc.u;
        ov = qc.v;
        ow = qc.w;
        if (corr) {
          ou = ou - (qc.pe - qc.p0) * dtdxi;
          ov = ov - (qc.pn - qc.p0) * dtdyi;
          ow = ow - (qc.pt - qc.p0) * dt * m.dzc_inv[k];
        }
        if (defer) {
          if (fu) ou = ou + f_u;
          if (fv) ov = ov + f_v;
          if (fw) ow = ow + f_w;
          if (k == g.nz) ow = T(0);  // the wall value of the boundary update
        }
      }
      {
        const RkCoef q = coef[k];
        float iu, iv, iw;
        if (crf)
          tendency_fast<true, true>(U, V, W, q, cu, iu, iv, iw, &zc);
        else
          tendency_fast<false, true>(U, V, W, q, cu, iu, iv, iw, &zc);
        if (STAGE == 1) {
          du[a] = U(0, 0, 0) + iu;
          dv[a] = V(0, 0, 0) + iv;
          dw[a] = W(0, 0, 0) + iw;
        } else {
          const T nu = (U(0, 0, 0) + ou) * 0.5f + iu;
          du[a] = nu;
          dv[a] = (V(0, 0, 0) + ov) * 0.5f + iv;
          dw[a] = (W(0, 0, 0) + ow) * 0.5f + iw;
          if (acc) msum += (double)nu * dzflzi[k];
        }
      }
    }
    __syncthreads();
    if (more) commit(sm);
    __syncthreads();
    const int t = sm;
    sm = s0;
    s0 = sp;
    sp = t;
  }
  if (acc) {
    // the ring is free after the last barrier of the loop
    double* red = reinterpret_cast<double*>(&sh[0][0][0]);
    static_assert(sizeof(sh) >= NT * sizeof(double), "reduction scratch");
    red[threadIdx.x] = msum;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = red[0];
  }
}

// The FP64 stages: k_rk's ring, folding and deferred projection, in the
// reference's arithmetic and operation order (tendency_z, src/mod_calcVel.cc:
// 81,103,125 and :194,216,238; bit-identical to the reference's stages), with
// the plane loads scheduled late.  A fetch only issues the loads of plane k+2
// into registers; the projection, forcing and boundary rules apply when that
// plane is committed to the ring, after plane k has been computed, and plane
// k's new values are stored after the commit.  With that arithmetic inside the
// fetch (k_rk's form) every wave waits for its loads there, and a commit after
// the stores waits for their completion too: that form ran 4-5 % slower in FP64
// at 768x384x384 (4 waves per SIMD).  Every plane load is unconditional (a
// thread without a tile cell loads the corner cell (0, 0) and writes LDS slot
// PL; w is loaded on every plane and replaced by the wall value where it
// applies): loads that some paths skip leave the compiler's wait counting
// unsure which are still in flight, and it then waits for them before the next
// fetch reuses their registers.  The FP64 tendency carries the lower z face,
// so plane k-1 is not read while plane k is computed and one barrier per plane
// suffices.  Stage 2 issues the point values of plane k before the fetch, so
// that waiting for them leaves the plane in flight.
// tile rows of k_rk_late: 8 (two 512-thread blocks per CU at 126 VGPRs); 4 rows
// measured 9 % and 16 rows 5 % slower at 768x384x384
constexpr int kLateTy = 8;
template <int STAGE, bool PROJ>
__global__ __launch_bounds__(kRkTx * kLateTy) void k_rk_late(
    Geom g, const double* __restrict__ su, const double* __restrict__ sv,
    const double* __restrict__ sw, double* __restrict__ du, double* __restrict__ dv,
    double* __restrict__ dw, MeshT<double> m, double dxi, double dyi, double rei, double dt,
    double ucrf, int kchunk, int ntx, int nty, int fold, int defer, double bc_u,
    const double* __restrict__ force, int forced_mask, const RkCoef* __restrict__,
    RkUniform, int crf, const double* __restrict__ pp, int corr, double dtdxi, double dtdyi,
    int, const double* __restrict__, double* __restrict__, int) {
  using T = double;
  constexpr int TX = kRkTx, TY = kLateTy, SX = TX + 2, SY = TY + 2, PL = SX * SY,
                NT = TX * TY;
  constexpr int NPF = (PL + NT - 1) / NT;
  // stage 1 launches the PROJ variant exactly when the projection is deferred
  constexpr bool PJ = STAGE == 1 && PROJ;
  corr = PROJ ? corr : 0;
  const int bcload = fold;
  __shared__ T sh[3][3][PL + 1];  // [field][ring slot][tile cell]; slot PL: no cell
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  const int i0 = ti * TX, j0 = tj * TY;  // tile cell (0,0) is the halo corner (i0, j0)
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool active = i <= g.nx && j <= g.ny;
  const bool any = fold || defer;
  const bool fu = any && (forced_mask & 1), fv = any && (forced_mask & 2),
             fw = any && (forced_mask & 4);
  const T f_u = fu ? force[0] : T(0), f_v = fv ? force[1] : T(0), f_w = fw ? force[2] : T(0);

  long gaddr[NPF];  // plane-relative source address of this thread's tile cells
  int cell[NPF];
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int e = threadIdx.x + n * NT;
    int gx = i0 + e % SX, gy = j0 + e / SX;
    const bool inside = e < PL && gx <= g.nx + 1 && gy <= g.ny + 1;
    if (bcload) {  // halo cells read their periodic image
      gx = gx == 0 ? g.nx : (gx == g.nx + 1 ? 1 : gx);
      gy = gy == 0 ? g.ny : (gy == g.ny + 1 ? 1 : gy);
    }
    gaddr[n] = inside ? gx + g.sx * gy : 0;
    cell[n] = inside ? e : PL;
  }
  T ru[NPF], rv[NPF], rw[NPF];
  T rp0[NPF], rpe[NPF], rpn[NPF], rpt[NPF];  // p at the cell, +x, +y, +z (PJ)
  auto plane_src = [&](int kk) {
    const bool ghost = bcload && (kk == 0 || kk == g.nz + 1);
    return ghost ? (kk == 0 ? 1 : g.nz) : kk;
  };
  auto fetch = [&](int kk) {
    const long pofs = g.sxy * plane_src(kk);
#pragma unroll
    for (int n = 0; n < NPF; ++n) {
      const long a = gaddr[n] + pofs;
      ru[n] = su[a];
      rv[n] = sv[a];
      rw[n] = sw[a];
      if constexpr (PJ) {
        rp0[n] = pp[a];
        rpe[n] = pp[a + 1];
        rpn[n] = pp[a + g.sx];
        rpt[n] = pp[a + g.sxy];
      }
    }
  };
  // k_rk's fetch arithmetic, applied to the loaded registers of plane kk
  auto commit = [&](int kk, int slot) {
    const bool ghost = bcload && (kk == 0 || kk == g.nz + 1);
    const int ks = plane_src(kk);
    const bool wzero = bcload && (kk == 0 || kk >= g.nz);
    const T dzci = PJ ? m.dzc_inv[ks] : T(0);
#pragma unroll
    for (int n = 0; n < NPF; ++n) {
      T qu = ru[n], qv = rv[n], qw = rw[n];
      if constexpr (PJ) {
        const T p0 = rp0[n];
        qu = qu - (rpe[n] - p0) * dtdxi;
        qv = qv - (rpn[n] - p0) * dtdyi;
        // same association as k_correct: ((p(k+1) - p(k)) * dt) * dzc_inv(k)
        qw = qw - (rpt[n] - p0) * dt * dzci;
      }
      if (fold) {
        if (fu) qu = qu + f_u;
        if (fv) qv = qv + f_v;
        if (fw) qw = qw + f_w;
      }
      if (wzero) qw = T(0);
      if (ghost) {
        qu = T(2) * bc_u - qu;  // src/mod_updateBound.cc:92,123
        qv = T(2) * T(0) - qv;
      }
      sh[0][slot][cell[n]] = qu;
      sh[1][slot][cell[n]] = qv;
      sh[2][slot][cell[n]] = qw;
    }
  };
  if (k0 > k1) return;
  fetch(k0 - 1);
  commit(k0 - 1, 0);
  fetch(k0);
  commit(k0, 1);
  fetch(k0 + 1);
  commit(k0 + 1, 2);
  __syncthreads();
  int sm = 0, s0 = 1, sp = 2;
  const int c = (ty + 1) * SX + tx + 1;
  // stage 2's point values and, with corr, the pressures around the point,
  // loaded for every thread (a thread without a point reads cell (1, 1))
  T qu0 = T(0), qv0 = T(0), qw0 = T(0), qp0 = T(0), qpe = T(0), qpn = T(0), qpt = T(0);
  auto load_point = [&](int kk) {
    const long a = active ? g.idx(i, j, kk) : g.idx(1, 1, kk);
    qu0 = du[a];
    qv0 = dv[a];
    qw0 = dw[a];
    if (corr) {
      qp0 = pp[a];
      qpe = pp[a + 1];
      qpn = pp[a + g.sx];
      qpt = pp[a + g.sxy];
    }
  };
  // the lower z-face terms of the thread's cell, carried up the column
  ZFace<T> zf{};
  if (active) {
    const PlaneAcc<T, SX> U{sh[0][sm], sh[0][s0], sh[0][sp], c};
    const PlaneAcc<T, SX> V{sh[1][sm], sh[1][s0], sh[1][sp], c};
    const PlaneAcc<T, SX> W{sh[2][sm], sh[2][s0], sh[2][sp], c};
    zf = zface_lower<T>(U, V, W, m, k0);
  }
  __syncthreads();  // plane k0-1 is read above and rewritten in the loop
  T nu = T(0), nv = T(0), nw = T(0);  // plane k's new values until they are stored
  auto compute = [&](int k) {
    if (!active) return;
    const PlaneAcc<T, SX> U{sh[0][sm], sh[0][s0], sh[0][sp], c};
    const PlaneAcc<T, SX> V{sh[1][sm], sh[1][s0], sh[1][sp], c};
    const PlaneAcc<T, SX> W{sh[2][sm], sh[2][s0], sh[2][sp], c};
    // reference arithmetic, src/mod_calcVel.cc:81,103,125 and :194,216,238
    T tu, tv, tw;
    if (crf)
      tendency_z<T, true, true>(U, V, W, m, k, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
    else
      tendency_z<T, false, true>(U, V, W, m, k, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
    if (STAGE == 1) {
      nu = U(0, 0, 0) + dt * tu;
      nv = V(0, 0, 0) + dt * tv;
      nw = W(0, 0, 0) + dt * tw;
    } else {
      T ou = qu0, ov = qv0, ow = qw0;
      if (corr) {
        ou = ou - (qpe - qp0) * dtdxi;
        ov = ov - (qpn - qp0) * dtdyi;
        ow = ow - (qpt - qp0) * dt * m.dzc_inv[k];
      }
      if (defer) {
        if (fu) ou = ou + f_u;
        if (fv) ov = ov + f_v;
        if (fw) ow = ow + f_w;
        if (k == g.nz) ow = T(0);  // the wall value of the boundary update
      }
      nu = (U(0, 0, 0) + dt * tu + ou) * T(0.5);
      nv = (V(0, 0, 0) + dt * tv + ov) * T(0.5);
      nw = (W(0, 0, 0) + dt * tw + ow) * T(0.5);
    }
  };
  auto store = [&](int k) {
    if (active) {
      const long a = g.idx(i, j, k);
      du[a] = nu;
      dv[a] = nv;
      dw[a] = nw;
    }
  };
  // the last plane, whose k+2 lies beyond the chunk's ring, runs after the loop
  for (int k = k0; k < k1; ++k) {
    if (STAGE == 2) load_point(k);
    fetch(k + 2);
    compute(k);
    commit(k + 2, sm);
    store(k);
    __syncthreads();
    const int t = sm;
    sm = s0;
    s0 = sp;
    sp = t;
  }
  if (STAGE == 2) load_point(k1);
  compute(k1);
  store(k1);
}

// Both RK stages of an FP32-mixed step inside a run (src/mod_calcVel.cc:61-125
// then :174-238) in one pass.  The stored velocity is the previous step's
// predictor u*; its projection, forcing and boundary update are applied on
// load as in k_rk's stage 1 (fold and corr).  A block owns a 64 x 8 tile of
// stage-2 points and marches a chunk of k planes.  Stage 1 runs on the tile
// plus a one-cell ring (66 x 10 cells) from a three-plane LDS ring of the
// loaded source (68 x 12 cells); its result u1 goes to a second three-plane
// LDS ring and never to HBM.  The u1 values on the ring are stage 1 at the
// periodic position -- the x/y boundary update of u1 -- and the planes 0 and
// nz+1 of u1 and w1 on plane nz follow the boundary rules
// (src/mod_updateBound.cc:92,123,137), as the stage-2 loads of k_rk rebuild
// them.  Stage 2's point values are the loaded source at the tile point: the
// same projection and forcing k_rk's stage 2 applies to them.  Per point, u,
// v, w and p are read once and the new u, v, w written once.  The result goes
// to separate arrays, since other blocks still read the source.  With `acc`
// the block sums the new u times dzflzi(k) in FP64 into partial[block].
template <int TY>
__global__ __launch_bounds__(kRkTx* TY) void k_rk12(
    Geom g, const float* __restrict__ su, const float* __restrict__ sv,
    const float* __restrict__ sw, const float* __restrict__ pp, float* __restrict__ du,
    float* __restrict__ dv, float* __restrict__ dw, const float* __restrict__ dzc_inv, float dt,
    float dtdxi, float dtdyi, float bc_u, const double* __restrict__ force, int forced_mask,
    const RkCoef* __restrict__ coef1, const RkCoef* __restrict__ coef2, RkUniform cu1,
    RkUniform cu2, int crf, int kchunk, int ntx, int nty, int acc,
    const double* __restrict__ dzflzi, double* __restrict__ partial) {
  constexpr int TX = kRkTx, NT = TX * TY;
  constexpr int RX = TX + 2, RY = TY + 2, RPL = RX * RY;  // u1 region: tile + 1
  constexpr int UX = TX + 4, UY = TY + 4, UPL = UX * UY;  // source: tile + 2
  constexpr int NPF = (UPL + NT - 1) / NT;
  constexpr int NRING = RPL - NT;  // region cells outside the tile
  static_assert(NRING == 2 * RX + 2 * TY && NRING <= NT, "one ring cell per thread");
  __shared__ float ssrc[3][3][UPL];  // [field][slot][cell]; cell (cx, cy) = (i0-1+cx, j0-1+cy)
  __shared__ float su1[3][3][RPL];   // [field][plane % 3][cell]; cell (rx, ry) = (i0+rx, j0+ry)
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  const int i0 = ti * TX, j0 = tj * TY;
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool active = i <= g.nx && j <= g.ny;
  const bool fu = forced_mask & 1, fv = forced_mask & 2, fw = forced_mask & 4;
  const float f_u = fu ? (float)force[0] : 0.f, f_v = fv ? (float)force[1] : 0.f,
              f_w = fw ? (float)force[2] : 0.f;
  if (k0 > k1) {
    if (acc && t == 0) partial[b] = 0.0;
    return;
  }
  auto wrap = [](int x, int n) { return ((x - 1) % n + n) % n + 1; };
  int gaddr[NPF];  // plane-relative address of the periodic image (-1: past the tile)
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int e = t + n * NT;
    gaddr[n] = e < UPL ? wrap(i0 - 1 + e % UX, g.nx) + (int)g.sx * wrap(j0 - 1 + e / UX, g.ny) : -1;
  }
  // Source plane kk with the rules of k_rk's folded, projected stage-1 loads.
  // The fetch only issues the loads (w on every plane, from plane 1 or nz for
  // the ghost planes); the projection, forcing and boundary rules apply at the
  // commit, after stage 2 of the plane below has been computed, and stage 2's
  // results are stored after the commit: arithmetic on the loaded values in the
  // fetch, or stores ahead of the commit, would make a wave wait there.
  float ru[NPF], rv[NPF], rw[NPF], rp0[NPF], rpe[NPF], rpn[NPF], rpt[NPF];
  auto fetch = [&](int kk) {
    const bool ghost = kk == 0 || kk == g.nz + 1;
    const int ks = ghost ? (kk == 0 ? 1 : g.nz) : kk;
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        const long a = gaddr[n] + g.sxy * ks;
        ru[n] = su[a];
        rv[n] = sv[a];
        rw[n] = sw[a];
        rp0[n] = pp[a];
        rpe[n] = pp[a + 1];
        rpn[n] = pp[a + g.sx];
        rpt[n] = pp[a + g.sxy];
      }
  };
  auto commit = [&](int kk, int slot) {
    const bool ghost = kk == 0 || kk == g.nz + 1;
    const int ks = ghost ? (kk == 0 ? 1 : g.nz) : kk;
    const bool wzero = kk == 0 || kk >= g.nz;
    const float dzci = dzc_inv[ks];
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        const float p0 = rp0[n];
        float qu = ru[n] - (rpe[n] - p0) * dtdxi;
        float qv = rv[n] - (rpn[n] - p0) * dtdyi;
        float qw = rw[n] - (rpt[n] - p0) * dt * dzci;
        if (fu) qu = qu + f_u;
        if (fv) qv = qv + f_v;
        if (fw) qw = qw + f_w;
        if (wzero) qw = 0.f;
        if (ghost) {
          qu = 2.f * bc_u - qu;  // src/mod_updateBound.cc:92,123
          qv = 2.f * 0.f - qv;
        }
        const int e = t + n * NT;
        ssrc[0][slot][e] = qu;
        ssrc[1][slot][e] = qv;
        ssrc[2][slot][e] = qw;
      }
  };
  // this thread's stage-1 cells: A = its own tile point, B = one ring cell (t < NRING)
  const int ra = (ty + 1) * RX + tx + 1;
  int bx, by;
  if (t < RX) {
    bx = t;
    by = 0;
  } else if (t < 2 * RX) {
    bx = t - RX;
    by = RY - 1;
  } else if (t < 2 * RX + TY) {
    bx = 0;
    by = t - 2 * RX + 1;
  } else {
    bx = RX - 1;
    by = t - 2 * RX - TY + 1;
  }
  const bool has_b = t < NRING;
  const int rb = by * RX + bx;
  int sm = 0, s0 = 1, sp = 2;  // source slots of planes q-1, q, q+1
  const int qs = max(k0 - 1, 1), qe = k1 + 1;  // u1 planes formed: stage 1 up to nz, rules at nz+1
  fetch(qs - 1);
  commit(qs - 1, 0);
  fetch(qs);
  commit(qs, 1);
  fetch(qs + 1);
  commit(qs + 1, 2);
  __syncthreads();
  ZConv za{}, zb{}, z2{};  // carried lower-face products: stage 1 at A and B, stage 2
  // stage 1 of region cell (rx, ry) on plane q, w1 = 0 on the wall plane nz.
  // The carry restarts where k_rk's stage 1 restarts it (the first plane of
  // each chunk of kchunk planes), and plane k0-1 continues the chain of the
  // chunk below, so u1 is bit-identical to k_rk's.
  auto stage1 = [&](int q, int rx, int ry, ZConv& z, float& o1, float& o2, float& o3) {
    const int c = (ry + 1) * UX + rx + 1;
    const PlaneAcc<float, UX> U{ssrc[0][sm], ssrc[0][s0], ssrc[0][sp], c};
    const PlaneAcc<float, UX> V{ssrc[1][sm], ssrc[1][s0], ssrc[1][sp], c};
    const PlaneAcc<float, UX> W{ssrc[2][sm], ssrc[2][s0], ssrc[2][sp], c};
    const RkCoef cq = coef1[q];
    if ((q - 1) % kchunk == 0) {
      z = zconv_lower(U, V, W, cq);
    } else if (q == qs) {
      const PlaneAcc<float, UX> Ub{ssrc[0][sm], ssrc[0][sm], ssrc[0][s0], c};
      const PlaneAcc<float, UX> Vb{ssrc[1][sm], ssrc[1][sm], ssrc[1][s0], c};
      const PlaneAcc<float, UX> Wb{ssrc[2][sm], ssrc[2][sm], ssrc[2][s0], c};
      z = zconv_upper(Ub, Vb, Wb, coef1[q - 1]);
    }
    float iu, iv, iw;
    if (crf)
      tendency_fast<true, true>(U, V, W, cq, cu1, iu, iv, iw, &z);
    else
      tendency_fast<false, true>(U, V, W, cq, cu1, iu, iv, iw, &z);
    o1 = U(0, 0, 0) + iu;
    o2 = V(0, 0, 0) + iv;
    o3 = q >= g.nz ? 0.f : W(0, 0, 0) + iw;
  };
  // u1 of plane q at region cell r; with q == 1 also plane 0 by the rules
  auto put1 = [&](int q, int r, float o1, float o2, float o3) {
    const int s = q % 3;
    su1[0][s][r] = o1;
    su1[1][s][r] = o2;
    su1[2][s][r] = o3;
    if (q == 1) {
      su1[0][0][r] = 2.f * bc_u - o1;
      su1[1][0][r] = 2.f * 0.f - o2;
      su1[2][0][r] = 0.f;
    }
  };
  // plane nz+1 of u1 by the rules from plane nz (the same thread wrote cell r of it)
  auto ghost_top = [&](int r) {
    const int s = g.nz % 3, d = (g.nz + 1) % 3;
    su1[0][d][r] = 2.f * bc_u - su1[0][s][r];
    su1[1][d][r] = 2.f * 0.f - su1[1][s][r];
    su1[2][d][r] = 0.f;
  };
  const int cp = (ty + 2) * UX + tx + 2;  // the tile point in the source ring
  double msum = 0.0;
  for (int q = qs; q <= qe; ++q) {
    // the next iteration's stage 1 reads source planes q .. q+2
    const bool more = q + 1 <= qe && q + 1 <= g.nz;
    const int k = q - 1;  // stage-2 plane
    const bool s2 = k >= k0;
    float pu = 0.f, pv = 0.f, pw = 0.f;
    if (s2) {
      pu = ssrc[0][sm][cp];
      pv = ssrc[1][sm][cp];
      pw = ssrc[2][sm][cp];
    }
    if (q <= g.nz) {
      float o1, o2, o3;
      stage1(q, tx + 1, ty + 1, za, o1, o2, o3);
      put1(q, ra, o1, o2, o3);
      if (has_b) {
        stage1(q, bx, by, zb, o1, o2, o3);
        put1(q, rb, o1, o2, o3);
      }
    } else {
      ghost_top(ra);
      if (has_b) ghost_top(rb);
    }
    __syncthreads();
    // fetched here rather than before stage 1: its registers are then not
    // live across both stages
    if (more) fetch(q + 2);
    float nu = 0.f, nv = 0.f, nw = 0.f;
    if (s2 && active) {
      const PlaneAcc<float, RX> U{su1[0][(k + 2) % 3], su1[0][k % 3], su1[0][(k + 1) % 3], ra};
      const PlaneAcc<float, RX> V{su1[1][(k + 2) % 3], su1[1][k % 3], su1[1][(k + 1) % 3], ra};
      const PlaneAcc<float, RX> W{su1[2][(k + 2) % 3], su1[2][k % 3], su1[2][(k + 1) % 3], ra};
      const RkCoef cq = coef2[k];
      if (k == k0) z2 = zconv_lower(U, V, W, cq);
      float iu, iv, iw;
      if (crf)
        tendency_fast<true, true>(U, V, W, cq, cu2, iu, iv, iw, &z2);
      else
        tendency_fast<false, true>(U, V, W, cq, cu2, iu, iv, iw, &z2);
      nu = (U(0, 0, 0) + pu) * 0.5f + iu;
      nv = (V(0, 0, 0) + pv) * 0.5f + iv;
      nw = (W(0, 0, 0) + pw) * 0.5f + iw;
      if (acc) msum += (double)nu * dzflzi[k];
    }
    if (more) commit(q + 2, sm);  // source plane q+2 replaces q-1
    if (s2 && active) {
      const long a = g.idx(i, j, k);
      du[a] = nu;
      dv[a] = nv;
      dw[a] = nw;
    }
    __syncthreads();
    const int tmp = sm;
    sm = s0;
    s0 = sp;
    sp = tmp;
  }
  if (acc) {
    // the block size is folded onto its largest power of two first
    double* red = reinterpret_cast<double*>(&ssrc[0][0][0]);
    static_assert(sizeof(ssrc) >= NT * sizeof(double), "reduction scratch");
    constexpr int P2 = NT >= 1024 ? 1024 : (NT >= 512 ? 512 : 256);
    static_assert(NT >= P2 && NT < 2 * P2, "block size");
    red[t] = msum;
    __syncthreads();
    if (t < NT - P2) red[t] += red[t + P2];
    __syncthreads();
    for (int h = P2 / 2; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) partial[b] = red[0];
  }
}

// RK stage 2 fused with calcRHS (src/mod_calcVel.cc:140-242 then
// src/mod_calcRHS.cc:27-29) for the step sequence.  Besides its 64x8 interior
// the block evaluates u_new on the column i0 and v_new on the row j0 (the
// backward neighbours calcRHS needs); at the domain edge these are computed
// from periodic images of the stencil and equal the boundary copy bit for bit.
// w_new(k-1) is carried in a register (a chunk first recomputes it for its
// plane k0-1; below the wall it is the boundary value 0).  Loads rebuild the
// u1 halos from interior cells (so the boundary update of u1 is not needed),
// w_new on plane nz is the wall value 0 (the middle boundary update's only
// effect on the interior), and the deferred force of the previous step is
// added to the point values of u.  The new state goes to separate buffers
// (nu, nv, nw): the backward neighbours read old values of cells that other
// blocks update, so the update cannot be in place.
//
// With PROJ the previous step's projection is deferred into this kernel as
// well: the stored u, v, w are the predictor u* and every point value read is
// u* - dt*grad(p) (src/mod_calcVel.cc:272-274, k_correct's association) before
// the force applies; p (pin, valid halos) is the previous pressure, held in a
// fourth field of the LDS ring, and the right-hand side goes to a separate
// buffer (pout).  With `acc` the block sums
// its new u times dzflzi(k) in FP64 into partial[block]: the bulk mean of the
// next projected velocity (the projection changes every x-row's sum only by
// p(nx+1) - p(1) = 0).
template <typename T, int TY, bool PROJ>
__global__ __launch_bounds__(kRkTx* TY) void k_rk2_rhs(
    Geom g, const T* __restrict__ su, const T* __restrict__ sv, const T* __restrict__ sw,
    const T* __restrict__ du, const T* __restrict__ dv, const T* __restrict__ dw,
    T* __restrict__ nu_out, T* __restrict__ nv_out, T* __restrict__ nw_out,
    const T* __restrict__ pin, T* __restrict__ pout, MeshT<T> m, T dxi, T dyi, T rei, T dt,
    T ucrf, int kchunk, int ntx, int nty, int defer, T bc_u, const double* __restrict__ force,
    int forced_mask, const RkCoef* __restrict__ coef, RkUniform cu, int crf, T dti, T dtidxi,
    T dtidyi, T pdx, T pdy, int acc, const double* __restrict__ dzflzi,
    double* __restrict__ partial) {
  constexpr int TX = kRkTx, SX = TX + 3, SY = TY + 3, PL = SX * SY, NT = TX * TY;
  constexpr int NPF = (PL + NT - 1) / NT;
  constexpr int NF = PROJ ? 4 : 3;
  __shared__ T sh[NF][3][PL];  // u1, v1, w1 (and p) ring; cell (cx, cy) = (i0-1+cx, j0-1+cy)
  __shared__ T un[TY][TX + 1];          // u_new at i = i0 .. i0+TX of the tile rows
  __shared__ T vn[TY + 1][TX];          // v_new at j = j0 .. j0+TY of the tile columns
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  const int i0 = ti * TX, j0 = tj * TY;
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool active = i <= g.nx && j <= g.ny;
  const bool fu = defer && (forced_mask & 1), fv = defer && (forced_mask & 2),
             fw = defer && (forced_mask & 4);
  const T f_u = fu ? (T)force[0] : T(0), f_v = fv ? (T)force[1] : T(0),
          f_w = fw ? (T)force[2] : T(0);
  auto wrapx = [&](int x) { return x < 1 ? x + g.nx : (x > g.nx ? x - g.nx : x); };
  auto wrapy = [&](int y) { return y < 1 ? y + g.ny : (y > g.ny ? y - g.ny : y); };

  long gaddr[NPF];
  int cell[NPF];
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int e = threadIdx.x + n * NT;
    const int gx = i0 - 1 + e % SX, gy = j0 - 1 + e / SX;
    cell[n] = e;
    const bool inside = e < PL && gx <= g.nx + 1 && gy <= g.ny + 1;
    gaddr[n] = inside ? wrapx(gx) + g.sx * wrapy(gy) : -1;
  }
  T pre[NF][NPF];
  // u1, v1, w1 of plane kk with the boundary rules applied, and p of plane kk as
  // stored (updateBoundP ran; x/y halo cells read their periodic image, which
  // holds the same value).  The fetch only issues the loads (w on every plane,
  // from plane 1 or nz for the ghost planes); the boundary rules apply at the
  // commit, after the plane below has been computed: arithmetic on the loaded
  // values inside the fetch would make every wave wait for its loads there.
  auto fetch = [&](int kk) {
    const bool ghost = kk == 0 || kk == g.nz + 1;
    const int ks = ghost ? (kk == 0 ? 1 : g.nz) : kk;
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        const long a = gaddr[n] + g.sxy * ks;
        pre[0][n] = su[a];
        pre[1][n] = sv[a];
        pre[2][n] = sw[a];
        if constexpr (PROJ) pre[NF - 1][n] = pin[gaddr[n] + g.sxy * kk];
      }
  };
  auto commit = [&](int kk, int slot) {
    const bool ghost = kk == 0 || kk == g.nz + 1;
    const bool wzero = kk == 0 || kk >= g.nz;
#pragma unroll
    for (int n = 0; n < NPF; ++n)
      if (gaddr[n] >= 0) {
        T qu = pre[0][n], qv = pre[1][n], qw = pre[2][n];
        if (wzero) qw = T(0);
        if (ghost) {
          qu = T(2) * bc_u - qu;  // src/mod_updateBound.cc:92,123
          qv = T(2) * T(0) - qv;
        }
        sh[0][slot][cell[n]] = qu;
        sh[1][slot][cell[n]] = qv;
        sh[2][slot][cell[n]] = qw;
        if constexpr (PROJ) sh[NF - 1][slot][cell[n]] = pre[NF - 1][n];
      }
  };
  // the projection's pressure differences at tile cell c of ring slots s0
  // (plane k) and sp (plane k+1)
  auto dpx = [&](int s0, int c) { return sh[NF - 1][s0][c + 1] - sh[NF - 1][s0][c]; };
  auto dpy = [&](int s0, int c) { return sh[NF - 1][s0][c + SX] - sh[NF - 1][s0][c]; };
  auto dpz = [&](int s0, int sp, int c) { return sh[NF - 1][sp][c] - sh[NF - 1][s0][c]; };
  // old u at tile cell (1, cy), old v at (cx, 1) of plane kk (slot s0), with
  // the deferred projection and force; x/y halo cells wrap
  auto old_u = [&](int s0, int cy, int x, int y, int kk) {
    T q = du[g.idx(wrapx(x), wrapy(y), kk)];
    if constexpr (PROJ) q = q - dpx(s0, cy * SX + 1) * pdx;
    return fu ? q + f_u : q;
  };
  auto old_v = [&](int s0, int cx, int x, int y, int kk) {
    T q = dv[g.idx(wrapx(x), wrapy(y), kk)];
    if constexpr (PROJ) q = q - dpy(s0, SX + cx) * pdy;
    return fv ? q + f_v : q;
  };
  // the three stage-2 updates at tile cell (cx, cy) of plane kk.  FP64: zf
  // receives the upper z-face terms; with `carry` it also supplies the lower
  // ones (the thread's own column, marching up)
  auto update = [&](int sm, int s0, int sp, int cx, int cy, int kk, T& nu, T& nv, T& nw,
                    T ou, T ov, T ow, ZFace<T>& zf, bool carry) {
    const int c = cy * SX + cx;
    const PlaneAcc<T, SX> U{sh[0][sm], sh[0][s0], sh[0][sp], c};
    const PlaneAcc<T, SX> V{sh[1][sm], sh[1][s0], sh[1][sp], c};
    const PlaneAcc<T, SX> W{sh[2][sm], sh[2][s0], sh[2][sp], c};
    if constexpr (std::is_same<T, double>::value) {
      T tu, tv, tw;
      if (!carry)
        tendency_z<T, true, false>(U, V, W, m, kk, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
      else if (crf)
        tendency_z<T, true, true>(U, V, W, m, kk, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
      else
        tendency_z<T, false, true>(U, V, W, m, kk, dxi, dyi, rei, ucrf, tu, tv, tw, zf);
      nu = (U(0, 0, 0) + dt * tu + ou) * T(0.5);
      nv = (V(0, 0, 0) + dt * tv + ov) * T(0.5);
      nw = (W(0, 0, 0) + dt * tw + ow) * T(0.5);
    } else {
      const RkCoef q = coef[kk];
      float iu, iv, iw;
      if (crf)
        tendency_fast<true>(U, V, W, q, cu, iu, iv, iw);
      else
        tendency_fast<false>(U, V, W, q, cu, iu, iv, iw);
      nu = (U(0, 0, 0) + ou) * 0.5f + iu;
      nv = (V(0, 0, 0) + ov) * 0.5f + iv;
      nw = (W(0, 0, 0) + ow) * 0.5f + iw;
    }
  };
  if (k0 > k1) {
    if (acc && threadIdx.x == 0) partial[b] = 0.0;
    return;
  }
  const int cxi = tx + 2, cyi = ty + 2;  // tile cell of the thread's interior point
  double msum = 0.0;  // FP64 partial of the bulk mean (acc)
  ZFace<T> zf{}, zx{};  // the column's carried z-face terms; scratch of the extra cells
  int sm = 0, s0 = 1, sp = 2;
  T wprev = T(0);  // w_new(i, j, k-1); 0 below the wall (w(0) = 0)
  if (k0 > 1) {
    // recompute w_new on plane k0-1 from u1 planes k0-2, k0-1, k0
    fetch(k0 - 2);
    commit(k0 - 2, 0);
    fetch(k0 - 1);
    commit(k0 - 1, 1);
    fetch(k0);
    commit(k0, 2);
    __syncthreads();
    if (active) {
      T nu, nv, nw;
      T ow = dw[g.idx(i, j, k0 - 1)];
      // same association as k_correct: ((p(k+1) - p(k)) * dt) * dzc_inv(k)
      if constexpr (PROJ) ow = ow - dpz(1, 2, cyi * SX + cxi) * dt * m.dzc_inv[k0 - 1];
      if (fw) ow = (k0 - 1 == g.nz) ? T(0) : ow + f_w;
      update(0, 1, 2, cxi, cyi, k0 - 1, nu, nv, nw, T(0), T(0), ow, zf, false);
      wprev = (k0 - 1 == g.nz) ? T(0) : nw;
    }
    __syncthreads();
    fetch(k0 + 1);
    commit(k0 + 1, 0);  // slot 0 held plane k0-2
    __syncthreads();
    sm = 1;
    s0 = 2;
    sp = 0;
  } else {
    fetch(k0 - 1);
    commit(k0 - 1, 0);
    fetch(k0);
    commit(k0, 1);
    fetch(k0 + 1);
    commit(k0 + 1, 2);
    __syncthreads();
    if (active) {
      const int c = cyi * SX + cxi;
      const PlaneAcc<T, SX> U{sh[0][0], sh[0][1], sh[0][2], c};
      const PlaneAcc<T, SX> V{sh[1][0], sh[1][1], sh[1][2], c};
      const PlaneAcc<T, SX> W{sh[2][0], sh[2][1], sh[2][2], c};
      zf = zface_lower<T>(U, V, W, m, k0);
    }
  }
  for (int k = k0; k <= k1; ++k) {
    const bool more = k + 2 <= k1 + 1;
    if (more) fetch(k + 2);
    T nw = T(0);
    if (active) {
      const long a = g.idx(i, j, k);
      T ou = du[a], ov = dv[a], ow = dw[a];
      if constexpr (PROJ) {
        // same association as k_correct: ((p(k+1) - p(k)) * dt) * dzc_inv(k)
        const int c = cyi * SX + cxi;
        ou = ou - dpx(s0, c) * pdx;
        ov = ov - dpy(s0, c) * pdy;
        ow = ow - dpz(s0, sp, c) * dt * m.dzc_inv[k];
      }
      if (fu) ou = ou + f_u;
      if (fv) ov = ov + f_v;
      if (fw) ow = (k == g.nz) ? T(0) : ow + f_w;
      T nu, nv;
      update(sm, s0, sp, cxi, cyi, k, nu, nv, nw, ou, ov, ow, zf, true);
      if (k == g.nz) nw = T(0);  // src/mod_updateBound.cc:137 (w(nz) = 0)
      un[ty][tx + 1] = nu;
      vn[ty + 1][tx] = nv;
      if (acc) msum += (double)nu * dzflzi[k];
      nu_out[a] = nu;
      nv_out[a] = nv;
      nw_out[a] = nw;
    }
    // backward neighbours: v_new on row j0 (wave 0), u_new on column i0 (wave 1)
    if (ty == 0 && i <= g.nx) {
      T nu, nv, nw2;
      update(sm, s0, sp, cxi, 1, k, nu, nv, nw2, T(0), old_v(s0, cxi, i, j0, k), T(0), zx, false);
      vn[0][tx] = nv;
    } else if (ty == 1 && tx < TY && j0 + tx + 1 <= g.ny) {
      T nu, nv, nw2;
      update(sm, s0, sp, 1, tx + 2, k, nu, nv, nw2, old_u(s0, tx + 2, i0, j0 + tx + 1, k), T(0),
             T(0), zx, false);
      un[tx][0] = nu;
    }
    __syncthreads();
    if (active) {
      // calcRHS, src/mod_calcRHS.cc:27-29
      pout[g.idx(i, j, k)] = (un[ty][tx + 1] - un[ty][tx]) * dtidxi +
                             (vn[ty + 1][tx] - vn[ty][tx]) * dtidyi + (nw - wprev) * dti * m.dzf_inv[k];
      wprev = nw;
    }
    if (more) commit(k + 2, sm);
    __syncthreads();
    const int t = sm;
    sm = s0;
    s0 = sp;
    sp = t;
  }
  if (acc) {
    // the ring is free after the last barrier of the loop; the block size
    // (64 x TY) is folded onto its largest power of two first
    double* red = reinterpret_cast<double*>(&sh[0][0][0]);
    static_assert(sizeof(sh) >= NT * sizeof(double), "reduction scratch");
    constexpr int P2 = NT >= 1024 ? 1024 : (NT >= 512 ? 512 : 256);
    static_assert(NT >= P2 && NT < 2 * P2, "block size");
    red[threadIdx.x] = msum;
    __syncthreads();
    if ((int)threadIdx.x < NT - P2) red[threadIdx.x] += red[threadIdx.x + P2];
    __syncthreads();
    for (int h = P2 / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = red[0];
  }
}

// calcRHS, src/mod_calcRHS.cc:27-29.  w(k-1) is carried from the previous plane.
template <typename T>
__global__ __launch_bounds__(kRkTx* kRkTy) void k_calc_rhs(Geom g, const T* __restrict__ u,
                                                           const T* __restrict__ v,
                                                           const T* __restrict__ w,
                                                           T* __restrict__ p,
                                                           const T* __restrict__ dzf_inv, T dti,
                                                           T dtidxi, T dtidyi, int kchunk,
                                                           int ntx, int nty) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  const int i = ti * kRkTx + (int)(threadIdx.x % kRkTx) + 1;
  const int j = tj * kRkTy + (int)(threadIdx.x / kRkTx) + 1;
  if (i > g.nx || j > g.ny) return;
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  long c = g.idx(i, j, k0);
  T wb = k0 <= k1 ? w[c - g.sxy] : T(0);
  for (int k = k0; k <= k1; ++k, c += g.sxy) {
    const T w0 = w[c];
    p[c] = (u[c] - u[c - 1]) * dtidxi + (v[c] - v[c - g.sx]) * dtidyi + (w0 - wb) * dti * dzf_inv[k];
    wb = w0;
  }
}

// correctVel, src/mod_calcVel.cc:272-274.  p(k+1) becomes the next plane's p(k).
// With ACC the block also sums the corrected u times dzflzi(k) (FP64) into
// partial[block] -- the first half of the bulk mean that forceVel needs.
template <typename T, bool ACC>
__global__ __launch_bounds__(kRkTx* correct_ty<T>()) void k_correct(Geom g, const T* __restrict__ p,
                                                          T* __restrict__ u, T* __restrict__ v,
                                                          T* __restrict__ w,
                                                          const T* __restrict__ dzc_inv, T dtdxi,
                                                          T dtdyi, T dt, int kchunk, int ntx,
                                                          int nty, const double* __restrict__ dzflzi,
                                                          double* __restrict__ partial) {
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = (int)(b % (unsigned)ntx), tj = (int)((b / (unsigned)ntx) % (unsigned)nty);
  const int kc = (int)(b / (unsigned)(ntx * nty));
  constexpr int TY = correct_ty<T>();
  const int i = ti * kRkTx + (int)(threadIdx.x % kRkTx) + 1;
  const int j = tj * TY + (int)(threadIdx.x / kRkTx) + 1;
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  double acc = 0.0;
  if (i <= g.nx && j <= g.ny && k0 <= k1) {
    long c = g.idx(i, j, k0);
    T p0 = p[c];
    for (int k = k0; k <= k1; ++k, c += g.sxy) {
      const T pt = p[c + g.sxy];
      const T un = u[c] - (p[c + 1] - p0) * dtdxi;
      u[c] = un;
      v[c] -= (p[c + g.sx] - p0) * dtdyi;
      w[c] -= (pt - p0) * dt * dzc_inv[k];
      if (ACC) acc += (double)un * dzflzi[k];
      p0 = pt;
    }
  }
  if (ACC) {
    __shared__ double sh[kRkTx * TY];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kRkTx * TY / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = sh[0];
  }
}

// ---------------------------------------------------------------------------
// Bulk mean (Mean, src/mod_utils.cc:30-51) in FP64: per-block partial sums of
// f*dzflzi(k) over a chunk of rows of one plane, then one block combines the
// partials in a fixed order, so the result is the same on every run.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double s, double* sh) {
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  return sh[0];
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_mean_partial(Geom g, const T* __restrict__ f,
                                                              const double* __restrict__ dzflzi,
                                                              double* __restrict__ partial,
                                                              int rows_per_chunk) {
  __shared__ double sh[kRedThreads];
  const int k = blockIdx.y + 1;
  const int j0 = blockIdx.x * rows_per_chunk;
  const int j1 = min(j0 + rows_per_chunk, g.ny);
  const double wk = dzflzi[k];
  double s = 0.0;
  for (int j = j0; j < j1; ++j) {
    const T* row = f + g.idx(1, j + 1, k);
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) s += (double)row[i] * wk;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// out[0] = sum(partial)/nxg/nyg; when force_slot >= 0 also stores
// out[force_slot] = vel_force - mean (forceVel, src/mod_calcVel.cc:297).
__global__ __launch_bounds__(kRedThreads) void k_mean_final(const double* __restrict__ partial,
                                                            long n, double nxg, double nyg,
                                                            double* out, double vel_force,
                                                            double* force_out) {
  __shared__ double sh[kRedThreads];
  double s = 0.0;
  for (long t = threadIdx.x; t < n; t += blockDim.x) s += partial[t];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double mean = s / nxg / nyg;
    out[0] = mean;
    if (force_out) force_out[0] = vel_force - mean;
  }
}

// forceVel update (src/mod_calcVel.cc:299-315): interior += force.
template <typename T>
__global__ __launch_bounds__(kTx* kTy) void k_add_force(Geom g, T* __restrict__ f,
                                                        const double* __restrict__ force,
                                                        int kchunk) {
  const int i = blockIdx.x * kTx + threadIdx.x + 1;
  const int j = blockIdx.y * kTy + threadIdx.y + 1;
  if (i > g.nx || j > g.ny) return;
  const T fv = (T)force[0];
  const int k0 = blockIdx.z * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  for (int k = k0; k <= k1; ++k) {
    const long c = g.idx(i, j, k);
    f[c] += fv;
  }
}

// transform2CRF and other uniform shifts: interior += value.
template <typename T>
__global__ __launch_bounds__(kTx* kTy) void k_add_value(Geom g, T* __restrict__ f, T value,
                                                        int kchunk) {
  const int i = blockIdx.x * kTx + threadIdx.x + 1;
  const int j = blockIdx.y * kTy + threadIdx.y + 1;
  if (i > g.nx || j > g.ny) return;
  const int k0 = blockIdx.z * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  for (int k = k0; k <= k1; ++k) f[g.idx(i, j, k)] += value;
}

// ---------------------------------------------------------------------------
// Boundary update at one GPU.  The reference's sequence -- y self-exchange,
// x periodic copies over the whole y/z extent, then the z walls over the whole
// x/y extent (src/mod_updateBound.cc:229-296) -- leaves every halo cell equal
// to a closed form of interior values:
//   side halo (i or j in {0, n+1}, 1 <= k <= nz): f(wrap i, wrap j, k)
//   z planes: centred u, v: 2*bc - f(wrap i, wrap j, 1|nz); face w: bc on
//   planes 0, nz and nz+1 (src/mod_updateBound.cc:75-144); p: zero gradient.
// Each halo cell is written once from interior cells, so one launch covers
// all of them with no ordering hazard.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap(int i, int n) { return i == 0 ? n : (i == n + 1 ? 1 : i); }

// Perimeter cell q of an (nx+2)x(ny+2) plane -> (i, j).
__device__ __forceinline__ void perimeter(int q, int nx, int ny, int& i, int& j) {
  const int xl = nx + 2;
  if (q < xl) {
    i = q;
    j = 0;
  } else if (q < 2 * xl) {
    i = q - xl;
    j = ny + 1;
  } else {
    const int r = q - 2 * xl;  // rows 1..ny, two cells each
    j = r / 2 + 1;
    i = (r & 1) ? nx + 1 : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_vel(Geom g, T* __restrict__ u, T* __restrict__ v,
                                                T* __restrict__ w, T bc_u, int walls_only) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nper = 2 * (g.nx + 2) + 2 * g.ny;
  const long nside = walls_only ? 0 : (long)nper * g.nz;
  const long nplane = (long)(g.nx + 2) * (g.ny + 2);
  const T zero = T(0), two = T(2);
  if (t < nside) {
    const int k = (int)(t / nper) + 1;
    int i, j;
    perimeter((int)(t % nper), g.nx, g.ny, i, j);
    const long d = g.idx(i, j, k);
    const long s = g.idx(wrap(i, g.nx), wrap(j, g.ny), k);
    u[d] = u[s];
    v[d] = v[s];
    w[d] = (k == g.nz) ? zero : w[s];
  } else if (t < nside + nplane) {
    const long q = t - nside;
    const int i = (int)(q % (g.nx + 2));
    const int j = (int)(q / (g.nx + 2));
    // after the exchange the wall rows read the periodic image; alone
    // (imposeBCVel) they read the halo cells as they stand
    const int wi = walls_only ? i : wrap(i, g.nx), wj = walls_only ? j : wrap(j, g.ny);
    u[g.idx(i, j, 0)] = two * bc_u - u[g.idx(wi, wj, 1)];
    u[g.idx(i, j, g.nz + 1)] = two * bc_u - u[g.idx(wi, wj, g.nz)];
    v[g.idx(i, j, 0)] = two * zero - v[g.idx(wi, wj, 1)];
    v[g.idx(i, j, g.nz + 1)] = two * zero - v[g.idx(wi, wj, g.nz)];
    w[g.idx(i, j, 0)] = zero;
    w[g.idx(i, j, g.nz)] = zero;
    w[g.idx(i, j, g.nz + 1)] = zero;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_p(Geom g, T* __restrict__ p) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nper = 2 * (g.nx + 2) + 2 * g.ny;
  const long nside = (long)nper * g.nz;
  const long nplane = (long)(g.nx + 2) * (g.ny + 2);
  if (t < nside) {
    const int k = (int)(t / nper) + 1;
    int i, j;
    perimeter((int)(t % nper), g.nx, g.ny, i, j);
    p[g.idx(i, j, k)] = p[g.idx(wrap(i, g.nx), wrap(j, g.ny), k)];
  } else if (t < nside + nplane) {
    const long q = t - nside;
    const int i = (int)(q % (g.nx + 2));
    const int j = (int)(q / (g.nx + 2));
    const int wi = wrap(i, g.nx), wj = wrap(j, g.ny);
    p[g.idx(i, j, 0)] = p[g.idx(wi, wj, 1)];
    p[g.idx(i, j, g.nz + 1)] = p[g.idx(wi, wj, g.nz)];
  }
}

// ---------------------------------------------------------------------------
// Initial condition (src/mod_initFlow.cc:47-102): u = profile(k) + noise,
// v = w = noise.  Noise: std::minstd_rand (a = 48271, m = 2^31-1) seeded with
// seed and advanced L+1 steps, L the interior linear index.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mulmod31(uint64_t a, uint64_t b) {
  return (a * b) % 2147483647ULL;
}

__device__ double noise_uniform(uint32_t seed, uint64_t offset) {
  uint64_t x = seed % 2147483647ULL;
  if (x == 0) x = 1;
  uint64_t base = 48271ULL, e = offset + 1, r = 1;
  while (e) {
    if (e & 1) r = mulmod31(r, base);
    base = mulmod31(base, base);
    e >>= 1;
  }
  x = mulmod31(x, r);
  return (double)(x - 1) / 2147483646.0;
}

template <typename T>
__global__ __launch_bounds__(kTx* kTy) void k_init(Geom g, const double* __restrict__ prof,
                                                   T* __restrict__ u, T* __restrict__ v,
                                                   T* __restrict__ w, uint32_t s0, uint32_t s1,
                                                   uint32_t s2, int with_noise, double noise,
                                                   int kchunk) {
  const int i = blockIdx.x * kTx + threadIdx.x + 1;
  const int j = blockIdx.y * kTy + threadIdx.y + 1;
  if (i > g.nx || j > g.ny) return;
  const int k0 = blockIdx.z * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  for (int k = k0; k <= k1; ++k) {
    const long c = g.idx(i, j, k);
    double uu = prof[k], vv = 0.0, ww = 0.0;
    if (with_noise) {
      const uint64_t L = ((uint64_t)(k - 1) * g.ny + (j - 1)) * g.nx + (i - 1);
      uu += 2.0 * (noise_uniform(s0, L) - 0.5) * noise;
      vv += 2.0 * (noise_uniform(s1, L) - 0.5) * noise;
      ww += 2.0 * (noise_uniform(s2, L) - 0.5) * noise;
    }
    u[c] = (T)uu;
    v[c] = (T)vv;
    w[c] = (T)ww;
  }
}

// u *= 1/(mean*u_ref) when mean != 0 (src/mod_initFlow.cc:91-100).
template <typename T>
__global__ __launch_bounds__(kTx* kTy) void k_scale_inv_mean(Geom g, T* __restrict__ f,
                                                             const double* __restrict__ mean,
                                                             double u_ref, int kchunk) {
  const int i = blockIdx.x * kTx + threadIdx.x + 1;
  const int j = blockIdx.y * kTy + threadIdx.y + 1;
  if (i > g.nx || j > g.ny) return;
  const double m = mean[0];
  if (m == 0.0) return;
  const double mult = 1.0 / (m * u_ref);
  const int k0 = blockIdx.z * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  for (int k = k0; k <= k1; ++k) {
    const long c = g.idx(i, j, k);
    f[c] = (T)((double)f[c] * mult);
  }
}

// ---------------------------------------------------------------------------
// Runtime checks (src/mod_utils.cc:77-217).  Max with location: a candidate
// replaces the running best only when strictly larger (NaN never does), and
// ties between threads go to the smaller linear index, which reproduces the
// reference's k-j-i loop's first maximum.
// ---------------------------------------------------------------------------
struct ValIdx {
  double v;
  long l;
};

__device__ __forceinline__ bool better(double va, long la, double vb, long lb) {
  return va > vb || (va == vb && la < lb);
}

template <typename T, int KIND>
__global__ __launch_bounds__(kRedThreads) void k_max_partial(Geom g, const T* __restrict__ u,
                                                             const T* __restrict__ v,
                                                             const T* __restrict__ w,
                                                             const T* __restrict__ dzf_inv,
                                                             T dxi, T dyi, ValIdx* out) {
  __shared__ double sv[kRedThreads];
  __shared__ long sl[kRedThreads];
  const long n = (long)g.nx * g.ny * g.nz;
  double best = 0.0;
  long bl = 0;
  for (long L = (long)blockIdx.x * blockDim.x + threadIdx.x; L < n;
       L += (long)gridDim.x * blockDim.x) {
    const int i = (int)(L % g.nx) + 1;
    const int j = (int)((L / g.nx) % g.ny) + 1;
    const int k = (int)(L / ((long)g.nx * g.ny)) + 1;
    const long c = g.idx(i, j, k);
    T val;
    if (KIND == 0) {  // divergence, src/mod_utils.cc:177-179
      val = fabs((u[c] - u[c - 1]) * dxi + (v[c] - v[c - g.sx]) * dyi +
                 (w[c] - w[c - g.sxy]) * dzf_inv[k]);
    } else {  // CFL, src/mod_utils.cc:121-123
      val = fabs((u[c] + u[c - 1]) * T(0.5)) * dxi + fabs((v[c] + v[c - g.sx]) * T(0.5)) * dyi +
            fabs((w[c] + w[c - g.sxy]) * T(0.5)) * dzf_inv[k];
    }
    const double dv = (double)val;
    if (dv > best) {
      best = dv;
      bl = L;
    }
  }
  sv[threadIdx.x] = best;
  sl[threadIdx.x] = bl;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h &&
        better(sv[threadIdx.x + h], sl[threadIdx.x + h], sv[threadIdx.x], sl[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + h];
      sl[threadIdx.x] = sl[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = ValIdx{sv[0], sl[0]};
}

__global__ __launch_bounds__(kRedThreads) void k_max_final(const ValIdx* __restrict__ in, int n,
                                                           ValIdx* out) {
  __shared__ double sv[kRedThreads];
  __shared__ long sl[kRedThreads];
  double best = 0.0;
  long bl = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    if (better(in[t].v, in[t].l, best, bl)) {
      best = in[t].v;
      bl = in[t].l;
    }
  sv[threadIdx.x] = best;
  sl[threadIdx.x] = bl;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h &&
        better(sv[threadIdx.x + h], sl[threadIdx.x + h], sv[threadIdx.x], sl[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + h];
      sl[threadIdx.x] = sl[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = ValIdx{sv[0], sl[0]};
}

// NaN count over the whole halo-padded array.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_nan_partial(const T* __restrict__ f, long n,
                                                             double* out) {
  __shared__ double sh[kRedThreads];
  double cnt = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x)
    if (isnan(f[t])) cnt += 1.0;
  cnt = block_sum(cnt, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = cnt;
}

// ---------------------------------------------------------------------------
// host-side helpers
// ---------------------------------------------------------------------------
inline int kchunk_for(const pll_context* c) {
  // enough blocks to fill 256 CUs several times over, k-chunks of >= 4 planes
  const long cols = (long)((c->g.nx + kTx - 1) / kTx) * ((c->g.ny + kTy - 1) / kTy);
  long want = 4096 / (cols > 0 ? cols : 1);
  if (want < 1) want = 1;
  long kc = (c->g.nz + want - 1) / want;
  if (kc < 4) kc = 4;
  if (kc > c->g.nz) kc = c->g.nz;
  return (int)kc;
}

inline dim3 col_grid(const pll_context* c, int kc) {
  return dim3((c->g.nx + kTx - 1) / kTx, (c->g.ny + kTy - 1) / kTy, (c->g.nz + kc - 1) / kc);
}

template <typename T>
inline MeshT<T> mesh_of(const pll_context* c) {
  const T* base = static_cast<const T*>(c->d_mesh);
  const long n = c->g.nz + 2;
  return MeshT<T>{base, base + n, base + 2 * n, base + 3 * n};
}

template <typename T>
inline T* F(pll_context* c, int f) {
  return static_cast<T*>(c->fld[f]);
}

// Tiled launch shape of the stencil kernels: ntx*nty tiles of 64 x ty columns,
// each marching a chunk of k planes; about 2048 blocks in total.
struct TileGrid {
  int ntx, nty, nkc, kchunk;
  unsigned blocks() const { return (unsigned)ntx * nty * nkc; }
};

inline TileGrid tile_grid(const pll_context* c, int ty = kRkTy) {
  TileGrid t;
  t.ntx = (c->g.nx + kRkTx - 1) / kRkTx;
  t.nty = (c->g.ny + ty - 1) / ty;
  const long cols = (long)t.ntx * t.nty;
  long want = (2048 + cols - 1) / cols;
  int kc = (int)((c->g.nz + want - 1) / want);
  if (kc < 4) kc = 4;
  if (kc > c->g.nz) kc = c->g.nz;
  t.kchunk = kc;
  t.nkc = (c->g.nz + kc - 1) / kc;
  return t;
}

inline int forced_mask(const pll_context* c) {
  return (c->is_forced[0] ? 1 : 0) | (c->is_forced[1] ? 2 : 0) | (c->is_forced[2] ? 4 : 0);
}

// Per-plane coefficients of the FP32-mixed operator for stage `stage`
// (S = dt or dt/2), computed in FP64 and rounded once.
RkUniform rk_uniform(const pll_context* c, int stage) {
  const double S = stage == 1 ? c->dt : 0.5 * c->dt;
  return RkUniform{(float)(S * 0.25 * c->dx_inv), (float)(S * 0.25 * c->dy_inv),
                   (float)(S * c->re_inv * c->dx_inv * c->dx_inv),
                   (float)(S * c->re_inv * c->dy_inv * c->dy_inv),
                   (float)(S * 0.5 * c->u_crf * c->dx_inv)};
}

int rk_coef_setup(pll_context* c) {
  const int nz = c->g.nz;
  std::vector<RkCoef> h(2 * (size_t)(nz + 2));
  for (int stage = 1; stage <= 2; ++stage) {
    const double S = stage == 1 ? c->dt : 0.5 * c->dt, rei = c->re_inv;
    for (int k = 1; k <= nz; ++k) {
      const double* f = c->dzf.data();
      const double* fi = c->dzf_inv.data();
      const double* ci = c->dzc_inv.data();
      RkCoef& q = h[(stage - 1) * (size_t)(nz + 2) + k];
      q.a1 = (float)(f[k + 1] * ci[k]);
      q.a2 = (float)(f[k] * ci[k]);
      q.b1 = (float)(f[k - 1] * ci[k - 1]);
      q.b2 = (float)(f[k] * ci[k - 1]);
      q.zuv = (float)(S * 0.25 * fi[k]);
      q.cp = (float)(S * rei * ci[k] * fi[k]);
      q.cm = (float)(S * rei * ci[k - 1] * fi[k]);
      q.zw = (float)(S * 0.25 * ci[k]);
      q.dp = (float)(S * rei * fi[k + 1] * ci[k]);
      q.dm = (float)(S * rei * fi[k] * ci[k]);
      // the diagonal from the rounded FP32 coefficients (one rounding off their sum)
      const RkUniform cu = rk_uniform(c, stage);
      const double vxy = 2.0 * (double)cu.vx + 2.0 * (double)cu.vy;
      q.cdu = (float)(vxy + (double)q.cp + (double)q.cm);
      q.cdw = (float)(vxy + (double)q.dp + (double)q.dm);
    }
  }
  PLL_HIP_TRY(ctx_malloc(c, &c->d_rkcoef, h.size() * sizeof(RkCoef)));
  PLL_HIP_TRY(hipMemcpy(c->d_rkcoef, h.data(), h.size() * sizeof(RkCoef), hipMemcpyHostToDevice));
  return PLL_OK;
}

template <typename T>
int rk_t(pll_context* c, int stage, int fold, int corr, int acc, int bcload) {
  const int ty = std::is_same<T, double>::value ? kLateTy : kRkTy;
  const TileGrid t = tile_grid(c, ty);
  if (acc && (long)t.blocks() > c->npartial)
    return fail(PLL_ERR_STATE, "reduction scratch smaller than the stencil grid");
  // the projected loads read p at the periodic image of halo cells, which only
  // the folded (image-addressed) loads provide
  if (corr && !fold) return fail(PLL_ERR_STATE, "deferred projection needs folded loads");
  const double* dzflzi = c->d_mesh64 + 3 * (c->g.nz + 2);
  const T dtdxi = (T)(c->dt * c->dx_inv), dtdyi = (T)(c->dt * c->dy_inv);  // src/mod_calcVel.cc:256-258
  const MeshT<T> m = mesh_of<T>(c);
  const T bc_u = (T)(0.0 - c->u_crf);  // src/mod_updateBound.cc:80
  const double* force = c->d_scalar + 3;
  const RkCoef* coef = static_cast<const RkCoef*>(c->d_rkcoef) + (stage - 1) * (c->g.nz + 2);
  const RkUniform cu = rk_uniform(c, stage);
  const int crf = c->u_crf != 0.0;
  // FP64 runs the late-commit form of the stages (k_rk_late), FP32 k_rk
  constexpr bool kF64 = std::is_same<T, double>::value;
  if (stage == 1) {
    TimerScope ts(c, K_RK1);
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(t.blocks()), dim3(kRkTx * ty), 0, c->stream, c->g,
                         F<T>(c, PLL_U), F<T>(c, PLL_V), F<T>(c, PLL_W), F<T>(c, PLL_U1),
                         F<T>(c, PLL_V1), F<T>(c, PLL_W1), m, (T)c->dx_inv, (T)c->dy_inv,
                         (T)c->re_inv, (T)c->dt, (T)c->u_crf, t.kchunk, t.ntx, t.nty, fold, 0,
                         bc_u, force, forced_mask(c), coef, cu, crf, F<T>(c, PLL_P), corr,
                         dtdxi, dtdyi, 0, dzflzi, c->d_partial, fold);
    };
    // the deferred projection is compiled into its own variant of stage 1
    if constexpr (kF64)
      go(corr ? k_rk_late<1, true> : k_rk_late<1, false>);
    else
      go(corr ? k_rk<T, 1, true> : k_rk<T, 1>);
  } else {
    TimerScope ts(c, K_RK2);
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(t.blocks()), dim3(kRkTx * ty), 0, c->stream, c->g,
                         F<T>(c, PLL_U1), F<T>(c, PLL_V1), F<T>(c, PLL_W1), F<T>(c, PLL_U),
                         F<T>(c, PLL_V), F<T>(c, PLL_W), m, (T)c->dx_inv, (T)c->dy_inv,
                         (T)c->re_inv, (T)c->dt, (T)c->u_crf, t.kchunk, t.ntx, t.nty, 0, fold,
                         bc_u, force, forced_mask(c), coef, cu, crf, F<T>(c, PLL_P), corr,
                         dtdxi, dtdyi, acc, dzflzi, c->d_partial, bcload);
    };
    if constexpr (kF64)
      go(k_rk_late<2, false>);
    else
      go(k_rk<T, 2>);
  }
  PLL_HIP_TRY(hipGetLastError());
  if (acc) {
    // forceVel's bulk mean and force of u (src/mod_calcVel.cc:287,297)
    TimerScope ts(c, K_MEAN_FINAL);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial,
                       (long)t.blocks(), (double)c->g.nx, (double)c->g.ny, c->d_scalar + 0,
                       c->vel_force[0], c->d_scalar + 3);
    PLL_HIP_TRY(hipGetLastError());
  }
  return PLL_OK;
}

// RK stage 2 + calcRHS into the spare velocity and pressure buffers, then
// swap them in.  corr: the previous step's projection is applied on load;
// acc: the bulk mean of the new u and the force of u go to d_scalar[0], [3].
template <typename T>
int rk2_rhs_t(pll_context* c, int defer, int corr, int acc) {
  constexpr int kTall = std::is_same<T, double>::value ? kRhsTallTy : kRkTy;
  const bool tall = (long)tile_grid(c, kTall).blocks() >= kRhsTallBlocks;
  const int ty = tall ? kTall : kRkTy;
  const TileGrid t = tile_grid(c, ty);
  if (acc && (long)t.blocks() > c->npartial)
    return fail(PLL_ERR_STATE, "reduction scratch smaller than the stencil grid");
  if (!c->fld_alt[3]) return fail(PLL_ERR_STATE, "no spare pressure buffer");
  // the projection reads p at the periodic image of halo cells, which only the
  // folded loads of stage 2 provide
  if (corr && !defer) return fail(PLL_ERR_STATE, "deferred projection needs folded loads");
  auto kernel = corr ? (tall ? k_rk2_rhs<T, kTall, true> : k_rk2_rhs<T, kRkTy, true>)
                     : (tall ? k_rk2_rhs<T, kTall, false> : k_rk2_rhs<T, kRkTy, false>);
  const MeshT<T> m = mesh_of<T>(c);
  const T bc_u = (T)(0.0 - c->u_crf);  // src/mod_updateBound.cc:80
  const RkCoef* coef = static_cast<const RkCoef*>(c->d_rkcoef) + (c->g.nz + 2);
  const RkUniform cu = rk_uniform(c, 2);
  const double dti = 1.0 / c->dt;  // src/mod_calcRHS.cc:11-13
  const double dtidxi = dti * c->dx_inv;
  const double dtidyi = dti * c->dy_inv;
  const double pdx = c->dt * c->dx_inv, pdy = c->dt * c->dy_inv;  // src/mod_calcVel.cc:256-258
  const double* dzflzi = c->d_mesh64 + 3 * (c->g.nz + 2);
  {
    TimerScope ts(c, K_RK2_RHS);
    hipLaunchKernelGGL(kernel, dim3(t.blocks()), dim3(kRkTx * ty), 0, c->stream, c->g,
                       F<T>(c, PLL_U1), F<T>(c, PLL_V1), F<T>(c, PLL_W1), F<T>(c, PLL_U),
                       F<T>(c, PLL_V), F<T>(c, PLL_W), static_cast<T*>(c->fld_alt[0]),
                       static_cast<T*>(c->fld_alt[1]), static_cast<T*>(c->fld_alt[2]),
                       F<T>(c, PLL_P), static_cast<T*>(c->fld_alt[3]), m, (T)c->dx_inv,
                       (T)c->dy_inv, (T)c->re_inv, (T)c->dt, (T)c->u_crf, t.kchunk, t.ntx, t.nty,
                       defer, bc_u, c->d_scalar + 3, forced_mask(c), coef, cu,
                       (int)(c->u_crf != 0.0), (T)dti, (T)dtidxi, (T)dtidyi, (T)pdx, (T)pdy, acc,
                       dzflzi, c->d_partial);
  }
  PLL_HIP_TRY(hipGetLastError());
  for (int d = 0; d < 3; ++d) std::swap(c->fld[PLL_U + d], c->fld_alt[d]);
  std::swap(c->fld[PLL_P], c->fld_alt[3]);
  if (acc) {
    // forceVel's bulk mean and force of u (src/mod_calcVel.cc:287,297)
    TimerScope ts(c, K_MEAN_FINAL);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial,
                       (long)t.blocks(), (double)c->g.nx, (double)c->g.ny, c->d_scalar + 0,
                       c->vel_force[0], c->d_scalar + 3);
    PLL_HIP_TRY(hipGetLastError());
  }
  return PLL_OK;
}

// Both stages of an FP32-mixed step inside a run (the previous step's
// projection, forcing and boundary update applied on load) into the spare
// velocity buffers, then swapped in; acc: the bulk mean of the new u and the
// force of u go to d_scalar[0], [3].
constexpr int kRk12Ty = 8;  // tile rows of the fused stages
int rk12_f32(pll_context* c, int acc) {
  const TileGrid t = tile_grid(c, kRk12Ty);
  if (acc && (long)t.blocks() > c->npartial)
    return fail(PLL_ERR_STATE, "reduction scratch smaller than the stencil grid");
  if (!c->fld_alt[0] || !c->fld_alt[1] || !c->fld_alt[2])
    return fail(PLL_ERR_STATE, "no spare velocity buffers");
  const MeshT<float> m = mesh_of<float>(c);
  const RkCoef* coef = static_cast<const RkCoef*>(c->d_rkcoef);
  const double* dzflzi = c->d_mesh64 + 3 * (c->g.nz + 2);
  {
    TimerScope ts(c, K_RK12);
    hipLaunchKernelGGL(k_rk12<kRk12Ty>, dim3(t.blocks()), dim3(kRkTx * kRk12Ty), 0, c->stream, c->g,
                       F<float>(c, PLL_U), F<float>(c, PLL_V), F<float>(c, PLL_W),
                       F<float>(c, PLL_P), static_cast<float*>(c->fld_alt[0]),
                       static_cast<float*>(c->fld_alt[1]), static_cast<float*>(c->fld_alt[2]),
                       m.dzc_inv, (float)c->dt, (float)(c->dt * c->dx_inv),
                       (float)(c->dt * c->dy_inv),  // src/mod_calcVel.cc:256-258
                       (float)(0.0 - c->u_crf),     // src/mod_updateBound.cc:80
                       c->d_scalar + 3, forced_mask(c), coef, coef + (c->g.nz + 2),
                       rk_uniform(c, 1), rk_uniform(c, 2), (int)(c->u_crf != 0.0), t.kchunk,
                       t.ntx, t.nty, acc, dzflzi, c->d_partial);
  }
  PLL_HIP_TRY(hipGetLastError());
  for (int d = 0; d < 3; ++d) std::swap(c->fld[PLL_U + d], c->fld_alt[d]);
  if (acc) {
    // forceVel's bulk mean and force of u (src/mod_calcVel.cc:287,297)
    TimerScope ts(c, K_MEAN_FINAL);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial,
                       (long)t.blocks(), (double)c->g.nx, (double)c->g.ny, c->d_scalar + 0,
                       c->vel_force[0], c->d_scalar + 3);
    PLL_HIP_TRY(hipGetLastError());
  }
  return PLL_OK;
}

template <typename T>
int rhs_t(pll_context* c) {
  const TileGrid t = tile_grid(c);
  const double dti = 1.0 / c->dt;  // src/mod_calcRHS.cc:11-13
  const double dtidxi = dti * c->dx_inv;
  const double dtidyi = dti * c->dy_inv;
  TimerScope ts(c, K_RHS);
  hipLaunchKernelGGL((k_calc_rhs<T>), dim3(t.blocks()), dim3(kRkTx * kRkTy), 0, c->stream, c->g,
                     F<T>(c, PLL_U), F<T>(c, PLL_V), F<T>(c, PLL_W), F<T>(c, PLL_P),
                     mesh_of<T>(c).dzf_inv, (T)dti, (T)dtidxi, (T)dtidyi, t.kchunk, t.ntx, t.nty);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int correct_t(pll_context* c, bool with_mean) {
  constexpr int ty = correct_ty<T>();
  const TileGrid t = tile_grid(c, ty);
  const double dtdxi = c->dt * c->dx_inv;  // src/mod_calcVel.cc:256-258
  const double dtdyi = c->dt * c->dy_inv;
  const double* dzflzi = c->d_mesh64 + 3 * (c->g.nz + 2);
  if (with_mean && (long)t.blocks() > c->npartial)
    return fail(PLL_ERR_STATE, "reduction scratch smaller than the projection grid");
  {
    TimerScope ts(c, K_CORRECT);
    if (with_mean)
      hipLaunchKernelGGL((k_correct<T, true>), dim3(t.blocks()), dim3(kRkTx * ty), 0, c->stream,
                         c->g, F<T>(c, PLL_P), F<T>(c, PLL_U), F<T>(c, PLL_V), F<T>(c, PLL_W),
                         mesh_of<T>(c).dzc_inv, (T)dtdxi, (T)dtdyi, (T)c->dt, t.kchunk, t.ntx,
                         t.nty, dzflzi, c->d_partial);
    else
      hipLaunchKernelGGL((k_correct<T, false>), dim3(t.blocks()), dim3(kRkTx * ty), 0,
                         c->stream, c->g, F<T>(c, PLL_P), F<T>(c, PLL_U), F<T>(c, PLL_V),
                         F<T>(c, PLL_W), mesh_of<T>(c).dzc_inv, (T)dtdxi, (T)dtdyi, (T)c->dt,
                         t.kchunk, t.ntx, t.nty, dzflzi, c->d_partial);
  }
  if (with_mean) {
    // forceVel's bulk mean and force of u (src/mod_calcVel.cc:287,297)
    TimerScope ts(c, K_MEAN_FINAL);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial,
                       (long)t.blocks(), (double)c->g.nx, (double)c->g.ny, c->d_scalar + 0,
                       c->vel_force[0], c->d_scalar + 3);
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

// forceVel's update of u with the force already on the device.
template <typename T>
int add_force_u_t(pll_context* c) {
  const int kc = kchunk_for(c);
  TimerScope ts(c, K_FORCE);
  hipLaunchKernelGGL((k_add_force<T>), col_grid(c, kc), dim3(kTx, kTy), 0, c->stream, c->g,
                     F<T>(c, PLL_U), c->d_scalar + 3, kc);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int mean_t(pll_context* c, int field, double nxg, double nyg, double* d_out, double vel_force,
           double* d_force) {
  const int chunks = c->mean_chunks;
  const int rpc = (c->g.ny + chunks - 1) / chunks;
  if (chunks < 1 || (long)chunks * c->g.nz > c->npartial || !c->d_partial || !F<T>(c, field) || !d_out)
    return fail(PLL_ERR_STATE, "bulk-mean reduction scratch does not fit the grid");
  {
    TimerScope ts(c, K_MEAN_PARTIAL);
    hipLaunchKernelGGL((k_mean_partial<T>), dim3(chunks, c->g.nz), dim3(kRedThreads), 0,
                       c->stream, c->g, F<T>(c, field), c->d_mesh64 + 3 * (c->g.nz + 2),
                       c->d_partial, rpc);
  }
  {
    TimerScope ts(c, K_MEAN_FINAL);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial,
                       (long)chunks * c->g.nz, nxg, nyg, d_out, vel_force, d_force);
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int force_t(pll_context* c) {
  const int kc = kchunk_for(c);
  for (int d = 0; d < 3; ++d) {
    if (!c->is_forced[d]) continue;
    int rc = mean_t<T>(c, PLL_U + d, (double)c->g.nx, (double)c->g.ny, c->d_scalar + d,
                       c->vel_force[d], c->d_scalar + 3 + d);
    if (rc) return rc;
  }
  for (int d = 0; d < 3; ++d) {
    if (!c->is_forced[d]) continue;
    TimerScope ts(c, K_FORCE);
    hipLaunchKernelGGL((k_add_force<T>), col_grid(c, kc), dim3(kTx, kTy), 0, c->stream, c->g,
                       F<T>(c, PLL_U + d), c->d_scalar + 3 + d, kc);
  }
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

inline long bc_threads(const pll_context* c) {
  const long nper = 2L * (c->g.nx + 2) + 2L * c->g.ny;
  return nper * c->g.nz + (long)(c->g.nx + 2) * (c->g.ny + 2);
}

template <typename T>
int bc_vel_t(pll_context* c, int which, int walls_only) {
  const long n = walls_only ? (long)(c->g.nx + 2) * (c->g.ny + 2) : bc_threads(c);
  const int base = which ? PLL_U1 : PLL_U;
  const double bc_u = 0.0 - c->u_crf;  // src/mod_updateBound.cc:80
  TimerScope ts(c, K_BC_VEL);
  hipLaunchKernelGGL((k_bc_vel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->g, F<T>(c, base), F<T>(c, base + 1), F<T>(c, base + 2), (T)bc_u,
                     walls_only);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int bc_p_t(pll_context* c) {
  const long n = bc_threads(c);
  TimerScope ts(c, K_BC_P);
  hipLaunchKernelGGL((k_bc_p<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->g, F<T>(c, PLL_P));
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int add_scalar_t(pll_context* c, int field, double value) {
  const int kc = kchunk_for(c);
  TimerScope ts(c, K_MISC);
  hipLaunchKernelGGL((k_add_value<T>), col_grid(c, kc), dim3(kTx, kTy), 0, c->stream, c->g,
                     F<T>(c, field), (T)value, kc);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int init_t(pll_context* c, const double* d_prof, int with_noise, double noise,
           const uint32_t seeds[3]) {
  const int kc = kchunk_for(c);
  TimerScope ts(c, K_INIT);
  hipLaunchKernelGGL((k_init<T>), col_grid(c, kc), dim3(kTx, kTy), 0, c->stream, c->g, d_prof,
                     F<T>(c, PLL_U), F<T>(c, PLL_V), F<T>(c, PLL_W), seeds[0], seeds[1],
                     seeds[2], with_noise, noise, kc);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T>
int scale_inv_mean_t(pll_context* c, int field, const double* d_mean, double u_ref) {
  const int kc = kchunk_for(c);
  TimerScope ts(c, K_INIT);
  hipLaunchKernelGGL((k_scale_inv_mean<T>), col_grid(c, kc), dim3(kTx, kTy), 0, c->stream, c->g,
                     F<T>(c, field), d_mean, u_ref, kc);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}

template <typename T, int KIND>
int max_t(pll_context* c, pll_value_index* out) {
  ValIdx* red = static_cast<ValIdx*>(c->d_red);
  {
    TimerScope ts(c, K_CHECK);
    hipLaunchKernelGGL((k_max_partial<T, KIND>), dim3(kRedBlocks), dim3(kRedThreads), 0,
                       c->stream, c->g, F<T>(c, PLL_U), F<T>(c, PLL_V), F<T>(c, PLL_W),
                       mesh_of<T>(c).dzf_inv, (T)c->dx_inv, (T)c->dy_inv, red);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kRedThreads), 0, c->stream, red, kRedBlocks,
                       red + kRedBlocks);
  }
  PLL_HIP_TRY(hipGetLastError());
  ValIdx h;
  PLL_HIP_TRY(hipMemcpyAsync(&h, red + kRedBlocks, sizeof(ValIdx), hipMemcpyDeviceToHost,
                             c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  const long nxy = (long)c->g.nx * c->g.ny;
  out->value = (KIND == 1) ? h.v * (double)(T)c->dt : h.v;
  out->rank = 0;
  out->ig = (int)(h.l % c->g.nx) + 1;
  out->jg = (int)((h.l / c->g.nx) % c->g.ny) + 1;
  out->kg = (int)(h.l / nxy) + 1;
  return PLL_OK;
}

template <typename T>
int nan_t(pll_context* c, int field, int* passed) {
  double* red = static_cast<double*>(c->d_red);
  {
    TimerScope ts(c, K_CHECK);
    hipLaunchKernelGGL((k_nan_partial<T>), dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream,
                       F<T>(c, field), c->nfield, red);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, red,
                       (long)kRedBlocks, 1.0, 1.0, red + kRedBlocks, 0.0, (double*)nullptr);
  }
  PLL_HIP_TRY(hipGetLastError());
  double cnt = 0.0;
  PLL_HIP_TRY(hipMemcpyAsync(&cnt, red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  *passed = (cnt == 0.0) ? 1 : 0;
  return PLL_OK;
}

}  // namespace

// ---- dispatch on the context precision ------------------------------------------
#define PLL_DISPATCH(c, call_fp64, call_fp32) \
  ((c)->prec == PLL_FP64 ? (call_fp64) : (call_fp32))

int rk_setup(pll_context* c) { return c->prec == PLL_FP64 ? PLL_OK : rk_coef_setup(c); }

int launch_rk(pll_context* c, int stage, int fold, int corr, int acc, int bcload) {
  if (c->closure) return launch_rk_closure(c, stage, fold, corr, acc, bcload);
  return PLL_DISPATCH(c, rk_t<double>(c, stage, fold, corr, acc, bcload),
                      rk_t<float>(c, stage, fold, corr, acc, bcload));
}
int launch_calc_rhs(pll_context* c) { return PLL_DISPATCH(c, rhs_t<double>(c), rhs_t<float>(c)); }
int launch_force_from_partials(pll_context* c, long n) {
  if (n < 1 || n > c->npartial || !c->d_partial)
    return fail(PLL_ERR_STATE, "bulk-mean partials do not fit the scratch");
  TimerScope ts(c, K_MEAN_FINAL);
  hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kRedThreads), 0, c->stream, c->d_partial, n,
                     (double)c->g.nx, (double)c->g.ny, c->d_scalar + 0, c->vel_force[0],
                     c->d_scalar + 3);
  PLL_HIP_TRY(hipGetLastError());
  return PLL_OK;
}
int launch_rk12(pll_context* c, int acc) {
  if (c->prec != PLL_FP32_MIXED) return fail(PLL_ERR_STATE, "the fused stages are the FP32-mixed step's");
  return rk12_f32(c, acc);
}
int launch_rk2_rhs(pll_context* c, int defer, int corr, int acc) {
  if (c->prec != PLL_FP64) return fail(PLL_ERR_STATE, "the fused stage 2 + calcRHS is the FP64 step's");
  return rk2_rhs_t<double>(c, defer, corr, acc);
}
int launch_correct_vel(pll_context* c, bool with_mean) {
  return PLL_DISPATCH(c, correct_t<double>(c, with_mean), correct_t<float>(c, with_mean));
}
int launch_add_force_u(pll_context* c) {
  return PLL_DISPATCH(c, add_force_u_t<double>(c), add_force_u_t<float>(c));
}
int launch_mean(pll_context* c, int field, double nxg, double nyg, double* d_out) {
  return PLL_DISPATCH(c, mean_t<double>(c, field, nxg, nyg, d_out, 0.0, nullptr),
                      mean_t<float>(c, field, nxg, nyg, d_out, 0.0, nullptr));
}
int launch_force(pll_context* c) { return PLL_DISPATCH(c, force_t<double>(c), force_t<float>(c)); }
int launch_bc_vel(pll_context* c, int which, int walls_only) {
  return PLL_DISPATCH(c, bc_vel_t<double>(c, which, walls_only),
                      bc_vel_t<float>(c, which, walls_only));
}
int launch_bc_p(pll_context* c) { return PLL_DISPATCH(c, bc_p_t<double>(c), bc_p_t<float>(c)); }
int launch_add_scalar(pll_context* c, int field, double value) {
  return PLL_DISPATCH(c, add_scalar_t<double>(c, field, value),
                      add_scalar_t<float>(c, field, value));
}
int launch_init_flow(pll_context* c, const double* d_prof, int with_noise, double noise,
                     const uint32_t seeds[3]) {
  return PLL_DISPATCH(c, init_t<double>(c, d_prof, with_noise, noise, seeds),
                      init_t<float>(c, d_prof, with_noise, noise, seeds));
}
int launch_scale_by_inverse_mean(pll_context* c, int field, const double* d_mean, double u_ref) {
  return PLL_DISPATCH(c, scale_inv_mean_t<double>(c, field, d_mean, u_ref),
                      scale_inv_mean_t<float>(c, field, d_mean, u_ref));
}
int launch_max_div(pll_context* c, pll_value_index* out) {
  return PLL_DISPATCH(c, (max_t<double, 0>(c, out)), (max_t<float, 0>(c, out)));
}
int launch_max_cfl(pll_context* c, pll_value_index* out) {
  return PLL_DISPATCH(c, (max_t<double, 1>(c, out)), (max_t<float, 1>(c, out)));
}
int launch_check_nan(pll_context* c, int field, int* passed) {
  return PLL_DISPATCH(c, nan_t<double>(c, field, passed), nan_t<float>(c, field, passed));
}

}  // namespace pll
