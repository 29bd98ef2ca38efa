// pll_internal.hpp -- shared definitions of the HIP library behind include/pll.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pll.h"

namespace pll {

// Halo-padded field geometry: (nx+2)*(ny+2)*(nz+2), x fastest, logical 0..n+1.
struct Geom {
  int nx, ny, nz;
  long sx, sxy;
  __host__ __device__ long idx(int i, int j, int k) const { return i + sx * j + sxy * k; }
};

// Mesh coefficient arrays (length nz+2, index 0..nz+1) in the field precision.
template <typename T>
struct MeshT {
  const T* dzf;
  const T* dzf_inv;
  const T* dzc_inv;
  const T* dzflzi;
};

// Stockham plan of a complex transform of length L held in LDS.
constexpr int kMaxPass = 20;
struct FftPlan {
  int L;
  int npass;
  int radix[kMaxPass];
  int ns[kMaxPass];     // span before the pass (product of earlier radices)
  int twoff[kMaxPass];  // offset of the pass' twiddles: tw[off + k*(R-1) + r-1]
};

// Blocks are dispatched round-robin over the 8 XCDs of the GPU; this gives each
// XCD a contiguous range of logical blocks, so neighbouring tiles share its L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n) {
  const unsigned q = n / 8, r = n % 8, x = b % 8, l = b / 8;
  return x * q + (x < r ? x : r) + l;
}

// Row chunks per plane of the bulk-mean reduction: the largest power of two
// <= ny with chunks*nz reaching 2048 partial sums (fewer when ny is small).
// Fewer than 4096 partials unless nz alone exceeds that (then nz).
inline int mean_chunks_for(int ny, int nz) {
  int chunks = 1;
  while (chunks * 2 <= ny && (long)chunks * nz < 2048) chunks *= 2;
  return chunks;
}

// Kernel identifiers for the per-kernel timers.
enum KernelId {
  K_RK1 = 0,
  K_RK2,
  K_RHS,
  K_FFT_X_FWD,
  K_FFT_Y_FWD,
  K_TRID,
  K_FFT_Y_BWD,
  K_FFT_X_BWD,
  K_CORRECT,
  K_MEAN_PARTIAL,
  K_MEAN_FINAL,
  K_FORCE,
  K_BC_VEL,
  K_BC_P,
  K_CHECK,
  K_INIT,
  K_MISC,
  K_RK2_RHS,        // stage 2 fused with calcRHS
  K_FFT_X_FWD_RHS,  // calcRHS fused into the forward x transform
  K_RK_CLOSURE,     // RK stage with the learned eddy viscosity
  K_RK12,           // both RK stages of the FP32-mixed step
  K_COUNT
};
extern const char* const kKernelNames[K_COUNT];

// Pending event pair of one timed launch.
struct PendingTimer {
  int kid;
  hipEvent_t start, stop;
};

}  // namespace pll

struct pll_context {
  pll_config cfg;
  pll::Geom g;
  long nfield;  // halo-padded element count
  long nspec;   // nx*ny*nz
  int prec;     // pll_precision
  size_t esz;   // field element size in bytes
  hipStream_t stream = nullptr;

  void* fld[PLL_NFIELD] = {};
  void* fld_alt[4] = {};  // spare u, v, w, p (FP64): the fused stage 2 writes here, then the roles swap
  void* spec = nullptr;

  // host mesh in FP64 (index 0..nz+1)
  std::vector<double> dzf_global, dzf, dzf_inv, dzc, dzc_inv, dzflzi, zc_global;
  double dx, dy, dx_inv, dy_inv, re_inv, dt, u_crf;
  double vel_force[3];
  int is_forced[3];

  void* d_mesh = nullptr;      // 4*(nz+2) in field precision: dzf, dzf_inv, dzc_inv, dzflzi
  void* d_rkcoef = nullptr;    // FP32-mixed stencil coefficients, 2 stages x (nz+2)
  double* d_mesh64 = nullptr;  // same, FP64

  // tridiagonal solve: a, c (nz), pre-factored pivot table [k][cy][cx]
  int cxn = 0, cyn = 0;  // nx/2+1, ny/2+1
  double* d_a = nullptr;
  double* d_c = nullptr;
  double* d_piv = nullptr;  // FP64 pivots (FP32-mixed: their reciprocals), cxn*cyn*nz
  // the partitioned solve of the step: b (nz), compact eigenvalues (cxn, cyn)
  // and the pivots of the planes below its segments (k_trid.hip)
  double* d_seg = nullptr;
  int seg_pin = 0;  // the (0,0) column's last pivot is pinned to 0

  // FFT plans and twiddles (in field precision)
  pll::FftPlan plan_x;  // complex length nx/2
  pll::FftPlan plan_y;  // complex length ny
  void* d_tw_x = nullptr;
  void* d_tw_y = nullptr;
  void* d_rtw_x = nullptr;  // exp(-2 pi i k/nx), k = 0..nx/2
  void* d_twl_x = nullptr;  // exp(-2 pi i t/L), L = nx/2 (fixed-length kernels)
  void* d_twl_y = nullptr;  // exp(-2 pi i t/L), L = ny
  bool x_fixed = false, y_fixed = false;
  int rhs_bcfold = 0;  // calcRHS in the x transform reads halos through the boundary rules
  int x_batch = 1, y_tile = 2;

  // reductions
  double* d_partial = nullptr;
  long npartial = 0;
  int mean_chunks = 1;
  double* d_scalar = nullptr;  // [0..2] means, [3..5] forces, [6..] scratch
  void* d_red = nullptr;       // check-reduction scratch
  long nred = 0;

  // statistics: FP64 per-point sums of the 11 sampled quantities (interior,
  // nz*ny*nx, x fastest); null when not selected
  double* d_stat[11] = {};

  // learned closure (not owned): when attached, the RK stages add the
  // eddy-viscosity stress with the network evaluated in the stage kernel
  pll_closure* closure = nullptr;
  long closure_version = -1;  // the closure's weight version this stream has waited for
  void* d_feat = nullptr;   // feature coefficients per plane (field precision), 3*(nz+2)
  float* d_nut = nullptr;   // eddy viscosity of the last closure stage, nx*ny*nz

  // device bytes allocated for this context (ctx_malloc)
  size_t dev_bytes = 0;

  // timers
  bool timing = false;
  std::vector<pll::PendingTimer> pending;
  std::vector<hipEvent_t> event_pool;
  double timer_ms[pll::K_COUNT] = {};
  long timer_launches[pll::K_COUNT] = {};
};

namespace pll {

// ---- error reporting --------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
const char* last_error_cstr();

#define PLL_HIP_TRY(expr)                                  \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::pll::hip_fail(_e, #expr); \
  } while (0)

// hipMalloc of a buffer the context owns, counted in pll_device_bytes.
template <typename P>
hipError_t ctx_malloc(pll_context* c, P** p, size_t bytes) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e == hipSuccess) c->dev_bytes += bytes;
  return e;
}

// ---- timers -------------------------------------------------------------------
void timer_begin(pll_context* c, int kid, PendingTimer* out);
void timer_end(pll_context* c, PendingTimer* t);
int timer_collect(pll_context* c);

struct TimerScope {
  pll_context* c;
  PendingTimer t;
  bool on;
  TimerScope(pll_context* ctx, int kid) : c(ctx), on(ctx->timing) {
    if (on) timer_begin(c, kid, &t);
  }
  ~TimerScope() {
    if (on) timer_end(c, &t);
  }
};

// ---- launchers (implemented in the .hip files) -------------------------------------
// stencils and pointwise kernels (k_stencil.hip)
// fold: stage 1 applies the previous step's deferred force and rebuilds the
// halos on load; stage 2 adds the deferred force to its point value of u.
// corr: apply the deferred projection on load; acc (stage 2): bulk mean of the
// new u into d_scalar[0], force into d_scalar[3]
// bcload (stage 2, FP32): rebuild the source's halos on load (u1's boundary
// update not run)
int launch_rk(pll_context* c, int stage, int fold, int corr = 0, int acc = 0, int bcload = 0);
// the stage of a context with an attached closure (k_closure.hip); fold,
// corr, acc, bcload as for launch_rk
int launch_rk_closure(pll_context* c, int stage, int fold = 0, int corr = 0, int acc = 0,
                      int bcload = 0);
// the closure stages' bulk-mean partials fit d_partial (the fused step sequence)
bool closure_fused_available(const pll_context* c);
// forceVel's bulk mean of u and its force from n FP64 block partials in
// d_partial, into d_scalar[0] and d_scalar[3] (k_stencil.hip)
int launch_force_from_partials(pll_context* c, long n);

int rk_setup(pll_context* c);  // per-plane coefficients of the FP32-mixed stencil
int launch_calc_rhs(pll_context* c);
// RK stage 2 fused with calcRHS (FP64 step sequence only): writes the new u, v,
// w and the right-hand side to the spare buffers and swaps them in; defer adds
// the previous step's force, corr applies its projection on load, acc reduces
// the bulk mean of the new u and the force of u into d_scalar[0], [3].
int launch_rk2_rhs(pll_context* c, int defer, int corr = 0, int acc = 0);
int launch_rk12(pll_context* c, int acc);
// with_mean: also reduce the bulk mean of the corrected u and store the force
// of u (vel_force[0] - mean) on the device.
int launch_correct_vel(pll_context* c, bool with_mean);
int launch_add_force_u(pll_context* c);
int launch_mean(pll_context* c, int field, double nxg, double nyg, double* d_out);
int launch_force(pll_context* c);
int launch_bc_vel(pll_context* c, int which, int walls_only = 0);
int launch_bc_p(pll_context* c);
int launch_add_scalar(pll_context* c, int field, double value);
int launch_init_flow(pll_context* c, const double* d_prof, int with_noise, double noise,
                     const uint32_t seeds[3]);
int launch_scale_by_inverse_mean(pll_context* c, int field, const double* d_mean, double u_ref);
int launch_max_div(pll_context* c, pll_value_index* out);
int launch_max_cfl(pll_context* c, pll_value_index* out);
int launch_check_nan(pll_context* c, int field, int* passed);

// transforms and the tridiagonal solve (k_fft.hip, k_trid.hip)
int fft_setup(pll_context* c);
int launch_fft_x_fwd(pll_context* c);
// calcRHS fused into the forward x transform (reads u, v, w; p is not written)
bool fft_x_rhs_available(const pll_context* c);
int launch_fft_x_fwd_rhs(pll_context* c);
// inverse x transform inside the step (x halos of p written with their periodic values)
int launch_fft_x_bwd_step(pll_context* c);
int launch_fft_x_bwd(pll_context* c);
int launch_fft_y_fwd(pll_context* c);
int launch_fft_y_bwd(pll_context* c);
int trid_setup(pll_context* c);
int launch_trid(pll_context* c);
// the solve inside the step: columns partitioned into register-held segments
int launch_trid_step(pll_context* c);
// The pressure solve of the step (src/mod_poisson_solver.cc:454-510) from p
// (or, with rhs, from calcRHS computed in the x transform's loads) back to p
// with its periodic x halos.
int launch_poisson_step(pll_context* c, bool rhs);

}  // namespace pll
