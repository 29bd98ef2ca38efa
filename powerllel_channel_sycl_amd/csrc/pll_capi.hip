// pll_capi.hip -- the C-ABI of include/pll.h: context lifetime, field
// transfer, and one entry point per reference routine of the hot path.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pll_internal.hpp"

namespace pll {

const char* const kKernelNames[K_COUNT] = {
    "rk1",         "rk2",       "calc_rhs",   "fft_x_fwd",    "fft_y_fwd", "solve_trid",
    "fft_y_bwd",   "fft_x_bwd", "correct_vel", "mean_partial", "mean_final", "force_add",
    "bc_vel",      "bc_p",      "check",      "init",         "misc",        "rk2_rhs",
    "fft_x_fwd_rhs", "rk_closure", "rk12"};

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
  return e == hipErrorOutOfMemory ? PLL_ERR_NOMEM : PLL_ERR_HIP;
}

// ---- timers --------------------------------------------------------------------
static hipEvent_t take_event(pll_context* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timer_begin(pll_context* c, int kid, PendingTimer* t) {
  t->kid = kid;
  t->start = take_event(c);
  t->stop = take_event(c);
  if (t->start) (void)hipEventRecord(t->start, c->stream);
}

void timer_end(pll_context* c, PendingTimer* t) {
  if (!t->start || !t->stop) return;
  (void)hipEventRecord(t->stop, c->stream);
  c->pending.push_back(*t);
}

int timer_collect(pll_context* c) {
  if (c->pending.empty()) return PLL_OK;
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto& t : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      c->timer_ms[t.kid] += ms;
      c->timer_launches[t.kid] += 1;
    }
    c->event_pool.push_back(t.start);
    c->event_pool.push_back(t.stop);
  }
  c->pending.clear();
  return PLL_OK;
}

}  // namespace pll

using namespace pll;

namespace {

bool factorable(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

int check_field(const pll_context* c, int f) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (f < 0 || f >= PLL_NFIELD) return fail(PLL_ERR_ARG, "field id out of range");
  return PLL_OK;
}

template <typename T>
int upload_as(pll_context* c, void* dst, const double* host, long n) {
  if (sizeof(T) == sizeof(double)) {
    PLL_HIP_TRY(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PLL_HIP_TRY(hipStreamSynchronize(c->stream));
    return PLL_OK;
  }
  std::vector<T> tmp(n);
  for (long t = 0; t < n; ++t) tmp[t] = (T)host[t];
  PLL_HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

template <typename T>
int download_as(pll_context* c, const void* src, double* host, long n) {
  if (sizeof(T) == sizeof(double)) {
    PLL_HIP_TRY(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PLL_HIP_TRY(hipStreamSynchronize(c->stream));
    return PLL_OK;
  }
  std::vector<T> tmp(n);
  PLL_HIP_TRY(hipMemcpyAsync(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  for (long t = 0; t < n; ++t) host[t] = (double)tmp[t];
  return PLL_OK;
}

int upload(pll_context* c, void* dst, const double* host, long n) {
  return c->prec == PLL_FP64 ? upload_as<double>(c, dst, host, n) : upload_as<float>(c, dst, host, n);
}
int download(pll_context* c, const void* src, double* host, long n) {
  return c->prec == PLL_FP64 ? download_as<double>(c, src, host, n)
                             : download_as<float>(c, src, host, n);
}

void release(pll_context* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& t : c->pending) {
    c->event_pool.push_back(t.start);
    c->event_pool.push_back(t.stop);
  }
  for (hipEvent_t e : c->event_pool)
    if (e) (void)hipEventDestroy(e);
  for (void*& p : c->fld)
    if (p) (void)hipFree(p);
  for (void*& p : c->fld_alt)
    if (p) (void)hipFree(p);
  for (double*& p : c->d_stat)
    if (p) (void)hipFree(p);
  void* bufs[] = {c->spec,   c->d_mesh, c->d_mesh64, c->d_a,    c->d_c,       c->d_piv, c->d_seg,
                  c->d_tw_x, c->d_tw_y,    c->d_rtw_x, c->d_twl_x, c->d_twl_y, c->d_rkcoef,
                  c->d_partial, c->d_scalar, c->d_red, c->d_feat, c->d_nut};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Mesh coefficients of initMesh at one rank (src/mod_mesh.cc:26-65).
void init_mesh(pll_context* c, const double* dzf_global) {
  const int nz = c->g.nz;
  c->dx = c->cfg.lx / c->cfg.nx;
  c->dy = c->cfg.ly / c->cfg.ny;
  c->dx_inv = 1.0 / c->dx;
  c->dy_inv = 1.0 / c->dy;
  c->dzf_global.assign(dzf_global, dzf_global + nz + 2);
  c->zc_global.assign(nz + 2, 0.0);
  c->zc_global[0] = -0.5 * c->dzf_global[0];
  for (int k = 1; k <= nz + 1; ++k)
    c->zc_global[k] = c->zc_global[k - 1] + 0.5 * (c->dzf_global[k - 1] + c->dzf_global[k]);
  c->dzf.assign(nz + 2, 0.0);
  c->dzf_inv.assign(nz + 2, 0.0);
  c->dzc.assign(nz + 2, 0.0);
  c->dzc_inv.assign(nz + 2, 0.0);
  c->dzflzi.assign(nz + 2, 0.0);
  for (int k = 0; k <= nz + 1; ++k) {
    c->dzf[k] = c->dzf_global[k];
    c->dzf_inv[k] = 1.0 / c->dzf[k];
  }
  for (int k = 0; k <= nz; ++k) {
    c->dzc[k] = 0.5 * (c->dzf[k] + c->dzf[k + 1]);
    c->dzc_inv[k] = 1.0 / c->dzc[k];
  }
  for (int k = 0; k <= nz + 1; ++k) c->dzflzi[k] = c->dzf[k] / c->cfg.lz;
}

template <typename T>
int upload_mesh(pll_context* c) {
  const long n = c->g.nz + 2;
  std::vector<T> m(4 * n);
  std::vector<double> m64(4 * n);
  for (long k = 0; k < n; ++k) {
    m64[k] = c->dzf[k];
    m64[n + k] = c->dzf_inv[k];
    m64[2 * n + k] = c->dzc_inv[k];
    m64[3 * n + k] = c->dzflzi[k];
  }
  for (long t = 0; t < 4 * n; ++t) m[t] = (T)m64[t];
  PLL_HIP_TRY(ctx_malloc(c, &c->d_mesh, m.size() * sizeof(T)));
  PLL_HIP_TRY(ctx_malloc(c, &c->d_mesh64, m64.size() * sizeof(double)));
  PLL_HIP_TRY(hipMemcpy(c->d_mesh, m.data(), m.size() * sizeof(T), hipMemcpyHostToDevice));
  PLL_HIP_TRY(hipMemcpy(c->d_mesh64, m64.data(), m64.size() * sizeof(double), hipMemcpyHostToDevice));
  return PLL_OK;
}

int create_impl(const pll_config* cfg, const double* dzf_global, pll_context** out) {
  if (!cfg || !dzf_global || !out) return fail(PLL_ERR_ARG, "null argument to pll_create");
  *out = nullptr;
  if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 2 || cfg->nx % 2 || cfg->ny % 2 || cfg->nz % 2)
    return fail(PLL_ERR_ARG, "nx, ny, nz must be even integers (nx, ny >= 4, nz >= 2)");
  if (!factorable(cfg->nx / 2) || !factorable(cfg->ny))
    return fail(PLL_ERR_ARG, "nx/2 and ny must be products of 2, 3 and 5");
  if (cfg->precision != PLL_FP64 && cfg->precision != PLL_FP32_MIXED)
    return fail(PLL_ERR_ARG, "unknown precision mode");
  if (!(cfg->dt > 0) || !(cfg->re > 0) || !(cfg->lx > 0) || !(cfg->ly > 0) || !(cfg->lz > 0))
    return fail(PLL_ERR_ARG, "dt, re and the box lengths must be positive");
  for (int k = 0; k < cfg->nz + 2; ++k)
    if (!(dzf_global[k] > 0)) return fail(PLL_ERR_ARG, "dzf_global must be positive");
  int ndev = 0;
  PLL_HIP_TRY(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(PLL_ERR_ARG, "device ordinal out of range");
  PLL_HIP_TRY(hipSetDevice(cfg->device));

  pll_context* c = new pll_context();
  c->cfg = *cfg;
  c->g.nx = cfg->nx;
  c->g.ny = cfg->ny;
  c->g.nz = cfg->nz;
  c->g.sx = cfg->nx + 2;
  c->g.sxy = (long)(cfg->nx + 2) * (cfg->ny + 2);
  c->nfield = c->g.sxy * (cfg->nz + 2);
  c->nspec = (long)cfg->nx * cfg->ny * cfg->nz;
  c->prec = cfg->precision;
  c->esz = c->prec == PLL_FP64 ? sizeof(double) : sizeof(float);
  c->re_inv = cfg->u_ref * cfg->l_ref / cfg->re;  // src/mod_parameters.cc:259
  c->dt = cfg->dt;
  c->u_crf = cfg->u_crf;
  for (int d = 0; d < 3; ++d) {
    c->vel_force[d] = cfg->vel_force[d];
    c->is_forced[d] = cfg->is_forced[d] ? 1 : 0;
  }
  int rc = PLL_OK;
  auto bail = [&](int code) {
    release(c);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreateWithFlags"));
  init_mesh(c, dzf_global);
  for (int f = 0; f < PLL_NFIELD; ++f) {
    e = ctx_malloc(c, &c->fld[f], c->nfield * c->esz);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(field)"));
    e = hipMemsetAsync(c->fld[f], 0, c->nfield * c->esz, c->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMemsetAsync(field)"));
  }
  // spare u, v, w (and p for the FP64 step's fused stage 2 + calcRHS)
  const int nspare = c->prec == PLL_FP64 ? 4 : 3;
  for (int d = 0; d < nspare; ++d) {
    e = ctx_malloc(c, &c->fld_alt[d], c->nfield * c->esz);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(spare field)"));
    e = hipMemsetAsync(c->fld_alt[d], 0, c->nfield * c->esz, c->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMemsetAsync(spare field)"));
  }
  e = ctx_malloc(c, &c->spec, c->nspec * c->esz);
  if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(spec)"));
  rc = c->prec == PLL_FP64 ? upload_mesh<double>(c) : upload_mesh<float>(c);
  if (rc) return bail(rc);
  // reduction scratch: mean partials (chunks x nz), scalars, check partials
  const int chunks = mean_chunks_for(c->g.ny, c->g.nz);
  c->mean_chunks = chunks;
  c->npartial = std::max<long>((long)chunks * c->g.nz, 1L << 15);
  if ((e = ctx_malloc(c, &c->d_partial, c->npartial * sizeof(double))) != hipSuccess)
    return bail(hip_fail(e, "hipMalloc(partial)"));
  if ((e = ctx_malloc(c, &c->d_scalar, 64 * sizeof(double))) != hipSuccess)
    return bail(hip_fail(e, "hipMalloc(scalar)"));
  if ((e = hipMemsetAsync(c->d_scalar, 0, 64 * sizeof(double), c->stream)) != hipSuccess)
    return bail(hip_fail(e, "hipMemsetAsync(scalar)"));
  c->nred = 2048 * 16;
  if ((e = ctx_malloc(c, &c->d_red, c->nred)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(red)"));
  if ((rc = fft_setup(c))) return bail(rc);
  if ((rc = trid_setup(c))) return bail(rc);
  if ((rc = rk_setup(c))) return bail(rc);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return bail(hip_fail(e, "hipStreamSynchronize"));
  *out = c;
  return PLL_OK;
}

int poisson_impl(pll_context* c) {
  int rc;
  if ((rc = launch_fft_x_fwd(c))) return rc;
  if ((rc = launch_fft_y_fwd(c))) return rc;
  if ((rc = launch_trid(c))) return rc;
  if ((rc = launch_fft_y_bwd(c))) return rc;
  return launch_fft_x_bwd(c);
}

// One step of src/main.cc:116-154.  When only u is forced (or nothing is),
// within a multi-step run the projection, the force update and the final
// boundary update of a step (:145-153) are folded into the next step's RK loads
// (`corr`, `fold`) and the bulk mean of the forcing is reduced by stage 2 from
// the predictor; the last step of a run applies them explicitly, so every run
// ends in the reference's state.
int finish_step(pll_context* c, bool fused, bool last);
int finish_deferred(pll_context* c, bool last);

int step_impl(pll_context* c, bool first, bool last) {
  int rc;
  const bool fused = !c->is_forced[1] && !c->is_forced[2];
  const int fold = (fused && !first) ? 1 : 0;
  if (c->closure && fused && fft_x_rhs_available(c) && closure_fused_available(c)) {
    // the closure stages (eddy viscosity evaluated on every cell of the
    // stage's source) in a deferred sequence: the previous step's projection,
    // forcing and boundary update (:145-153) in one pass before stage 1, u1's
    // boundary update written by stage 1, the predictor's bulk mean for the
    // forcing summed by stage 2, and calcRHS inside the forward x transform
    // (which rebuilds u's boundary on load)
    const int corr = fold;
    const int acc = (!last && c->is_forced[0]) ? 1 : 0;
    if ((rc = launch_rk(c, 1, fold, corr, 0, 1))) return rc;  // :145-153, :117, :121
    if ((rc = launch_rk(c, 2, 0, 0, acc, 0))) return rc;      // :125
    c->rhs_bcfold = 1;
    rc = launch_poisson_step(c, true);  // :133-137
    c->rhs_bcfold = 0;
    if (rc) return rc;
    // inside a run p's boundary update (:141) is left to the next step's
    // deferred pass, which reads p's halos through their rules
    return last ? finish_deferred(c, last) : PLL_OK;
  }
  if (c->closure) {
    // the closure stages (eddy viscosity evaluated in the stage kernel), then
    // the reference sequence src/main.cc:121-153 unfused
    if ((rc = launch_rk(c, 1, 0))) return rc;  // :117
    if ((rc = launch_bc_vel(c, 1))) return rc;  // :121
    if ((rc = launch_rk(c, 2, 0))) return rc;  // :125
    if ((rc = launch_bc_vel(c, 0))) return rc;  // :129
    if ((rc = launch_calc_rhs(c))) return rc;   // :133
    if ((rc = launch_poisson_step(c, false))) return rc;  // :137
    return finish_step(c, false, last);
  }
  if (fused && c->prec != PLL_FP64 && fft_x_rhs_available(c)) {
    // FP32-mixed: the projection of step n (:145), its forcing (:149) and
    // boundary update (:153) are applied by step n+1's stage loads (corr);
    // stage 2 sums the bulk mean of the predictor for that forcing, and
    // calcRHS runs inside the forward x transform's loads (:133-137).
    // The boundary updates of u1 (:121) and u (:129) are not run either:
    // stage 2 and calcRHS rebuild the halos they read from the boundary rules.
    // Inside a run both stages run in one kernel (u1 stays in LDS); the
    // first step (nothing deferred) and the last (u1 is left in its arrays)
    // run them separately.
    const int corr = fold;
    const int acc = (!last && c->is_forced[0]) ? 1 : 0;
    if (fold && !last) {
      if ((rc = launch_rk12(c, acc))) return rc;  // :117-125
    } else {
      if ((rc = launch_rk(c, 1, fold, corr, 0))) return rc;       // :117
      if ((rc = launch_rk(c, 2, fold, corr, acc, 1))) return rc;  // :125
    }
    c->rhs_bcfold = 1;
    rc = launch_poisson_step(c, true);  // :133-137
    c->rhs_bcfold = 0;
    if (rc) return rc;
    return finish_deferred(c, last);
  }
  if (fused && c->prec == PLL_FP64) {
    // FP64: the projection of step n (:145) and its forcing (:149) are applied
    // by step n+1's loads (corr, fold); :121-133 run in one kernel: the u1
    // boundary is rebuilt on load, the middle boundary update only matters for
    // calcRHS (computed in the same kernel) and for w(nz) = 0 (written there).
    // That kernel sums the predictor's bulk mean for the forcing.
    const int corr = fold;
    const int acc = (!last && c->is_forced[0]) ? 1 : 0;
    if ((rc = launch_rk(c, 1, fold, corr))) return rc;        // :117
    if ((rc = launch_rk2_rhs(c, fold, corr, acc))) return rc;  // :121-133
    if ((rc = launch_poisson_step(c, false))) return rc;       // :137
    return finish_deferred(c, last);
  }
  // unfused, or FP32-mixed without the calcRHS transform
  if ((rc = launch_rk(c, 1, fold))) return rc;  // src/main.cc:117
  if ((rc = launch_bc_vel(c, 1))) return rc;    // :121
  if ((rc = launch_rk(c, 2, fold))) return rc;  // :125
  if ((rc = launch_bc_vel(c, 0))) return rc;    // :129
  if ((rc = launch_calc_rhs(c))) return rc;     // :133
  if ((rc = launch_poisson_step(c, false))) return rc;  // :137
  return finish_step(c, fused, last);
}

// src/main.cc:141 after the pressure solve of a step whose projection, forcing
// and final boundary update (:145-153) the next step's loads apply; the last
// step of a run refreshes u1's halos (its boundary update :121 was not run)
// and applies them explicitly.
int finish_deferred(pll_context* c, bool last) {
  int rc;
  if ((rc = launch_bc_p(c))) return rc;  // :141
  if (!last) return PLL_OK;
  if ((rc = launch_bc_vel(c, 1))) return rc;                          // :121, u1's halos
  if ((rc = launch_correct_vel(c, c->is_forced[0] != 0))) return rc;  // :145 (+ mean of :149)
  if (c->is_forced[0] && (rc = launch_add_force_u(c))) return rc;     // :149
  return launch_bc_vel(c, 0);                                          // :153
}

// src/main.cc:141-153 after the pressure solve.
int finish_step(pll_context* c, bool fused, bool last) {
  int rc;
  if ((rc = launch_bc_p(c))) return rc;           // :141
  if (!fused) {
    if ((rc = launch_correct_vel(c, false))) return rc;  // :145
    if ((rc = launch_force(c))) return rc;               // :149
    return launch_bc_vel(c, 0);                          // :153
  }
  if ((rc = launch_correct_vel(c, c->is_forced[0] != 0))) return rc;  // :145 (+ mean of :149)
  if (!last) return PLL_OK;
  if (c->is_forced[0] && (rc = launch_add_force_u(c))) return rc;     // :149
  return launch_bc_vel(c, 0);                                          // :153
}

}  // namespace

extern "C" {

int pll_abi_version(void) { return PLL_ABI_VERSION; }

const char* pll_last_error(void) { return last_error_cstr(); }

void pll_config_default(pll_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  // examples/template/param.json
  cfg->nx = cfg->ny = cfg->nz = 128;
  cfg->lx = 12.566370614359172;
  cfg->ly = 6.283185307179586;
  cfg->lz = 2.0;
  cfg->re = 2800.0;
  cfg->u_ref = 1.0;
  cfg->l_ref = 1.0;
  cfg->dt = 1.0e-3;
  cfg->u_crf = 0.0;
  cfg->is_forced[0] = 1;
  cfg->vel_force[0] = 1.0;
  cfg->precision = PLL_FP64;
  cfg->pin_singular = 1;
  cfg->device = 0;
}

int pll_mesh_by_func(int mesh_type, int stretch_func, double stretch_ratio, int nz, double lz,
                     double* dzf_global) {
  // initMeshByFunc, src/mod_mesh.cc:95-125
  if (nz < 1 || !dzf_global) return fail(PLL_ERR_ARG, "bad arguments to pll_mesh_by_func");
  const double pi = M_PI;
  if (stretch_ratio == 0.0 || mesh_type == 0) {
    for (int k = 0; k <= nz + 1; ++k) dzf_global[k] = lz / nz;
    return PLL_OK;
  }
  std::vector<double> zf(nz + 2, 0.0);
  for (int k = 1; k <= nz; ++k) {
    const double z0 = k / (1.0 * nz);
    if (stretch_func == 0)
      zf[k] = 0.5 * (1.0 + std::tanh((z0 - 0.5) * stretch_ratio) / std::tanh(stretch_ratio / 2.0));
    else
      zf[k] = 0.5 * (1.0 + std::sin((z0 - 0.5) * stretch_ratio * pi) /
                               std::sin(stretch_ratio * pi / 2.0));
  }
  for (int k = 1; k <= nz; ++k) dzf_global[k] = lz * (zf[k] - zf[k - 1]);
  dzf_global[0] = dzf_global[1];
  dzf_global[nz + 1] = dzf_global[nz];
  return PLL_OK;
}

int pll_eigen_values(int n, double* lambda) {
  // get_eigen_values for "PP", src/mod_fft.cc:196-222
  if (n < 2 || n % 2 || !lambda) return fail(PLL_ERR_ARG, "bad arguments to pll_eigen_values");
  const double pi = std::acos(-1.0);
  std::vector<double> aux(n);
  for (int i = 0; i < n; ++i) aux[i] = 2.0 * (std::cos(2.0 * pi * (i - 0.0) / n) - 1.0);
  lambda[0] = aux[0];
  lambda[1] = aux[n / 2];
  for (int i = 2; i < n; ++i) lambda[i] = aux[i / 2];
  return PLL_OK;
}

int pll_create(const pll_config* cfg, const double* dzf_global, pll_context** out) {
  return create_impl(cfg, dzf_global, out);
}

int pll_destroy(pll_context* ctx) {
  if (!ctx) return PLL_OK;
  (void)hipSetDevice(ctx->cfg.device);
  release(ctx);
  return PLL_OK;
}

int pll_get_config(const pll_context* ctx, pll_config* cfg) {
  if (!ctx || !cfg) return fail(PLL_ERR_ARG, "null argument");
  *cfg = ctx->cfg;
  cfg->vel_force[0] = ctx->vel_force[0];
  cfg->vel_force[1] = ctx->vel_force[1];
  cfg->vel_force[2] = ctx->vel_force[2];
  return PLL_OK;
}

int pll_device_bytes(const pll_context* c, size_t* bytes) {
  if (!c || !bytes) return fail(PLL_ERR_ARG, "null argument");
  *bytes = c->dev_bytes;
  return PLL_OK;
}

int pll_field_upload(pll_context* c, int field, const double* host) {
  int rc = check_field(c, field);
  if (rc) return rc;
  if (!host) return fail(PLL_ERR_ARG, "null host buffer");
  return upload(c, c->fld[field], host, c->nfield);
}

int pll_field_download(pll_context* c, int field, double* host) {
  int rc = check_field(c, field);
  if (rc) return rc;
  if (!host) return fail(PLL_ERR_ARG, "null host buffer");
  return download(c, c->fld[field], host, c->nfield);
}

int pll_field_device_ptr(pll_context* c, int field, void** dev, int* elem_bytes) {
  int rc = check_field(c, field);
  if (rc) return rc;
  if (dev) *dev = c->fld[field];
  if (elem_bytes) *elem_bytes = (int)c->esz;
  return PLL_OK;
}

int pll_spec_upload(pll_context* c, const double* host) {
  if (!c || !host) return fail(PLL_ERR_ARG, "null argument");
  return upload(c, c->spec, host, c->nspec);
}

int pll_spec_download(pll_context* c, double* host) {
  if (!c || !host) return fail(PLL_ERR_ARG, "null argument");
  return download(c, c->spec, host, c->nspec);
}

int pll_init_flow(pll_context* c, int profile, double u0, int with_noise, double noise_intensity,
                  const uint32_t seeds[3]) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (profile < PLL_PROFILE_UNI || profile > PLL_PROFILE_LOG) return fail(PLL_ERR_ARG, "unknown profile");
  const uint32_t def_seeds[3] = {42u, 43u, 44u};
  if (!seeds) seeds = def_seeds;
  // velocity profile per k (src/mod_initFlow.cc:16-30,47-74)
  const int nz = c->g.nz;
  std::vector<double> prof(nz + 2, 0.0);
  bool is_mean = false;
  if (profile == PLL_PROFILE_UNI) {
    for (int k = 1; k <= nz; ++k) prof[k] = u0;
  } else if (profile == PLL_PROFILE_POI) {
    for (int k = 1; k <= nz; ++k) {
      const double zn = c->zc_global[k] / c->cfg.lz;
      prof[k] = 6.0 * (1.0 - zn) * zn * c->cfg.u_ref;
    }
    is_mean = true;
  } else {
    const double retau = 0.09 * std::pow(c->cfg.re, 0.88);
    for (int k = 1; k <= nz; ++k) {
      const double zn = c->zc_global[k] / c->cfg.lz;
      const double z = std::fmin(zn, 1.0 - zn) * 2.0 * retau;
      prof[k] = (z <= 11.6 ? z : 2.5 * std::log(z) + 5.5);
    }
    is_mean = true;
  }
  double* d_prof = nullptr;
  PLL_HIP_TRY(hipMalloc(&d_prof, prof.size() * sizeof(double)));
  hipError_t e = hipMemcpyAsync(d_prof, prof.data(), prof.size() * sizeof(double),
                                hipMemcpyHostToDevice, c->stream);
  int rc = (e != hipSuccess) ? hip_fail(e, "hipMemcpyAsync(profile)") : PLL_OK;
  if (!rc) rc = launch_init_flow(c, d_prof, with_noise, noise_intensity, seeds);
  // src/mod_initFlow.cc:91-100: Mean(nz, ny, ...) -- nz where nx is meant
  if (!rc && is_mean) rc = launch_mean(c, PLL_U, (double)nz, (double)c->g.ny, c->d_scalar + 6);
  if (!rc && is_mean) rc = launch_scale_by_inverse_mean(c, PLL_U, c->d_scalar + 6, c->cfg.u_ref);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_prof);
  return rc;
}

int pll_transform2crf(pll_context* c, double vel_crf) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (vel_crf == 0) return PLL_OK;  // src/mod_calcVel.cc:14
  int rc = launch_add_scalar(c, PLL_U, -vel_crf);
  if (rc) return rc;
  c->vel_force[0] -= vel_crf;
  return PLL_OK;
}

int pll_time_int_vel_rk1(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return launch_rk(c, 1, 0);
}

int pll_time_int_vel_rk2(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return launch_rk(c, 2, 0);
}

int pll_correct_vel(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return launch_correct_vel(c, false);
}

int pll_force_vel(pll_context* c, double mean_out[3]) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  int rc = launch_force(c);
  if (rc || !mean_out) return rc;
  double m[3];
  PLL_HIP_TRY(hipMemcpyAsync(m, c->d_scalar, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  for (int d = 0; d < 3; ++d) mean_out[d] = c->is_forced[d] ? m[d] : 0.0;
  return PLL_OK;
}

int pll_calc_rhs(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return launch_calc_rhs(c);
}

int pll_poisson_execute(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return poisson_impl(c);
}

int pll_fft_fwd_xpen(pll_context* c) { return c ? launch_fft_x_fwd(c) : fail(PLL_ERR_ARG, "null context"); }
int pll_fft_fwd_ypen(pll_context* c) { return c ? launch_fft_y_fwd(c) : fail(PLL_ERR_ARG, "null context"); }
int pll_solve_trid(pll_context* c) { return c ? launch_trid(c) : fail(PLL_ERR_ARG, "null context"); }
int pll_fft_bwd_ypen(pll_context* c) { return c ? launch_fft_y_bwd(c) : fail(PLL_ERR_ARG, "null context"); }
int pll_fft_bwd_xpen(pll_context* c) { return c ? launch_fft_x_bwd(c) : fail(PLL_ERR_ARG, "null context"); }

int pll_update_bound_vel(pll_context* c, int which) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (which != 0 && which != 1) return fail(PLL_ERR_ARG, "which must be 0 (u,v,w) or 1 (u1,v1,w1)");
  return launch_bc_vel(c, which);
}

int pll_impose_bc_vel(pll_context* c, int which) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (which != 0 && which != 1) return fail(PLL_ERR_ARG, "which must be 0 (u,v,w) or 1 (u1,v1,w1)");
  return launch_bc_vel(c, which, 1);
}

int pll_update_bound_p(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  return launch_bc_p(c);
}

int pll_mean(pll_context* c, int field, int nx_global, int ny_global, double* out) {
  int rc = check_field(c, field);
  if (rc) return rc;
  if (!out || nx_global <= 0 || ny_global <= 0) return fail(PLL_ERR_ARG, "bad arguments to pll_mean");
  rc = launch_mean(c, field, (double)nx_global, (double)ny_global, c->d_scalar + 7);
  if (rc) return rc;
  PLL_HIP_TRY(hipMemcpyAsync(out, c->d_scalar + 7, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

int pll_calc_max_div(pll_context* c, pll_value_index* out) {
  if (!c || !out) return fail(PLL_ERR_ARG, "null argument");
  return launch_max_div(c, out);
}

int pll_calc_max_cfl(pll_context* c, pll_value_index* out) {
  if (!c || !out) return fail(PLL_ERR_ARG, "null argument");
  return launch_max_cfl(c, out);
}

int pll_check_nan(pll_context* c, int field, int* passed) {
  int rc = check_field(c, field);
  if (rc) return rc;
  if (!passed) return fail(PLL_ERR_ARG, "null argument");
  return launch_check_nan(c, field, passed);
}

int pll_step(pll_context* c, int nsteps) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (nsteps < 0) return fail(PLL_ERR_ARG, "nsteps must be >= 0");
  for (int s = 0; s < nsteps; ++s) {
    int rc = step_impl(c, s == 0, s == nsteps - 1);
    if (rc) return rc;
  }
  return PLL_OK;
}

int pll_synchronize(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  PLL_HIP_TRY(hipStreamSynchronize(c->stream));
  return PLL_OK;
}

int pll_timing_enable(pll_context* c, int on) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  int rc = timer_collect(c);
  c->timing = on != 0;
  return rc;
}

int pll_timing_reset(pll_context* c) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  int rc = timer_collect(c);
  for (int k = 0; k < K_COUNT; ++k) {
    c->timer_ms[k] = 0.0;
    c->timer_launches[k] = 0;
  }
  return rc;
}

int pll_timing_count(pll_context* c, int* nkernels) {
  if (!c || !nkernels) return fail(PLL_ERR_ARG, "null argument");
  *nkernels = K_COUNT;
  return PLL_OK;
}

int pll_timing_read(pll_context* c, int index, const char** name, double* total_ms, long* launches) {
  if (!c) return fail(PLL_ERR_ARG, "null context");
  if (index < 0 || index >= K_COUNT) return fail(PLL_ERR_ARG, "kernel index out of range");
  int rc = timer_collect(c);
  if (rc) return rc;
  if (name) *name = kKernelNames[index];
  if (total_ms) *total_ms = c->timer_ms[index];
  if (launches) *launches = c->timer_launches[index];
  return PLL_OK;
}

}  // extern "C"
