// mod_modules.cc -- the reference's module routines over the C-ABI
// (src/mod_calcVel.cc, src/mod_calcRHS.cc, src/mod_poisson_solver.cc,
// src/mod_updateBound.cc, src/mod_utils.cc).
//
// The device context integrates fixed field slots: RK1 reads (u, v, w) and
// writes (u1, v1, w1), RK2 the reverse, calcRHS writes p.  A routine called
// with fields in other roles, or with a u_crf other than the context's,
// throws std::invalid_argument instead of computing something else.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "mod_calcRHS.hh"
#include "mod_calcVel.hh"
#include "mod_device.hh"
#include "mod_parameters.hh"
#include "mod_poisson_solver.hh"
#include "mod_updateBound.hh"
#include "mod_utils.hh"

namespace {

void roles(const char* fn, const Array3D& a, int sa, const Array3D& b, int sb, const Array3D& c,
           int sc) {
  if (a.slot() != sa || b.slot() != sb || c.slot() != sc)
    throw std::invalid_argument(std::string(fn) + ": fields passed in roles the device context "
                                                  "does not integrate");
}

void same_crf(const char* fn, fp crf) {
  pll_config cfg;
  mod_device::check(pll_get_config(mod_device::context(), &cfg), "pll_get_config");
  if (crf != cfg.u_crf)
    throw std::invalid_argument(std::string(fn) + ": u_crf differs from the device context's");
}

}  // namespace

// ---------------------------------------------------------------- mod_calcVel
namespace mod_calcVel {

void transform2CRF(fp vel_crf, Array3DH3& vel, fp& vel_force) {
  if (vel.slot() != PLL_U) throw std::invalid_argument("transform2CRF: expects the context's u");
  mod_device::check(pll_transform2crf(mod_device::context(), vel_crf), "transform2CRF");
  mod_device::sync();
  if (vel_crf != 0) vel_force -= vel_crf;
}

void timeIntVelRK1(fp crf, const Array3DH3& u, const Array3DH3& v, const Array3DH3& w,
                   Array3DH3& u1, Array3DH3& v1, Array3DH3& w1) {
  roles("timeIntVelRK1", u, PLL_U, v, PLL_V, w, PLL_W);
  roles("timeIntVelRK1", u1, PLL_U1, v1, PLL_V1, w1, PLL_W1);
  same_crf("timeIntVelRK1", crf);
  mod_device::check(pll_time_int_vel_rk1(mod_device::context()), "timeIntVelRK1");
  mod_device::sync();
}

void timeIntVelRK2(fp crf, Array3DH3& u, Array3DH3& v, Array3DH3& w, const Array3DH3& u1,
                   const Array3DH3& v1, const Array3DH3& w1) {
  roles("timeIntVelRK2", u, PLL_U, v, PLL_V, w, PLL_W);
  roles("timeIntVelRK2", u1, PLL_U1, v1, PLL_V1, w1, PLL_W1);
  same_crf("timeIntVelRK2", crf);
  mod_device::check(pll_time_int_vel_rk2(mod_device::context()), "timeIntVelRK2");
  mod_device::sync();
}

void correctVel(const Array3DH1& p, Array3DH3& u, Array3DH3& v, Array3DH3& w) {
  roles("correctVel", u, PLL_U, v, PLL_V, w, PLL_W);
  if (p.slot() != PLL_P) throw std::invalid_argument("correctVel: expects the context's p");
  mod_device::check(pll_correct_vel(mod_device::context()), "correctVel");
  mod_device::sync();
}

void forceVel(Array3DH3& u, Array3DH3& v, Array3DH3& w) {
  roles("forceVel", u, PLL_U, v, PLL_V, w, PLL_W);
  double mean[3];
  mod_device::check(pll_force_vel(mod_device::context(), mean), "forceVel");
}

}  // namespace mod_calcVel

// ---------------------------------------------------------------- mod_calcRHS
namespace mod_calcRHS {

void calcRHS(const Array3DH3& u, const Array3DH3& v, const Array3DH3& w, Array3DH1& rhs) {
  roles("calcRHS", u, PLL_U, v, PLL_V, w, PLL_W);
  if (rhs.slot() != PLL_P) throw std::invalid_argument("calcRHS: expects the context's p");
  mod_device::check(pll_calc_rhs(mod_device::context()), "calcRHS");
  mod_device::sync();
}

}  // namespace mod_calcRHS

// ------------------------------------------------------------- poisson solver

void init_poisson_solver(int nx_global, int ny_global, int nz_global, double dx_, double dy_,
                         double* dzf_global, const char bctype_x[], const char bctype_y[],
                         const char bctype_z[], std::array<std::array<int, 6>, 3>& neighbor_xyz) {
  (void)neighbor_xyz;  // one rank: every neighbour is the rank itself or a wall
  // src/mod_fft.cc:20-26 supports only periodic x and y; the channel's walls
  // are Neumann for the pressure (src/main.cc:56).
  if (std::strncmp(bctype_x, "PP", 2) != 0 || std::strncmp(bctype_y, "PP", 2) != 0)
    throw std::invalid_argument("init_poisson_solver: x and y must be periodic (\"PP\")");
  if (std::strncmp(bctype_z, "NN", 2) != 0)
    throw std::invalid_argument("init_poisson_solver: z must be Neumann at both walls (\"NN\")");
  pll_config cfg;
  pll_config_default(&cfg);
  cfg.nx = nx_global;
  cfg.ny = ny_global;
  cfg.nz = nz_global;
  cfg.lx = dx_ * nx_global;
  cfg.ly = dy_ * ny_global;
  cfg.lz = lz;
  cfg.re = re;
  cfg.u_ref = u_ref;
  cfg.l_ref = l_ref;
  cfg.dt = dt;
  cfg.u_crf = u_crf;
  for (int d = 0; d < 3; ++d) {
    cfg.is_forced[d] = is_forced[d] ? 1 : 0;
    cfg.vel_force[d] = vel_force[d];
  }
  cfg.precision = precision;
  cfg.device = device_id;
  mod_device::create(cfg, dzf_global);
}

void execute_poisson_solver(Array3DH1& p) {
  if (p.slot() != PLL_P) throw std::invalid_argument("execute_poisson_solver: expects the context's p");
  mod_device::check(pll_poisson_execute(mod_device::context()), "execute_poisson_solver");
  mod_device::sync();
}

void free_poisson_solver() { mod_device::destroy(); }

// ------------------------------------------------------------ mod_updateBound
namespace mod_updateBound {

void updateBoundVel(fp crf, Array3DH3& u, Array3DH3& v, Array3DH3& w) {
  same_crf("updateBoundVel", crf);
  int which;
  if (u.slot() == PLL_U && v.slot() == PLL_V && w.slot() == PLL_W)
    which = 0;
  else if (u.slot() == PLL_U1 && v.slot() == PLL_V1 && w.slot() == PLL_W1)
    which = 1;
  else
    throw std::invalid_argument("updateBoundVel: expects (u, v, w) or (u1, v1, w1)");
  mod_device::check(pll_update_bound_vel(mod_device::context(), which), "updateBoundVel");
  mod_device::sync();
}

void imposeBCVel(Array3DH3& u, Array3DH3& v, Array3DH3& w, fp crf) {
  same_crf("imposeBCVel", crf);
  int which;
  if (u.slot() == PLL_U && v.slot() == PLL_V && w.slot() == PLL_W)
    which = 0;
  else if (u.slot() == PLL_U1 && v.slot() == PLL_V1 && w.slot() == PLL_W1)
    which = 1;
  else
    throw std::invalid_argument("imposeBCVel: expects (u, v, w) or (u1, v1, w1)");
  mod_device::check(pll_impose_bc_vel(mod_device::context(), which), "imposeBCVel");
  mod_device::sync();
}

void updateBoundP(Array3DH1& p) {
  if (p.slot() != PLL_P) throw std::invalid_argument("updateBoundP: expects the context's p");
  mod_device::check(pll_update_bound_p(mod_device::context()), "updateBoundP");
  mod_device::sync();
}

}  // namespace mod_updateBound

// ------------------------------------------------------------------ mod_utils
namespace mod_utils {

fp Mean(int nx_global, int ny_global, const Array1DH1& dzflzi, const Array3DH3& var) {
  (void)dzflzi;
  double out = 0.0;
  mod_device::check(pll_mean(mod_device::context(), var.slot(), nx_global, ny_global, &out), "Mean");
  return out;
}

void initCheckCFLAndDiv(value_index_pair_t& cfl_max, value_index_pair_t& div_max) {
  div_max = cfl_max = value_index_pair_t{0.0, 0, 1, 1, 1};
}

void freeCheckCFLAndDiv() {}

bool checkNaN(const Array3DH3& var, std::string_view tag) {
  int passed = 0;
  mod_device::check(pll_check_nan(mod_device::context(), var.slot(), &passed), "checkNaN");
  if (!passed)
    std::printf("PowerLLEL.ERROR.checkNaN: NaN has been detected in <%.*s>!\n",
                static_cast<int>(tag.size()), tag.data());
  return passed != 0;
}

static void from_device(const pll_value_index& r, value_index_pair_t& out) {
  out.value = r.value;
  out.rank = r.rank;
  out.ig = r.ig;
  out.jg = r.jg;
  out.kg = r.kg;
}

void calcMaxCFL(fp dt_, fp dxi, fp dyi, const Array1DH1& dzfi, const Array3DH3& u,
                const Array3DH3& v, const Array3DH3& w, value_index_pair_t& cfl_max) {
  (void)dt_, (void)dxi, (void)dyi, (void)dzfi;  // the context's own dt and metrics
  roles("calcMaxCFL", u, PLL_U, v, PLL_V, w, PLL_W);
  pll_value_index r;
  mod_device::check(pll_calc_max_cfl(mod_device::context(), &r), "calcMaxCFL");
  from_device(r, cfl_max);
}

bool checkCFL(fp limit, const value_index_pair_t& cfl_max) {
  if (cfl_max.value > limit || std::isnan(cfl_max.value)) {
    std::printf("PowerLLEL.ERROR.checkCFL: Maximum CFL number exceeds the threshold specified by "
                "the user!\n");
    return false;
  }
  return true;
}

void calcMaxDiv(fp dxi, fp dyi, const Array1DH1& dzfi, const Array3DH3& u, const Array3DH3& v,
                const Array3DH3& w, value_index_pair_t& div_max) {
  (void)dxi, (void)dyi, (void)dzfi;
  roles("calcMaxDiv", u, PLL_U, v, PLL_V, w, PLL_W);
  pll_value_index r;
  mod_device::check(pll_calc_max_div(mod_device::context(), &r), "calcMaxDiv");
  from_device(r, div_max);
}

bool checkDiv(fp limit, const value_index_pair_t& div_max) {
  if (div_max.value > limit || std::isnan(div_max.value)) {
    std::printf("PowerLLEL.ERROR.checkDiv: Maximum divergence exceeds the threshold specified by "
                "the user!\n");
    return false;
  }
  return true;
}

}  // namespace mod_utils
