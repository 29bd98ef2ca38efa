// main.cc -- the channel-flow driver (src/main.cc:24-240) on one GPU.
//
//   pll_channel [param.json]
//   pll_channel --check [param.json]   parse, check and print the run set-up
//                                      as JSON (no GPU needed)
//   pll_channel --mesh [param.json]    parse, build the mesh and write its
//                                      files as initialisation does (no GPU)
//
// Same initialisation order, step sequence, checks and screen output as the
// reference.  With "fused_step" (default) the steps between two checks or
// outputs run as one pll_step call, which performs the identical sequence
// with fused kernels; with "fused_step": false every module routine is
// called one by one, exactly as src/main.cc:116-154 does.
#include <chrono>
#include <cstdio>
#include <exception>
#include <string>

#include "mod_calcRHS.hh"
#include "mod_calcVel.hh"
#include "mod_dataIO.hh"
#include "mod_device.hh"
#include "mod_initFlow.hh"
#include "mod_mesh.hh"
#include "mod_parameters.hh"
#include "mod_poisson_solver.hh"
#include "mod_statistics.hh"
#include "mod_updateBound.hh"
#include "mod_utils.hh"
#include "mod_variables.hh"

using namespace mod_variables;

namespace {

void step_modules() {
  mod_calcVel::timeIntVelRK1(u_crf, u, v, w, u1, v1, w1);
  mod_updateBound::updateBoundVel(u_crf, u1, v1, w1);
  mod_calcVel::timeIntVelRK2(u_crf, u, v, w, u1, v1, w1);
  mod_updateBound::updateBoundVel(u_crf, u, v, w);
  mod_calcRHS::calcRHS(u, v, w, p);
  execute_poisson_solver(p);
  mod_updateBound::updateBoundP(p);
  mod_calcVel::correctVel(p, u, v, w);
  mod_calcVel::forceVel(u, v, w);
  mod_updateBound::updateBoundVel(u_crf, u, v, w);
}

// the next step at which the loop body has work besides stepping
int next_event(int nt) {
  auto next = [nt](int every) { return every > 0 ? (nt / every + 1) * every : 1 << 30; };
  int e = nt_end;
  for (int every : {nt_check, nt_out_scrn, nt_out_inst, nt_out_save, nt_out_stat})
    e = std::min(e, next(every));
  // statistics start at nt_init_stat + 1 and sample every sample_interval
  // steps after nt_init_stat (mod_statistics::calcStat); steps in between stay fused
  if (nt <= nt_init_stat) e = std::min(e, nt_init_stat + 1);
  else if (sample_interval > 0)
    e = std::min(e, nt_init_stat + ((nt - nt_init_stat) / sample_interval + 1) * sample_interval);
  return e;
}

int check_only(const char* fn) {
  readInputParam(fn);
  allocVariables({nx, ny, nz});
  mod_mesh::initMesh();
  std::printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"lx\": %.17g, \"ly\": %.17g, \"lz\": %.17g,\n",
              nx, ny, nz, lx, ly, lz);
  std::printf(" \"re\": %.17g, \"re_inv\": %.17g, \"dt\": %.17g, \"nt_end\": %d, \"nt_check\": %d,\n", re,
              re_inv, dt, nt_end, nt_check);
  std::printf(" \"initial_field\": \"%s\", \"is_forced\": [%d, %d, %d], \"vel_force\": [%.17g, %.17g, %.17g],\n",
              initial_field.c_str(), (int)is_forced[0], (int)is_forced[1], (int)is_forced[2],
              vel_force[0], vel_force[1], vel_force[2]);
  std::printf(" \"precision\": %d, \"fused_step\": %d, \"closure\": \"%s\", \"dx\": %.17g, \"dy\": %.17g,\n",
              precision, (int)fused_step, closure_file.c_str(), mod_mesh::dx, mod_mesh::dy);
  const char* names[] = {"dzf_global", "zc_global", "dzc", "dzflzi"};
  const Array1DH1* arrs[] = {&mod_mesh::dzf_global, &mod_mesh::zc_global, &mod_mesh::dzc,
                             &mod_mesh::dzflzi};
  for (int a = 0; a < 4; ++a) {
    std::printf(" \"%s\": [", names[a]);
    for (int k = 0; k <= nz + 1; ++k) std::printf("%s%.17g", k ? ", " : "", (*arrs[a])(k));
    std::printf("]%s\n", a < 3 ? "," : "}");
  }
  return 0;
}

int run(const char* fn) {
  std::printf(
      "================================================================================\n"
      "                      PowerLLEL_channel (DNS, 2ndFD + RK2)                      \n"
      "================================================================================\n"
      "PowerLLEL.NOTE: Initialization starts ...\n");
  readInputParam(fn);
  allocVariables({nx, ny, nz});
  mod_mesh::initMesh();
  mod_mesh::outputMesh();
  std::array<std::array<int, 6>, 3> neighbor_xyz{};
  init_poisson_solver(nx, ny, nz, mod_mesh::dx, mod_mesh::dy, mod_mesh::dzf_global.data(), "PP",
                      "PP", "NN", neighbor_xyz);
  if (!closure_file.empty()) {
    mod_device::attach_closure(closure_file);
    std::printf("PowerLLEL.NOTE: Learned closure attached from %s\n", closure_file.c_str());
  }

  int nt_start;
  if (is_restart) {
    std::printf("PowerLLEL.NOTE: Initializing flow from checkpoint fields ...\n");
    nt_start = mod_dataIO::inputData(u, v, w) + 1;
  } else {
    std::printf("PowerLLEL.NOTE: Initializing flow according to input parameters ...\n");
    mod_initFlow::initFlow(u, v, w);
    nt_start = 1;
  }
  mod_calcVel::transform2CRF(u_crf, u, vel_force[0]);
  mod_updateBound::updateBoundVel(u_crf, u, v, w);
  mod_updateBound::updateBoundP(p);
  if (is_restart && nt_start > nt_init_stat + 1) {
    mod_statistics::allocStat();
    mod_dataIO::inputStatData(nt_start - 1);
  }

  mod_utils::value_index_pair_t cfl_max, div_max;
  mod_utils::initCheckCFLAndDiv(cfl_max, div_max);
  std::printf("PowerLLEL.NOTE: Initialization ends successfully!\n");
  std::printf("PowerLLEL.NOTE: Selected device is: HIP device %d, precision %s, %s steps\n",
              device_id, precision == PLL_FP64 ? "fp64" : "fp32_mixed",
              fused_step ? "fused" : "module-by-module");
  std::printf("PowerLLEL.NOTE: Simulation starts at nt = %9d!\n", nt_start);
  std::printf("********************************************************************************\n");
  std::printf("%9s  %17s  %14s  %10s  %10s\n", "nt", "speed(wSteps/Day)", "remaining time",
              "cfl_max", "div_max");
  std::fflush(stdout);

  using clock = std::chrono::steady_clock;
  auto wtime = clock::now();
  auto t_loop = wtime;
  int nt = nt_start;
  while (nt <= nt_end) {
    // steps nt .. last run without anything else in between
    int last = nt;
    if (fused_step) {
      last = std::max(nt, std::min(next_event(nt - 1), nt_end));
      mod_device::check(pll_step(mod_device::context(), last - nt + 1), "pll_step");
      mod_device::sync();
    } else {
      step_modules();
    }
    nt = last;

    if (nt % nt_check == 0) {
      bool ok = true;
      ok &= mod_utils::checkNaN(u, "u");
      mod_utils::calcMaxCFL(dt, mod_mesh::dx_inv, mod_mesh::dy_inv, mod_mesh::dzf_inv, u, v, w,
                            cfl_max);
      ok &= mod_utils::checkCFL(cfl_limit, cfl_max);
      mod_utils::calcMaxDiv(mod_mesh::dx_inv, mod_mesh::dy_inv, mod_mesh::dzf_inv, u, v, w,
                            div_max);
      ok &= mod_utils::checkDiv(div_limit, div_max);
      if (!ok) {
        std::printf("PowerLLEL.ERROR: nt = %9d, cfl_max = %10.3e at (%5d,%5d,%5d) from rank%5d\n",
                    nt, cfl_max.value, cfl_max.ig, cfl_max.jg, cfl_max.kg, cfl_max.rank);
        std::printf("PowerLLEL.ERROR: nt = %9d, div_max = %10.3e at (%5d,%5d,%5d) from rank%5d\n",
                    nt, div_max.value, div_max.ig, div_max.jg, div_max.kg, div_max.rank);
        return 1;
      }
    }

    if (nt > nt_init_stat) {
      if (nt == nt_init_stat + 1) {
        mod_statistics::allocStat();
        mod_statistics::initStat();
        std::printf("PowerLLEL.NOTE: Statistical process starts at nt = %d!\n", nt);
      }
      mod_statistics::calcStat(nt, u_crf, u, v, w, p);
    }

    if (nt % nt_out_scrn == 0) {
      const double wt = std::chrono::duration<double>(clock::now() - wtime).count();
      int t = static_cast<int>((nt_end - nt) * wt / nt_out_scrn);
      const int t_d = t / 86400;
      int t_res = t % 86400;
      const int t_h = t_res / 3600;
      t_res %= 3600;
      const int t_m = t_res / 60, t_s = t_res % 60;
      char remaining[32];
      std::snprintf(remaining, sizeof(remaining), "%3dd%2dh%2dm%2ds", t_d, t_h, t_m, t_s);
      const double speed = nt_out_scrn * 3600.0 * 24.0 / wt / 10000;
      std::printf("%9d  %17.3f  %14s  %10.3E  %10.3E\n", nt, speed, remaining, cfl_max.value,
                  div_max.value);
      std::fflush(stdout);
    }

    mod_dataIO::outputData(nt, u_crf, u, v, w, p);
    if (nt % nt_out_scrn == 0) wtime = clock::now();
    ++nt;
  }
  const double loop_s = std::chrono::duration<double>(clock::now() - t_loop).count();
  const int nsteps = nt_end - nt_start + 1;
  if (nsteps > 0)
    std::printf("PowerLLEL.NOTE: %d steps in %.3f s: %.3f ms/step, %.1f Mpoint-steps/s\n", nsteps,
                loop_s, 1e3 * loop_s / nsteps, 1e-6 * (double)nx * ny * nz * nsteps / loop_s);

  free_poisson_solver();
  mod_mesh::freeMesh();
  freeVariables();
  mod_statistics::freeStat();
  mod_utils::freeCheckCFLAndDiv();
  std::printf("********************************************************************************\n");
  std::printf("PowerLLEL.NOTE: Simulation ends successfully!\n");
  return 0;
}

int mesh_only(const char* fn) {
  readInputParam(fn);
  mod_mesh::initMesh();
  mod_mesh::outputMesh();
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  const bool check = mode == "--check", mesh = mode == "--mesh";
  const int a = (check || mesh) ? 2 : 1;
  const char* fn = argc > a ? argv[a] : "param.json";
  try {
    return check ? check_only(fn) : mesh ? mesh_only(fn) : run(fn);
  } catch (const std::exception& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "%s\n", e.what());
    mod_device::destroy();
    return 2;
  }
}
