#include "mod_statistics.hh"

#include <stdexcept>

#include "mod_device.hh"
#include "mod_parameters.hh"

namespace mod_statistics {

const char* stat_type(int i) {
  static const char* names[11] = {"u", "v", "w", "u2", "v2", "w2", "uv", "uw", "vw", "p", "p2"};
  return (i >= 0 && i < 11) ? names[i] : "?";
}

void allocStat(std::array<int, 3>) {
  int which[11];
  for (int i = 0; i < 11; ++i) which[i] = stat_which_var[i] ? 1 : 0;
  mod_device::check(pll_stat_enable(mod_device::context(), which), "allocStat");
  stat_allocated = true;
}

void freeStat() {
  if (!stat_allocated || !mod_device::has_context()) {
    stat_allocated = false;
    return;
  }
  int none[11] = {};
  mod_device::check(pll_stat_enable(mod_device::context(), none), "freeStat");
  stat_allocated = false;
}

void initStat() {
  stat_info.nts = nt_init_stat + 1;
  stat_info.nte = nt_init_stat;
  stat_info.nspl = 0;
}

void calcStat(int nt, fp, const Array3DH3&, const Array3DH3&, const Array3DH3&, const Array3DH1&) {
  if (sample_interval <= 0 || (nt - nt_init_stat) % sample_interval != 0) return;
  stat_info.nte += sample_interval;
  stat_info.nspl += 1;
  mod_device::check(pll_stat_accumulate(mod_device::context()), "calcStat");
}

std::vector<double> meanProfile(int var) {
  std::vector<double> out(nz, 0.0);
  if (!stat_which_var.at(var) || stat_info.nspl == 0) return out;
  mod_device::check(pll_stat_plane_sums(mod_device::context(), var, out.data()), "meanProfile");
  const double scale = 1.0 / ((double)nx * ny * stat_info.nspl);
  for (double& x : out) x *= scale;
  return out;
}

}  // namespace mod_statistics
