// mod_statistics.hh -- time statistics of the run (src/mod_statistics.hh).
//
// The per-point sums live on the device (pll_stat_*).  Unlike
// src/main.cc:180-190, which calls calcStat only at nt == nt_init_stat + 1,
// the driver samples every step after nt_init_stat, every sample_interval.
#pragma once

#include <array>
#include <vector>

#include "mod_type.hh"

namespace mod_statistics {
struct StatInfo {
  int nts;
  int nte;
  int nspl;
};
inline StatInfo stat_info{0, 0, 0};
inline bool stat_allocated = false;

const char* stat_type(int i);  // u v w u2 v2 w2 uv uw vw p p2

void allocStat(std::array<int, 3> sz = {0, 0, 0});
void freeStat();
void initStat();
void calcStat(int nt, fp u_crf, const Array3DH3& u, const Array3DH3& v, const Array3DH3& w,
              const Array3DH1& p);
// Plane average of quantity var over the samples so far: nz values, k = 1..nz.
std::vector<double> meanProfile(int var);
}  // namespace mod_statistics
