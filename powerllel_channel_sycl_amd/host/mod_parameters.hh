// mod_parameters.hh -- run parameters read from param.json
// (src/mod_parameters.hh:1-90, src/mod_parameters.cc:160-260).
//
// Same sections, keys and meaning as the reference's file.  One optional
// section is added, "DEVICE": {"precision": "fp64" | "fp32_mixed",
// "device": <HIP ordinal>, "fused_step": true | false, "closure": "<file>"};
// "closure" names a learned closure saved by pll_closure_save, attached to the
// run's context (an LES on this grid).
#pragma once

#include <array>
#include <cmath>
#include <string>
#include <string_view>

#include "mod_type.hh"

void readInputParam(std::string_view fn_input);

constexpr fp pi = M_PI;

// parallel computing parameters
inline int p_row = 1, p_col = 1, nthreads = 1;

// mesh parameters
inline bool read_mesh = false;
inline int mesh_type = 0;
inline fp stretch_ratio = 0.0;
inline int stretch_func = 0;
inline int nx = 0, ny = 0, nz = 0;
inline fp lx = 0.0, ly = 0.0, lz = 0.0;

// time step parameters
inline int nt_end = 0;
inline int nt_check = 1;
inline fp dt = 0.0;
inline fp cfl_limit = 1.0;
inline fp div_limit = 1.0;

// restart computing parameters
inline bool is_restart = false;
inline std::string fn_prefix_input_inst = "save";
inline std::string fn_prefix_input_stat = "save_stat";

// physical property parameters
inline fp re = 1.0, u_ref = 1.0, l_ref = 1.0, re_inv = 1.0;
inline std::string initial_field = "uni";
inline fp u0 = 0.0;
inline bool init_with_noise = false;
inline fp noise_intensity = 0.0;
inline bool smooth_wall_visc = false;
inline fp u_crf = 0.0;

// forced flow parameters
inline std::array<bool, 3> is_forced{false, false, false};
inline std::array<fp, 3> vel_force{0.0, 0.0, 0.0};

// statistics parameters
inline int nt_init_stat = 1 << 30;
inline int sample_interval = 1;
inline std::array<bool, 11> stat_which_var{};

// data output parameters
inline int nt_out_scrn = 1;
inline int nt_out_inst = 1 << 30;
inline int nt_out_stat = 1 << 30;
inline int nt_out_save = 1 << 30;
inline bool overwrite_save = true;
inline bool auto_cleanup = true;
inline int num_retained = 0;
inline int nt_out_moni = 1 << 30;
inline std::string fn_prefix_inst = "inst";
inline std::string fn_prefix_save = "save";
inline std::string fn_prefix_stat = "stat";

// monitor parameters (read for compatibility; the probes are not built)
inline bool out_forcing = false;
inline bool out_probe_point = false;
inline std::array<int, 3> probe_ijk{1, 1, 1};
inline std::array<bool, 2> out_skf_z{false, false};
inline bool out_region = false;
inline int nt_out_region = 1 << 30;
inline std::array<int, 6> region_ijk{1, 1, 1, 1, 1, 1};

// device parameters (this library's own section)
inline int precision = PLL_FP64;
inline int device_id = 0;
inline bool fused_step = true;
inline std::string closure_file;  // empty: no closure

// The sanity checks of src/mod_parameters.cc:170-224 plus the one-GPU rule;
// returns false after printing every failed check.
bool checkMeshAndPara();
