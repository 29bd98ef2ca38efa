#include "mod_parameters.hh"

#include <cstdio>
#include <stdexcept>

#include "json_lite.hh"

namespace {

using json_lite::Value;

[[noreturn]] void bad(const char* key, const char* want) {
  throw std::runtime_error(std::string("PowerLLEL.ERROR.readInputParam: <") + key +
                           "> must be " + want);
}

void get(const Value& v, const char* key, int& out) {
  if (v.kind != Value::Number || !v.integral) bad(key, "an integer");
  out = static_cast<int>(v.num);
}
void get(const Value& v, const char* key, double& out) {
  if (v.kind != Value::Number) bad(key, "a number");
  out = v.num;
}
void get(const Value& v, const char* key, bool& out) {
  if (v.kind != Value::Bool) bad(key, "true or false");
  out = v.b;
}
void get(const Value& v, const char* key, std::string& out) {
  if (v.kind != Value::String) bad(key, "a string");
  out = v.str;
}

// A missing key keeps its default, as in src/mod_parameters.cc:27-44.
template <typename T>
void read(const Value& node, const char* key, T& val) {
  if (const Value* v = node.find(key)) get(*v, key, val);
}
template <typename T, size_t N>
void read(const Value& node, const char* key, std::array<T, N>& val) {
  const Value* v = node.find(key);
  if (!v) return;
  if (v->kind != Value::Array || v->arr.size() < N) bad(key, "an array of the right length");
  for (size_t i = 0; i < N; ++i) get(v->arr[i], key, val[i]);
}

#define READ(node, val) read(node, #val, val)

const Value& section(const Value& d, const char* name) {
  const Value* s = d.find(name);
  if (!s || s->kind != Value::Object)
    throw std::runtime_error(std::string("PowerLLEL.ERROR.readInputParam: Problem with ") + name +
                             " in the input file!");
  return *s;
}

}  // namespace

bool checkMeshAndPara() {
  bool ok = true;
  auto check = [&](bool passed, const char* msg) {
    if (!passed) std::printf("PowerLLEL.ERROR.checkInputParam: %s\n", msg);
    ok &= passed;
  };
  check(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "<nx>, <ny>, <nz> should be even integers!");
  check(p_row > 0 && p_col > 0 && nx % p_row == 0, "<nx> should be exactly divisible by <p_row>!");
  check(p_row > 0 && ny % p_row == 0, "<ny> should be exactly divisible by <p_row>!");
  check(p_col > 0 && nz % p_col == 0, "<nz> should be exactly divisible by <p_col>!");
  check(p_col > 0 && nz % (p_col * nthreads) == 0,
        "<nz> should be exactly divisible by <p_col*nthreads>!");
  check(p_row == 1 && p_col == 1,
        "one process drives one GPU holding the whole domain: <p_row> and <p_col> must be 1!");
  check(dt > 0.0 && nt_check > 0 && nt_out_scrn > 0, "<dt>, <nt_check> and <nt_out_scrn> must be positive!");
  return ok;
}

void readInputParam(std::string_view fn_input) {
  Value d;
  try {
    d = json_lite::parse_file(std::string(fn_input));
  } catch (const json_lite::ParseError& e) {
    throw std::runtime_error(std::string("PowerLLEL.ERROR.readInputParam: Fail to open file ") +
                             std::string(fn_input) + "! (" + e.what() + ")");
  }
  {
    const Value& n = section(d, "PARA");
    READ(n, p_row);
    READ(n, p_col);
  }
  {
    const Value& n = section(d, "MESH");
    READ(n, read_mesh);
    READ(n, mesh_type);
    READ(n, stretch_func);
    READ(n, stretch_ratio);
    READ(n, nx);
    READ(n, ny);
    READ(n, nz);
    READ(n, lx);
    READ(n, ly);
    READ(n, lz);
  }
  {
    const Value& n = section(d, "TIME");
    READ(n, nt_end);
    READ(n, dt);
    READ(n, nt_check);
    READ(n, cfl_limit);
    READ(n, div_limit);
  }
  {
    const Value& n = section(d, "RESTART");
    READ(n, is_restart);
    READ(n, fn_prefix_input_inst);
    READ(n, fn_prefix_input_stat);
  }
  {
    const Value& n = section(d, "PHYSICAL");
    READ(n, re);
    READ(n, u_ref);
    READ(n, l_ref);
    READ(n, initial_field);
    READ(n, u0);
    READ(n, init_with_noise);
    READ(n, noise_intensity);
    READ(n, smooth_wall_visc);
    READ(n, u_crf);
  }
  {
    const Value& n = section(d, "FORCE");
    READ(n, is_forced);
    READ(n, vel_force);
  }
  {
    const Value& n = section(d, "STATISTICS");
    READ(n, nt_init_stat);
    READ(n, sample_interval);
    READ(n, stat_which_var);
  }
  {
    const Value& n = section(d, "OUTPUT");
    READ(n, nt_out_scrn);
    READ(n, nt_out_inst);
    READ(n, nt_out_stat);
    READ(n, nt_out_save);
    READ(n, overwrite_save);
    READ(n, auto_cleanup);
    READ(n, num_retained);
    READ(n, nt_out_moni);
  }
  {
    const Value& n = section(d, "MONITOR");
    READ(n, out_forcing);
    READ(n, out_probe_point);
    READ(n, probe_ijk);
    READ(n, out_skf_z);
    READ(n, out_region);
    READ(n, nt_out_region);
    READ(n, region_ijk);
  }
  if (const Value* n = d.find("DEVICE")) {
    std::string prec;
    read(*n, "precision", prec);
    if (prec == "fp64" || prec.empty())
      precision = PLL_FP64;
    else if (prec == "fp32_mixed")
      precision = PLL_FP32_MIXED;
    else
      bad("precision", "\"fp64\" or \"fp32_mixed\"");
    read(*n, "device", device_id);
    READ(*n, fused_step);
    read(*n, "closure", closure_file);
  }
  if (!checkMeshAndPara())
    throw std::runtime_error("PowerLLEL.ERROR.checkInputParam: invalid parameters");
  re_inv = u_ref * l_ref / re;
}
