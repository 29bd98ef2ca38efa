#include "mod_dataIO.hh"

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mod_device.hh"
#include "mod_mesh.hh"
#include "mod_parameters.hh"
#include "mod_statistics.hh"

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>


namespace mod_dataIO {
namespace {

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  std::vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

[[noreturn]] void io_error(const std::string& what) {
  throw std::runtime_error("PowerLLEL.ERROR.dataIO: " + what);
}

// The serial HDF5 C library is loaded at run time (PLL_HDF5_LIB, else the
// conda copy in the image), so that linking it does not put its directory --
// which carries an older libstdc++ than the HIP runtime needs -- on the
// library search path.  Types and constants are those of HDF5 1.10.
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0;

struct H5Api {
  herr_t (*open)();
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char*, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
  hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen2)(hid_t, const char*, hid_t);
  herr_t (*dclose)(hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  hid_t (*dget_space)(hid_t);
  hid_t* native_int;
  hid_t* native_double;
};

const H5Api& h5() {
  static H5Api api;
  static bool loaded = false;
  if (loaded) return api;
  const char* env = std::getenv("PLL_HDF5_LIB");
  const char* paths[] = {env, "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                         "libhdf5.so", "libhdf5_serial.so"};
  void* h = nullptr;
  for (const char* p : paths)
    if (p && (h = dlopen(p, RTLD_NOW | RTLD_LOCAL))) break;
  if (!h) io_error("the HDF5 library is not available (set PLL_HDF5_LIB)");
  auto sym = [h](const char* name) {
    void* f = dlsym(h, name);
    if (!f) io_error(std::string("HDF5 symbol missing: ") + name);
    return f;
  };
  api.open = reinterpret_cast<herr_t (*)()>(sym("H5open"));
  api.fcreate = reinterpret_cast<hid_t (*)(const char*, unsigned, hid_t, hid_t)>(sym("H5Fcreate"));
  api.fopen = reinterpret_cast<hid_t (*)(const char*, unsigned, hid_t)>(sym("H5Fopen"));
  api.fclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Fclose"));
  api.screate_simple =
      reinterpret_cast<hid_t (*)(int, const hsize_t*, const hsize_t*)>(sym("H5Screate_simple"));
  api.sclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Sclose"));
  api.sget_ndims = reinterpret_cast<int (*)(hid_t)>(sym("H5Sget_simple_extent_ndims"));
  api.sget_dims = reinterpret_cast<int (*)(hid_t, hsize_t*, hsize_t*)>(sym("H5Sget_simple_extent_dims"));
  api.dcreate2 = reinterpret_cast<hid_t (*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t)>(
      sym("H5Dcreate2"));
  api.dopen2 = reinterpret_cast<hid_t (*)(hid_t, const char*, hid_t)>(sym("H5Dopen2"));
  api.dclose = reinterpret_cast<herr_t (*)(hid_t)>(sym("H5Dclose"));
  api.dwrite = reinterpret_cast<herr_t (*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*)>(sym("H5Dwrite"));
  api.dread = reinterpret_cast<herr_t (*)(hid_t, hid_t, hid_t, hid_t, hid_t, void*)>(sym("H5Dread"));
  api.dget_space = reinterpret_cast<hid_t (*)(hid_t)>(sym("H5Dget_space"));
  if (api.open() < 0) io_error("H5open failed");
  api.native_int = static_cast<hid_t*>(sym("H5T_NATIVE_INT_g"));
  api.native_double = static_cast<hid_t*>(sym("H5T_NATIVE_DOUBLE_g"));
  loaded = true;
  return api;
}

struct H5File {
  hid_t id;
  H5File(const std::string& name, bool create) {
    id = create ? h5().fcreate(name.c_str(), kAccTrunc, kDefault, kDefault)
                : h5().fopen(name.c_str(), kAccRdonly, kDefault);
    if (id < 0) io_error("cannot " + std::string(create ? "create " : "open ") + name);
  }
  ~H5File() { h5().fclose(id); }
};

void write_int(hid_t f, const char* tag, int value) {
  const H5Api& H = h5();
  hsize_t one = 1;
  hid_t sp = H.screate_simple(1, &one, nullptr);
  hid_t ds = H.dcreate2(f, tag, *H.native_int, sp, kDefault, kDefault, kDefault);
  herr_t e = ds < 0 ? -1 : H.dwrite(ds, *H.native_int, kAll, kAll, kDefault, &value);
  if (ds >= 0) H.dclose(ds);
  H.sclose(sp);
  if (e < 0) io_error(std::string("cannot write dataset ") + tag);
}

int read_int(hid_t f, const char* tag) {
  const H5Api& H = h5();
  int value = 0;
  hid_t ds = H.dopen2(f, tag, kDefault);
  if (ds < 0) io_error(std::string("missing dataset ") + tag);
  herr_t e = H.dread(ds, *H.native_int, kAll, kAll, kDefault, &value);
  H.dclose(ds);
  if (e < 0) io_error(std::string("cannot read dataset ") + tag);
  return value;
}

void write_3d(hid_t f, const char* tag, const double* data) {
  const H5Api& H = h5();
  hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
  hid_t sp = H.screate_simple(3, dims, nullptr);
  hid_t ds = H.dcreate2(f, tag, *H.native_double, sp, kDefault, kDefault, kDefault);
  herr_t e = ds < 0 ? -1 : H.dwrite(ds, *H.native_double, kAll, kAll, kDefault, data);
  if (ds >= 0) H.dclose(ds);
  H.sclose(sp);
  if (e < 0) io_error(std::string("cannot write dataset ") + tag);
}

void write_1d(hid_t f, const char* tag, const double* data, long n) {
  const H5Api& H = h5();
  hsize_t dims = (hsize_t)n;
  hid_t sp = H.screate_simple(1, &dims, nullptr);
  hid_t ds = H.dcreate2(f, tag, *H.native_double, sp, kDefault, kDefault, kDefault);
  herr_t e = ds < 0 ? -1 : H.dwrite(ds, *H.native_double, kAll, kAll, kDefault, data);
  if (ds >= 0) H.dclose(ds);
  H.sclose(sp);
  if (e < 0) io_error(std::string("cannot write dataset ") + tag);
}

void read_3d(hid_t f, const char* tag, double* data) {
  const H5Api& H = h5();
  hid_t ds = H.dopen2(f, tag, kDefault);
  if (ds < 0) io_error(std::string("missing dataset ") + tag);
  hid_t sp = H.dget_space(ds);
  hsize_t dims[3] = {0, 0, 0};
  const int rank = H.sget_ndims(sp);
  if (rank == 3) H.sget_dims(sp, dims, nullptr);
  H.sclose(sp);
  if (rank != 3 || dims[0] != (hsize_t)nz || dims[1] != (hsize_t)ny || dims[2] != (hsize_t)nx) {
    H.dclose(ds);
    io_error(std::string("dataset ") + tag + " is not (nz, ny, nx)");
  }
  herr_t e = H.dread(ds, *H.native_double, kAll, kAll, kDefault, data);
  H.dclose(ds);
  if (e < 0) io_error(std::string("cannot read dataset ") + tag);
}

std::vector<double> interior(const Array3D& var, fp shift) {
  std::vector<double> out((size_t)nx * ny * nz);
  size_t o = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) out[o++] = var(i, j, k) + shift;
  return out;
}

void stat_file(const std::string& name, int var, bool input) {
  using namespace mod_statistics;
  const std::string tag = std::string(stat_type(var)) + "_stat";
  std::vector<double> buf((size_t)nx * ny * nz);
  if (input) {
    H5File f(name, false);
    stat_info.nts = read_int(f.id, "nts");
    stat_info.nte = read_int(f.id, "nte");
    stat_info.nspl = read_int(f.id, "nspl");
    read_3d(f.id, tag.c_str(), buf.data());
    mod_device::check(pll_stat_upload(mod_device::context(), var, buf.data()), "inputStatData");
  } else {
    mod_device::check(pll_stat_download(mod_device::context(), var, buf.data()), "outputStatData");
    H5File f(name, true);
    write_int(f.id, "nts", stat_info.nts);
    write_int(f.id, "nte", stat_info.nte);
    write_int(f.id, "nspl", stat_info.nspl);
    write_3d(f.id, tag.c_str(), buf.data());
  }
}

// Plane-averaged profiles: z, then the 11 means and the Reynolds stresses.
void write_profiles(const std::string& name) {
  using namespace mod_statistics;
  std::vector<std::vector<double>> m(11);
  for (int t = 0; t < 11; ++t) m[t] = meanProfile(t);
  FILE* f = std::fopen(name.c_str(), "w");
  if (!f) io_error("cannot write " + name);
  std::fprintf(f, "# nts %d nte %d nspl %d; plane averages of cell-centred samples\n",
               stat_info.nts, stat_info.nte, stat_info.nspl);
  std::fprintf(f, "# k zc u v w u2 v2 w2 uv uw vw p p2 u'u' v'v' w'w' u'v' u'w' v'w'\n");
  for (int k = 0; k < nz; ++k) {
    std::fprintf(f, "%5d %23.15E", k + 1, mod_mesh::zc_global.size() ? mod_mesh::zc_global(k + 1) : 0.0);
    for (int t = 0; t < 11; ++t) std::fprintf(f, " %23.15E", m[t][k]);
    const double u = m[0][k], v = m[1][k], w = m[2][k];
    std::fprintf(f, " %23.15E %23.15E %23.15E %23.15E %23.15E %23.15E\n", m[3][k] - u * u,
                 m[4][k] - v * v, m[5][k] - w * w, m[6][k] - u * v, m[7][k] - u * w, m[8][k] - v * w);
  }
  std::fclose(f);
}

}  // namespace

void writeMesh(const char* filename) {
  using mod_mesh::xc_global;
  using mod_mesh::yc_global;
  using mod_mesh::zc_global;
  H5File f(filename, true);
  // interior values, from index 1 (src/mod_hdf5.cc:124-137)
  write_1d(f.id, "xc", &xc_global(1), nx);
  write_1d(f.id, "yc", &yc_global(1), ny);
  write_1d(f.id, "zc", &zc_global(1), nz);
}

void writeField(const char* filename, const char* tag, int nt, const Array3D& var, fp shift) {
  std::vector<double> buf = interior(var, shift);
  H5File f(filename, true);
  write_int(f.id, "nt", nt);
  write_3d(f.id, tag, buf.data());
}

int readField(const char* filename, const char* tag, Array3D& var) {
  std::vector<double> buf((size_t)nx * ny * nz);
  int nt;
  {
    H5File f(filename, false);
    nt = read_int(f.id, "nt");
    read_3d(f.id, tag, buf.data());
  }
  size_t o = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) var(i, j, k) = buf[o++];
  return nt;
}

int inputData(Array3DH3& u, Array3DH3& v, Array3DH3& w) {
  std::printf("PowerLLEL.NOTE.inputData: Reading checkpoint fields ...\n");
  Array3D* f[3] = {&u, &v, &w};
  const char* tags[3] = {"u", "v", "w"};
  int nt0 = 0;
  for (int d = 0; d < 3; ++d) {
    const std::string name = fn_prefix_input_inst + "_" + tags[d] + ".h5";
    const int nt = readField(name.c_str(), tags[d], *f[d]);
    f[d]->to_device();
    std::printf("PowerLLEL.NOTE.inputData: Finish reading checkpoint field <%s>!\n", tags[d]);
    if (d == 0)
      nt0 = nt;
    else if (nt != nt0)
      io_error(fmt("inputData: The timestamp of <%s> does not match with that of <u>!", tags[d]));
  }
  return nt0;
}

void outputData(int nt, fp crf, Array3DH3& u, const Array3DH3& v, const Array3DH3& w,
                const Array3DH1& p) {
  const bool inst = nt % nt_out_inst == 0, save = nt % nt_out_save == 0;
  if (inst || save) {
    // fill the host mirrors once for both outputs
    const Array3D* fields[4] = {&u, &v, &w, &p};
    for (const Array3D* f : fields) const_cast<Array3D*>(f)->to_host();
  }
  if (inst) {
    std::printf("PowerLLEL.NOTE.outputData: Writing instantaneous fields ...\n");
    const std::string pre = fmt("%s_%08d_", fn_prefix_inst.c_str(), nt);
    writeField((pre + "u.h5").c_str(), "u", nt, u, crf);
    writeField((pre + "v.h5").c_str(), "v", nt, v);
    writeField((pre + "w.h5").c_str(), "w", nt, w);
    writeField((pre + "p.h5").c_str(), "p", nt, p);
  }
  if (save) {
    std::printf("PowerLLEL.NOTE.outputData: Writing checkpoint fields ...\n");
    const std::string dump = overwrite_save ? "_" : fmt("_%08d_", nt);
    const std::string pre = fn_prefix_save + dump;
    writeField((pre + "u.h5").c_str(), "u", nt, u, crf);
    writeField((pre + "v.h5").c_str(), "v", nt, v);
    writeField((pre + "w.h5").c_str(), "w", nt, w);
    std::printf("PowerLLEL.NOTE.outputData: Finish writing checkpoint fields!\n");
  }
  if (nt > nt_init_stat && mod_statistics::stat_allocated) outputStatData(nt);
  if (!overwrite_save && auto_cleanup && save) {
    const int nt_cleanup = nt - (num_retained + 1) * nt_out_save;
    if (nt_cleanup > 0) {
      const std::string pre = fmt("%s_%08d_", fn_prefix_save.c_str(), nt_cleanup);
      for (const char* t : {"u", "v", "w"}) std::remove((pre + t + ".h5").c_str());
      if (nt_cleanup > nt_init_stat)
        for (int i = 0; i < 11; ++i)
          if (stat_which_var[i])
            std::remove((pre + "stat_" + mod_statistics::stat_type(i) + ".h5").c_str());
    }
  }
}

void inputStatData(int nt_in_inst) {
  std::printf("PowerLLEL.NOTE.inputStatData: Reading checkpoint fields of statistics ...\n");
  for (int i = 0; i < 11; ++i) {
    if (!stat_which_var[i]) continue;
    stat_file(fn_prefix_input_stat + "_" + mod_statistics::stat_type(i) + ".h5", i, true);
    if (mod_statistics::stat_info.nte != nt_in_inst)
      io_error(fmt("inputStatData: The timestamp of <%s_stat> does not match with that of inst. "
                   "field!",
                   mod_statistics::stat_type(i)));
  }
}

void outputStatData(int nt) {
  using namespace mod_statistics;
  if (nt % nt_out_save == 0) {
    // same file-name rule as the field checkpoints, so that
    // fn_prefix_input_stat = "<fn_prefix_save>_stat" restarts from them
    const std::string dump = overwrite_save ? "_" : fmt("_%08d_", nt);
    for (int i = 0; i < 11; ++i)
      if (stat_which_var[i])
        stat_file(fn_prefix_save + dump + "stat_" + stat_type(i) + ".h5", i, false);
  }
  if ((nt - nt_init_stat) % nt_out_stat == 0) {
    std::printf("PowerLLEL.NOTE.outputStatData: Writing statistics fields ...\n");
    const std::string pre = fmt("%s_%08d-%08d_", fn_prefix_stat.c_str(), stat_info.nts, stat_info.nte);
    for (int i = 0; i < 11; ++i)
      if (stat_which_var[i]) stat_file(pre + stat_type(i) + ".h5", i, false);
    write_profiles(pre + "profiles.txt");
  }
}

}  // namespace mod_dataIO
