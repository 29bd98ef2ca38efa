#include "mod_device.hh"

#include "mod_type.hh"

namespace mod_device {
namespace {
pll_context* g_ctx = nullptr;
pll_closure* g_closure = nullptr;
}

void check(int rc, const char* what) {
  if (rc == PLL_OK) return;
  const char* msg = pll_last_error();
  throw Error(rc, std::string(what) + ": " + (msg ? msg : ""));
}

bool has_context() { return g_ctx != nullptr; }

pll_context* context() {
  if (!g_ctx)
    throw Error(PLL_ERR_STATE,
                "no device context: call init_poisson_solver after readInputParam and initMesh");
  return g_ctx;
}

void create(const pll_config& cfg, const double* dzf_global) {
  if (g_ctx) throw Error(PLL_ERR_STATE, "the device context already exists");
  check(pll_create(&cfg, dzf_global, &g_ctx), "pll_create");
}

void attach_closure(const std::string& path) {
  if (g_closure) throw Error(PLL_ERR_STATE, "a closure is already attached");
  check(pll_closure_load(path.c_str(), &g_closure), "pll_closure_load");
  check(pll_closure_attach(context(), g_closure), "pll_closure_attach");
}

void destroy() {
  if (g_ctx && g_closure) pll_closure_attach(g_ctx, nullptr);
  if (g_closure) pll_closure_destroy(g_closure);
  g_closure = nullptr;
  if (g_ctx) pll_destroy(g_ctx);
  g_ctx = nullptr;
}

void sync() { check(pll_synchronize(context()), "pll_synchronize"); }

}  // namespace mod_device

void Array3D::allocate(int slot, std::array<int, 3> sz) {
  slot_ = slot;
  sz_ = sz;
  host_.assign(static_cast<size_t>(sz[0] + 2) * (sz[1] + 2) * (sz[2] + 2), 0.0);
}

void Array3D::deallocate() {
  slot_ = -1;
  host_.clear();
  host_.shrink_to_fit();
}

void Array3D::to_host() {
  mod_device::check(pll_field_download(mod_device::context(), slot_, host_.data()),
                    "pll_field_download");
}

void Array3D::to_device() {
  mod_device::check(pll_field_upload(mod_device::context(), slot_, host_.data()),
                    "pll_field_upload");
}
