// mod_device.hh -- the one device context behind the host modules (the role
// of the reference's global SYCL queue, src/mod_oneapi.hh).
//
// Every module routine runs on this context.  Library failures become
// mod_device::Error (a std::runtime_error carrying the library's message);
// the reference aborts the MPI job instead (src/mod_type.hh:30-146).
#pragma once

#include <stdexcept>
#include <string>

#include "pll.h"

namespace mod_device {

struct Error : std::runtime_error {
  int status;
  Error(int rc, const std::string& what) : std::runtime_error(what), status(rc) {}
};

// Throws Error(rc, "<what>: <library message>") when rc != PLL_OK.
void check(int rc, const char* what);

bool has_context();
pll_context* context();  // throws when none has been created
void create(const pll_config& cfg, const double* dzf_global);
// Load a saved learned closure (pll_closure_load) and attach it to the context.
void attach_closure(const std::string& path);
void destroy();  // detaches and frees the closure first

// Synchronous semantics of the reference's module calls: wait for the stream.
void sync();

}  // namespace mod_device
