// Error reporting and version of the cov3d HIP library.
#include <stdarg.h>

#include "common.h"

namespace cov3d {

static thread_local char g_last_error[1024] = "no error";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace cov3d

extern "C" const char* cov3d_last_error(void) { return cov3d::g_last_error; }

extern "C" const char* cov3d_version(void) { return "cov3d_hip 0.1 gfx950"; }
