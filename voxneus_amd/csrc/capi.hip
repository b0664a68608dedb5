// Library-wide C-ABI plumbing: version string and the thread-local error message.
#include <stdarg.h>

#include "vxn_host.h"

namespace vxn {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace vxn

extern "C" const char* vxn_version(void) { return "voxneus_hip 0.1.0 gfx950"; }
extern "C" const char* vxn_last_error(void) { return vxn::g_err; }
