// Error reporting and version of libscarfhip.
#include <stdarg.h>
#include <stdio.h>

#include "sh_common.h"

namespace sh {
static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace sh

extern "C" int sh_version(void) { return 10000; }  // 1.0.0
extern "C" const char *sh_last_error(void) { return sh::g_err; }
