// Error reporting and version query of the C ABI.
#include <cstdarg>
#include <cstdio>

#include "pg_common.h"

namespace pg {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace pg

extern "C" {

int pg_abi_version(void) { return PG_ABI_VERSION; }

const char* pg_last_error(void) { return pg::g_err; }

}  // extern "C"
