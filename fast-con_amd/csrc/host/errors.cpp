// Per-thread error text for fc_last_error(). The library never exits the
// process; the reference's error()/cudaError() paths (BFS-4K/Util/fUtil.h:18,
// BFS-4K/Util/cudaUtil.cuh:90-96) become a status plus this message.
#include "common.hpp"

#include <cstdlib>

namespace fc {

static thread_local std::string g_last_error;

fc_status fail(fc_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

void clear_error() { g_last_error.clear(); }

}  // namespace fc

extern "C" const char* fc_last_error(void) { return fc::g_last_error.c_str(); }

extern "C" const char* fc_version(void) { return "fastcon 0.1 gfx950"; }

extern "C" void fc_free(void* p) { std::free(p); }
