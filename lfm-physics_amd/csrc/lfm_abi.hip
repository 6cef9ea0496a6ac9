// Library-wide entry points: ABI version and the per-thread error message.
#include <string>

#include "lfm_internal.h"

namespace lfm {
namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

}  // namespace lfm

extern "C" int lfm_abi_version(void) { return LFM_ABI_VERSION; }

extern "C" const char* lfm_last_error(void) { return lfm::g_last_error.c_str(); }
