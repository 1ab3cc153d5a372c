// Error plumbing shared by every entry point of libtgs_hip.so.
#include <stdarg.h>
#include <stdio.h>

#include "tgs_hip.h"

namespace tgs {
static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace tgs

extern "C" const char *tgs_last_error(void) { return tgs::g_err; }
extern "C" int tgs_abi_version(void) { return 2; }
