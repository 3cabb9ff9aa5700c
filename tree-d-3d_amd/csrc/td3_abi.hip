// Library-wide entry points: ABI version and the per-thread error string.

#include <stdarg.h>
#include <stdio.h>

#include "td3_common.h"

namespace {
thread_local char g_error[1024] = "";
}

void td3_set_error(const char* fmt, ...) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
}

extern "C" int td3_abi_version(void) { return TD3_ABI_VERSION; }

extern "C" const char* td3_last_error(void) { return g_error; }
