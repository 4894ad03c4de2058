#include "errors.h"
#include "../../include/lya2pcf.h"

static thread_local char g_err[1024] = "";

int lya_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* lya_last_error(void) { return g_err; }
extern "C" int lya_version(void) { return 1; }
