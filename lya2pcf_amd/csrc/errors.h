// Error reporting shared by the translation units of the lya2pcf library:
// the last failure message is kept per thread and returned by lya_last_error().
#pragma once
#include <cstdarg>
#include <cstdio>

int lya_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
