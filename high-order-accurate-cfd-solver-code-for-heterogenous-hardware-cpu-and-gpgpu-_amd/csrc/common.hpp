// Shared host-side helpers: error reporting into a thread-local message.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/cfdx.h"

namespace cfdx {

std::string &last_error_slot();

// Records a message for cfdx_last_error() and returns the code.
inline int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error_slot() = buf;
    return code;
}

}  // namespace cfdx
