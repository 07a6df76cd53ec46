// Thread-local error message for the C ABI.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "imba_hip.h"

namespace imba {

void set_error(const char* fmt, ...);
void clear_error();

// Returns `code` after recording the message (use as `return fail(IMBA_ERR_..., "...")`).
int fail(int code, const char* fmt, ...);

}  // namespace imba

#define IMBA_CHECK_ARG(cond, msg)                                    \
    do {                                                             \
        if (!(cond)) return ::imba::fail(IMBA_ERR_INVALID_ARG, "%s", msg); \
    } while (0)
