#include "status.h"

namespace imba {

static thread_local std::string t_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
}

void clear_error() { t_last_error.clear(); }

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
    return code;
}

}  // namespace imba

extern "C" const char* imba_last_error(void) { return imba::t_last_error.c_str(); }
extern "C" const char* imba_version(void) { return "imbatracer_amd 0.1 (gfx950)"; }
