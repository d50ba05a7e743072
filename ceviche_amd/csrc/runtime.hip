// Library-wide host utilities: last-error string and version.
#include "common.h"

#include <stdarg.h>
#include <stdio.h>

namespace cv {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace cv

extern "C" const char* cv_last_error(void) { return cv::g_err; }

extern "C" const char* cv_version(void) { return "ceviche_amd 0.1.0 (gfx950)"; }
