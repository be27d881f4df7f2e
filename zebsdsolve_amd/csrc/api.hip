// Error reporting and version of the C ABI (include/zebsd.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "common.h"

namespace zb {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: launch failed: %s", what, hipGetErrorString(e));
        return ZB_ELAUNCH;
    }
    return ZB_OK;
}

}  // namespace zb

extern "C" const char* zb_version(void) { return "0.1.0"; }

extern "C" int32_t zb_last_error(char* buf, int32_t cap) {
    int32_t n = (int32_t)strlen(zb::g_err);
    if (buf && cap > 0) {
        int32_t k = n < cap - 1 ? n : cap - 1;
        memcpy(buf, zb::g_err, k);
        buf[k] = 0;
    }
    return n;
}
