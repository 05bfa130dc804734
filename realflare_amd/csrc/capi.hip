// Error reporting shared by every C-ABI entry point.
#include <cstdarg>
#include <cstdio>

#include "rf_common.h"

namespace rf {

static thread_local char g_error[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char* what) {
    return set_error(RF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace rf

extern "C" const char* rf_last_error(void) { return rf::g_error; }

extern "C" int rf_abi_version(void) { return RF_ABI_VERSION; }
