// bbpd_common.hip -- error reporting and device queries of libbbpd_hip.
#include <hip/hip_runtime.h>
#include <string>

#include "bbpd_internal.hpp"

namespace bbpd {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return BBPD_ERR_HIP;
}

} // namespace bbpd

extern "C" const char *bbpd_last_error(void) { return bbpd::g_err.c_str(); }

extern "C" int bbpd_version(void) { return 100; }

extern "C" int bbpd_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
