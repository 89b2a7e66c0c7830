// Error state and device helpers shared by the LBM and LSM parts.
#include "psx_common.h"

#include <cstring>

namespace psx {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
        return PSX_EHIP;
    }
    return PSX_OK;
}

}  // namespace psx

extern "C" const char *psx_last_error(void) { return psx::g_error.c_str(); }

extern "C" int psx_device_info(char *name, int32_t name_len, int32_t *num_cu, double *hbm_gb) {
    int rc = psx::require_device();
    if (rc) return rc;
    int dev = 0;
    PSX_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PSX_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cu) *num_cu = prop.multiProcessorCount;
    if (hbm_gb) *hbm_gb = (double)prop.totalGlobalMem / 1e9;
    return PSX_OK;
}
