// ldpcp_common.cpp -- thread-local error text and device query for the C ABI.
#include "ldpcp_common.hpp"

#include <string.h>

namespace ldpcp {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

}  // namespace ldpcp

extern "C" const char* ldpcp_last_error(void) { return ldpcp::g_error; }

extern "C" int ldpcp_device_info(char* name, int32_t name_len, int32_t* n_cu, int64_t* hbm_bytes)
{
    int dev = 0;
    hipDeviceProp_t prop;
    LDPCP_HIP_TRY(hipGetDevice(&dev));
    LDPCP_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return LDPCP_OK;
}
