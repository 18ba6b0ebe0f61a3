// ldpcp_common.hpp -- error reporting and HIP checks shared by the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/ldpcp.h"

namespace ldpcp {

void set_error(const char* fmt, ...);

#define LDPCP_HIP_TRY(expr)                                                                    \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            ::ldpcp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                               __LINE__);                                                       \
            return LDPCP_EDEVICE;                                                               \
        }                                                                                       \
    } while (0)

}  // namespace ldpcp
