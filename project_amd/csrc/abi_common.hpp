// Host-side helpers shared by the C-ABI translation units: thread-local error text and
// HIP status checking.  No exception crosses the extern "C" boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/uppe.h"

namespace uppe {
void set_error(const char *fmt, ...);
}

#define UPPE_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            uppe::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),         \
                            __FILE__, __LINE__);                                           \
            return UPPE_EHIP;                                                              \
        }                                                                                  \
    } while (0)

#define UPPE_CHECK(cond, ...)                                                              \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            uppe::set_error(__VA_ARGS__);                                                  \
            return UPPE_EINVAL;                                                            \
        }                                                                                  \
    } while (0)
