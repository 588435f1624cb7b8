// capi_util.hpp -- error reporting and RAII helpers of the extern "C" layer
// (capi.cpp, capi_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lqcd.h"

namespace lqcd {

// message of the last failed call on this host thread (lqcd_last_error)
std::string& last_error_slot();

inline int fail(int code, const std::string& msg)
{
    last_error_slot() = msg;
    return code;
}

struct DeviceGuard {
    explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};

// Temporary device buffer freed on scope exit.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

} // namespace lqcd

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::lqcd::fail(_e == hipErrorOutOfMemory ? LQCD_ERR_OUT_OF_MEMORY : LQCD_ERR_DEVICE, \
                                std::string(#expr) + ": " + hipGetErrorString(_e));    \
    } while (0)

#define CHECK_ARG(cond, msg)                                                           \
    do {                                                                               \
        if (!(cond)) return ::lqcd::fail(LQCD_ERR_INVALID_ARGUMENT, msg);              \
    } while (0)
