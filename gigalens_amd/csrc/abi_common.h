// Error reporting shared by the translation units of the C ABI: every entry point
// returns a gl status code and leaves a thread-local message for gl_last_error().
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "gigalens_amd.h"

namespace glc {
int fail(int code, const std::string& msg);
}  // namespace glc

#define GL_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return glc::fail(GL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
