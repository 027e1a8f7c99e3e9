// Error plumbing shared by the host model and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace sr {

// Records the message as the library's last error (thread-safe, copied).
void set_error(const std::string &msg);

}  // namespace sr

#define SR_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::sr::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));               \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)
