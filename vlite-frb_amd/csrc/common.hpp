// Internal helpers shared by the HIP and host translation units of libvfrb.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "vfrb.h"

namespace vfrb {

// Records a printf-style message as this thread's last error.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

// Maps a HIP error to VFRB_ERR_HIP / VFRB_ERR_NOMEM and records it.
vfrb_status hip_fail(hipError_t err, const char* what, const char* file, int line);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace vfrb

#define VFRB_HIP(call)                                                      \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) return ::vfrb::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define VFRB_LAUNCH_CHECK(name)                                             \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return ::vfrb::hip_fail(e_, name, __FILE__, __LINE__); \
    } while (0)

#define VFRB_REQUIRE(cond, code, ...)                                       \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::vfrb::set_error(__VA_ARGS__);                                 \
            return (code);                                                  \
        }                                                                   \
    } while (0)
