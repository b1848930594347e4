// Error reporting shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

namespace mfx {
// Records a message for mfx_last_error() (thread local) and returns ``code``.
int fail(int code, const char* fmt, const char* a);
// 0 for hipSuccess, else records the HIP error string and returns the error code.
int hip_status(hipError_t e);
}  // namespace mfx
