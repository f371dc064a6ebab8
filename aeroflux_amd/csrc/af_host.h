// af_host.h -- host-side helpers shared by the engine (af_engine.hip) and the
// batched engine (af_batch.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/aeroflux.h"

namespace af {

// Records "aeroflux: ..." as this thread's last error and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t err, const char *what);
const char *last_error();

// NewMAC's field order and 128-byte padding (grid.go:101-134): fills the
// eight descriptors and returns the arena size in bytes.
int64_t mac_layout(af_field *f, int64_t nx, int64_t ny, int64_t nz, int64_t g);

}  // namespace af

#define HIP_TRY(expr, what)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return af::hip_fail(_e, what);  \
    } while (0)
