// af_ic.h -- device generator of the seeded band-limited initial velocity.
#pragma once

#include "af_internal.h"

namespace af {

// Device scratch: mode coefficients, per-axis phase tables, two partial-sum
// cubes and reduction partials.  Grows on demand (a growth frees the old
// buffer, which waits for the device); generations on one stream reuse it
// in stream order.
struct ICScratch {
    void *dev = nullptr;
    int64_t dev_bytes = 0;

    hipError_t reserve(int64_t bytes);
    void release();
};

int64_t ic_scratch_bytes(Geom G, int kmax);

// Writes the band-limited field into the interior faces of U, V, W (the
// other cells are left as they are) on stream s.
hipError_t launch_band_limited(DField U, DField V, DField W, Geom G, double h, uint64_t seed,
                               int kmax, double slope, double rms, ICScratch &sc, hipStream_t s);

}  // namespace af
