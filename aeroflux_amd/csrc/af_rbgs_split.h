// af_rbgs_split.h -- red-black passes on the colour-split layout.
#pragma once

#include "af_internal.h"

namespace af {

template <typename T>
struct SplitArgs {
    T *red, *black;                 // the two colours, interior cells only
    const T *rhs_red, *rhs_black;   // FP64: divergence; FP32: residual
    int NX, NY, NZ, NXH, g;         // NXH = (NX + 1) / 2
    double h2;
    const int *done = nullptr;      // if set and *done != 0, the launch does nothing
    int slab = 1 << 30;             // workgroup order: y-slabs of this many workgroup rows,
                                    // each walked through z (>= the rows: plane order)
};

// One colour pass.  mode 0: normal; 1: first pass from an all-zero field;
// 2: second pass from an all-zero field (own old values are +0).
template <typename T>
hipError_t launch_split_pass(const SplitArgs<T> &a, int color, int mode, hipStream_t s);

// Interior of a padded natural-layout field -> split arrays.
template <typename TS, typename TN>
hipError_t launch_to_split(const TN *nat, int sy, int sz, Geom G, TS *red, TS *black,
                           hipStream_t s);

// Split arrays -> interior of the FP64 natural-layout P (add: P += split).
template <typename TS>
hipError_t launch_from_split(double *nat, int sy, int sz, Geom G, const TS *red, const TS *black,
                             bool add, const int *done, hipStream_t s);

// FP32 right-hand side -(h2 * div) of the mixed-precision smoother, split.
hipError_t launch_rhs32_split(const double *div, int sy, int sz, Geom G, double h2, float *red,
                              float *black, hipStream_t s);

inline int64_t split_count(Geom G) { return (int64_t)G.NZ * G.NY * ((G.NX + 1) / 2); }

}  // namespace af
