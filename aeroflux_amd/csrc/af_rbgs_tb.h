// af_rbgs_tb.h -- temporally blocked red-black Gauss-Seidel launcher.
#pragma once

#include <hip/hip_runtime.h>

namespace af {

template <typename T>
struct TbArgs {
    const T *src;  // field before the K passes (ignored when zero_src)
    T *dst;        // field after the K passes (interior cells only)
    const T *rhs;  // FP64: divergence (p = (sum - h2*div)/6); FP32: residual
    int NX, NY, NZ, g;
    int sy, sz;    // strides of the padded P layout
    double h2;
    int color0;    // colour of the first of the K passes
    int cz;        // z-chunk per workgroup
    const int *done = nullptr;  // if set and *done != 0, the launch does nothing
};

// K passes per launch (2, 4 or 6), `rows` rows per thread (1 or 2).
template <typename T>
hipError_t launch_rbgs_tb(const TbArgs<T> &a, int k, int rows, bool zero_src, hipStream_t s);

bool rbgs_tb_supported(int k, int rows);

// Raises the dynamic-LDS limit of every instantiation; call once per device.
hipError_t rbgs_tb_init();

}  // namespace af
