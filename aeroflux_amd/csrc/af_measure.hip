// af_measure.hip -- attainable streaming bandwidth of the access patterns the
// step's kernels use (measurement only; no reference counterpart).  A
// grid-stride loop over 16-byte elements, no reuse, HIP events around
// back-to-back launches on a private stream.  The arrays hold pseudo-random
// doubles in [-1, 1], like the fields the step moves: the memory system
// moves all-zero data measurably faster (a 256^3 FP64 colour pass takes
// about 29 us on zeroed arrays and 32.5 us on such values).
#include <hip/hip_runtime.h>

#include "../../include/aeroflux.h"
#include "af_host.h"

namespace af {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

// Element i of an array: a hash of (i, seed) mapped to [-1, 1] in steps of
// 1e-6.
__global__ __launch_bounds__(256) void k_stream_fill(double *p, int64_t n, unsigned seed)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        unsigned long long h = (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        p[i] = (double)(h % 2000001ull) * 1e-6 - 1.0;
    }
}

__global__ __launch_bounds__(256) void k_stream_read(const d2 *a, int64_t n2, double *sink)
{
    double s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * 256) {
        const d2 v = a[i];
        s += v.x + v.y;
    }
    // One store per workgroup keeps the loads live.
    sink[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_stream_copy(const d2 *a, d2 *b, int64_t n2)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * 256)
        b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_stream_2r1w(const d2 *a, const d2 *b, d2 *c, int64_t n2)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * 256)
        c[i] = a[i] + b[i];
}

__global__ __launch_bounds__(256) void k_stream_4r3w(const d2 *p, d2 *u, d2 *v, d2 *w, int64_t n2)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * 256) {
        const d2 q = p[i];
        u[i] = u[i] - q;
        v[i] = v[i] + q;
        w[i] = w[i] - q;
    }
}

constexpr int kGrid = 1024;  // workgroups (4 per CU)

}  // namespace
}  // namespace af

using namespace af;

extern "C" int af_measure_stream(int kind, int64_t n, int reps, double *gbps)
{
    static const int kArrays[4] = {1, 2, 3, 4};
    static const double kMoved[4] = {1, 2, 3, 7};  // array passes per launch
    if (kind < 0 || kind > 3 || n < 2 || (n & 1) || reps < 1 || !gbps)
        return fail(AF_ERR_ARG, "aeroflux: bad stream measurement arguments");
    const size_t bytes = (size_t)n * 8;
    const int64_t n2 = n / 2;
    d2 *buf[4] = {};
    double *sink = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipStreamCreate(&s);
    for (int i = 0; i < kArrays[kind] && err == hipSuccess; i++) {
        err = hipMalloc(&buf[i], bytes);
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_stream_fill, dim3(kGrid), dim3(256), 0, s,
                               reinterpret_cast<double *>(buf[i]), n, 17u + (unsigned)i);
            err = hipGetLastError();
        }
    }
    if (err == hipSuccess) err = hipMalloc(&sink, kGrid * sizeof(double));
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    auto launch = [&]() {
        switch (kind) {
        case 0: hipLaunchKernelGGL(k_stream_read, dim3(kGrid), dim3(256), 0, s, buf[0], n2, sink); break;
        case 1: hipLaunchKernelGGL(k_stream_copy, dim3(kGrid), dim3(256), 0, s, buf[0], buf[1], n2); break;
        case 2:
            hipLaunchKernelGGL(k_stream_2r1w, dim3(kGrid), dim3(256), 0, s, buf[0], buf[1], buf[2], n2);
            break;
        default:
            hipLaunchKernelGGL(k_stream_4r3w, dim3(kGrid), dim3(256), 0, s, buf[0], buf[1], buf[2],
                               buf[3], n2);
        }
        return hipGetLastError();
    };
    float ms = 0;
    if (err == hipSuccess) err = launch();
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    for (int r = 0; r < reps && err == hipSuccess; r++) err = launch();
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (d2 *b : buf)
        if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    if (s) (void)hipStreamDestroy(s);
    if (err != hipSuccess) return hip_fail(err, "stream measurement");
    *gbps = kMoved[kind] * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return AF_OK;
}
