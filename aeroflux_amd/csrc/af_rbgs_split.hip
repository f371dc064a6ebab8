// af_rbgs_split.hip -- red-black Gauss-Seidel on a colour-split layout.
//
// A colour pass of gaussSeidelSweep (project.go:103-130) writes every cell of
// one colour from its six neighbours -- all of the other colour -- and its
// divergence; it never reads the cell's own old value (except where the
// Neumann mirror substitutes it for a ghost neighbour).  Storing the two
// colours in separate compacted arrays therefore lets a pass stream
// (read other colour + read divergence + write this colour) = 12 B per cell
// in FP64 instead of the 24 B a pass over the interleaved arena needs, with
// every access unit-stride.  The arithmetic and its order are the
// reference's, so the results are bit-identical.
//
// Split layout: interior cells only, row (y', z') of colour c holds the
// cells x' = x0 + 2k (x0 = first x' of colour c in that row) at index
// ((z'*NY + y') * NXH + k), NXH = ceil(NX/2).  For a cell of colour c at
// x', its x-neighbours are entries (x'-1)>>1 and (x'+1)>>1 of the
// other-colour row, and its y/z neighbours are entry x'>>1 of the adjacent
// other-colour rows -- the same k.
#include "af_rbgs_split.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = 64;
constexpr int BY = 4;

template <typename T>
__device__ __forceinline__ T relax_s(T sum, T rhs, double h2);

template <>
__device__ __forceinline__ double relax_s<double>(double sum, double rhs, double h2)
{
    return (sum - h2 * rhs) / 6.0;
}

template <>
__device__ __forceinline__ float relax_s<float>(float sum, float rhs, double)
{
    return (sum + rhs) * (1.0f / 6.0f);
}

// Parity of interior cell (x', y', z') in padded coordinates: red = 0.
__device__ __forceinline__ int parity(int xi, int yi, int zi, int g)
{
    return (xi + yi + zi + 3 * g) & 1;
}

// Planes per thread: a pass over one plane per thread leaves each wave too
// little work to keep the CUs full, so every thread updates ZR planes.
constexpr int ZR = 4;

// MODE 0: normal pass.  MODE 1: first pass of a solve from zero (every
// neighbour is +0).  MODE 2: second pass of a solve from zero (the cell's own
// old value, used for mirrored ghosts, is +0).  All ZR results are formed
// before any is stored, so every load can be issued up front.
template <typename T, int MODE>
__global__ __launch_bounds__(BX *BY) void k_split_pass(SplitArgs<T> a, int color, int3 nb)
{
    if (a.done && *a.done) return;
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int k = bx * BX + threadIdx.x;
    const int yi = by * BY + threadIdx.y;
    const int NX = a.NX, NY = a.NY, NZ = a.NZ, NXH = a.NXH;
    if (yi >= NY) return;
    T *cur = color ? a.black : a.red;
    const T *oth = color ? a.red : a.black;
    const T *rhs = color ? a.rhs_black : a.rhs_red;
    T out[ZR];
    int at[ZR];
#pragma unroll
    for (int r = 0; r < ZR; r++) {
        const int zi = bz * ZR + r;
        const int x0 = (color + yi + zi + 3 * a.g) & 1;
        const int xi = x0 + 2 * k;
        at[r] = -1;
        if (zi >= NZ || xi >= NX) continue;
        const int row = zi * NY + yi;
        const int i = row * NXH + k;
        at[r] = i;
        T sum;
        if (MODE == 1) {
            sum = T(0) + T(0) + T(0) + T(0) + T(0) + T(0);
        } else {
            const T c = MODE == 2 ? T(0) : ((xi == 0 || xi == NX - 1 || yi == 0 || yi == NY - 1 ||
                                             zi == 0 || zi == NZ - 1)
                                                ? cur[i]
                                                : T(0));
            const T xm = xi == 0 ? c : oth[row * NXH + ((xi - 1) >> 1)];
            const T xp = xi == NX - 1 ? c : oth[row * NXH + ((xi + 1) >> 1)];
            const T ym = yi == 0 ? c : oth[i - NXH];
            const T yp = yi == NY - 1 ? c : oth[i + NXH];
            const T zm = zi == 0 ? c : oth[i - NY * NXH];
            const T zp = zi == NZ - 1 ? c : oth[i + NY * NXH];
            sum = xm + xp + ym + yp + zm + zp;
        }
        out[r] = relax_s<T>(sum, rhs[i], a.h2);
    }
#pragma unroll
    for (int r = 0; r < ZR; r++)
        if (at[r] >= 0) cur[at[r]] = out[r];
}

// Natural interior -> split (both colours).  `nat` uses the padded layout.
template <typename TS, typename TN>
__global__ __launch_bounds__(BX *BY) void k_to_split(const TN *nat, int sy, int sz, int g, int NX,
                                                     int NY, int NXH, TS *red, TS *black, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int xi = bx * BX + threadIdx.x;
    const int yi = by * BY + threadIdx.y;
    const int zi = bz;
    if (xi >= NX || yi >= NY) return;
    const TN v = nat[(xi + g) + (yi + g) * sy + (zi + g) * sz];
    const int i = (zi * NY + yi) * NXH + (xi >> 1);
    if (parity(xi, yi, zi, g)) black[i] = (TS)v; else red[i] = (TS)v;
}

// Split -> natural interior of P (FP64).  ADD: P += split (the correction
// of the mixed-precision solve); else P = split.
template <typename TS, bool ADD>
__global__ __launch_bounds__(BX *BY) void k_from_split(double *nat, int sy, int sz, int g, int NX,
                                                       int NY, int NXH, const TS *red,
                                                       const TS *black, const int *done, int3 nb)
{
    if (done && *done) return;
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int xi = bx * BX + threadIdx.x;
    const int yi = by * BY + threadIdx.y;
    const int zi = bz;
    if (xi >= NX || yi >= NY) return;
    const int i = (zi * NY + yi) * NXH + (xi >> 1);
    const double v = (double)(parity(xi, yi, zi, g) ? black[i] : red[i]);
    const int n = (xi + g) + (yi + g) * sy + (zi + g) * sz;
    nat[n] = ADD ? nat[n] + v : v;
}

__global__ __launch_bounds__(BX *BY) void k_rhs32_split(const double *div, int sy, int sz, int g,
                                                        int NX, int NY, int NXH, double h2,
                                                        float *red, float *black, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int xi = bx * BX + threadIdx.x;
    const int yi = by * BY + threadIdx.y;
    const int zi = bz;
    if (xi >= NX || yi >= NY) return;
    const float v = (float)(-(h2 * div[(xi + g) + (yi + g) * sy + (zi + g) * sz]));
    const int i = (zi * NY + yi) * NXH + (xi >> 1);
    if (parity(xi, yi, zi, g)) black[i] = v; else red[i] = v;
}

inline int3 nblocks(int nx, int ny, int nz)
{
    return make_int3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
}
inline dim3 grid1(int3 nb) { return dim3((unsigned)(nb.x * nb.y * nb.z)); }

}  // namespace

template <typename T>
hipError_t launch_split_pass(const SplitArgs<T> &a, int color, int mode, hipStream_t s)
{
    const int3 nb = nblocks(a.NXH, a.NY, (a.NZ + ZR - 1) / ZR);
    const dim3 blk(BX, BY);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_split_pass<T, 0>), grid1(nb), blk, 0, s, a, color, nb); break;
    case 1: hipLaunchKernelGGL((k_split_pass<T, 1>), grid1(nb), blk, 0, s, a, color, nb); break;
    default: hipLaunchKernelGGL((k_split_pass<T, 2>), grid1(nb), blk, 0, s, a, color, nb); break;
    }
    return hipGetLastError();
}

template <typename TS, typename TN>
hipError_t launch_to_split(const TN *nat, int sy, int sz, Geom G, TS *red, TS *black, hipStream_t s)
{
    const int NXH = (G.NX + 1) / 2;
    const int3 nb = nblocks(G.NX, G.NY, G.NZ);
    hipLaunchKernelGGL((k_to_split<TS, TN>), grid1(nb), dim3(BX, BY), 0, s, nat, sy, sz, G.g, G.NX,
                       G.NY, NXH, red, black, nb);
    return hipGetLastError();
}

template <typename TS>
hipError_t launch_from_split(double *nat, int sy, int sz, Geom G, const TS *red, const TS *black,
                             bool add, const int *done, hipStream_t s)
{
    const int NXH = (G.NX + 1) / 2;
    const int3 nb = nblocks(G.NX, G.NY, G.NZ);
    if (add)
        hipLaunchKernelGGL((k_from_split<TS, true>), grid1(nb), dim3(BX, BY), 0, s, nat, sy, sz,
                           G.g, G.NX, G.NY, NXH, red, black, done, nb);
    else
        hipLaunchKernelGGL((k_from_split<TS, false>), grid1(nb), dim3(BX, BY), 0, s, nat, sy, sz,
                           G.g, G.NX, G.NY, NXH, red, black, done, nb);
    return hipGetLastError();
}

hipError_t launch_rhs32_split(const double *div, int sy, int sz, Geom G, double h2, float *red,
                              float *black, hipStream_t s)
{
    const int NXH = (G.NX + 1) / 2;
    const int3 nb = nblocks(G.NX, G.NY, G.NZ);
    hipLaunchKernelGGL(k_rhs32_split, grid1(nb), dim3(BX, BY), 0, s, div, sy, sz, G.g, G.NX, G.NY,
                       NXH, h2, red, black, nb);
    return hipGetLastError();
}

template hipError_t launch_split_pass<double>(const SplitArgs<double> &, int, int, hipStream_t);
template hipError_t launch_split_pass<float>(const SplitArgs<float> &, int, int, hipStream_t);
template hipError_t launch_to_split<double, double>(const double *, int, int, Geom, double *,
                                                    double *, hipStream_t);
template hipError_t launch_from_split<double>(double *, int, int, Geom, const double *,
                                              const double *, bool, const int *, hipStream_t);
template hipError_t launch_from_split<float>(double *, int, int, Geom, const float *, const float *,
                                             bool, const int *, hipStream_t);

}  // namespace af
