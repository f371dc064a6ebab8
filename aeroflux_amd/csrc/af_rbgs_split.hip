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
// reference's, so the results are bit-identical.  The per-thread bodies and
// the layout's index rules are in af_bodies.h.
#include "af_bodies.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;
constexpr int BY = 4;

// Planes per thread: each thread walks zr<T>() planes of its column in z,
// reading the other colour's column once per plane into a register queue
// (body::split_pass_column); the queue's two prologue loads are paid once
// per zr<T>() planes.  Measured at 256^3 (one pass, us): FP64 4/8/16
// planes 30.8/29.7/29.6, FP32 8/16/32 planes 15.6/16.5/16.8.
template <typename T>
constexpr int zr()
{
    return 8;
}

template <typename T, int MODE>
__global__ __launch_bounds__(BX *BY) void k_split_pass(SplitArgs<T> a, int color, int3 nb)
{
    if (a.done && *a.done) return;
    int bx, by, bz;
    xcd_block3_slab(nb.x, nb.y, nb.z, a.slab, bx, by, bz);
    // FP64: the chunks software-pipelined (body::split_pass_column_pipelined).
    if constexpr (sizeof(T) == 8)
        body::split_pass_column_pipelined<T, MODE, zr<T>()>(a, color, bx * BX + threadIdx.x,
                                                            by * BY + threadIdx.y, bz * zr<T>());
    else
        body::split_pass_column<T, MODE, zr<T>()>(a, color, bx * BX + threadIdx.x,
                                                  by * BY + threadIdx.y, bz * zr<T>());
}

// Natural interior -> split (both colours).  `nat` uses the padded layout.
// Each thread converts the cell pair of split entry bx * BX + threadIdx.x
// (body::to_split_pair); the grid covers NXH entries per row.
__global__ __launch_bounds__(BX *BY) void k_to_split(const double *nat, int sy, int sz, int g,
                                                     int NX, int NY, int NXH, double *red,
                                                     double *black, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    body::to_split_pair(nat, sy, sz, g, NX, NY, NXH, red, black, bx * BX + threadIdx.x,
                        by * BY + threadIdx.y, bz);
}

// Split -> natural interior, a cell pair per thread (body::from_split_pair).
template <typename TS, bool ADD>
__global__ __launch_bounds__(BX *BY) void k_from_split(double *nat, int sy, int sz, int g, int NX,
                                                       int NY, int NXH, const TS *red,
                                                       const TS *black, const int *done, int3 nb)
{
    if (done && *done) return;
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    body::from_split_pair<TS, ADD>(nat, sy, sz, g, NX, NY, NXH, red, black, bx * BX + threadIdx.x,
                                   by * BY + threadIdx.y, bz);
}

// Gated form (the mixed solve's later correction cycles, which are skipped
// once the target is met): a fixed grid of workgroups strides over the
// tiles, so a skipped launch dispatches a few thousand workgroups that exit
// at once instead of one per tile.
template <typename TS, bool ADD>
__global__ __launch_bounds__(BX *BY) void k_from_split_gated(double *nat, int sy, int sz, int g,
                                                             int NX, int NY, int NXH,
                                                             const TS *red, const TS *black,
                                                             const int *done, int3 nb)
{
    if (*done) return;
    const int total = nb.x * nb.y * nb.z;
    for (int l = blockIdx.x; l < total; l += gridDim.x) {
        const int bx = l % nb.x, r = l / nb.x;
        const int by = r % nb.y, bz = r / nb.y;
        body::from_split_pair<TS, ADD>(nat, sy, sz, g, NX, NY, NXH, red, black,
                                       bx * BX + threadIdx.x, by * BY + threadIdx.y, bz);
    }
}

constexpr int kGatedGrid = 2048;

__global__ __launch_bounds__(BX *BY) void k_rhs32_split(const double *div, int sy, int sz, int g,
                                                        int NX, int NY, int NXH, double h2,
                                                        float *red, float *black, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    body::rhs32_pair(div, sy, sz, g, NX, NY, NXH, h2, red, black, bx * BX + threadIdx.x,
                     by * BY + threadIdx.y, bz);
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
    const int3 nb = nblocks(a.NXH, a.NY, (a.NZ + zr<T>() - 1) / zr<T>());
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
    static_assert(sizeof(TS) == 8 && sizeof(TN) == 8, "FP64 natural -> FP64 split only");
    const int NXH = (G.NX + 1) / 2;
    const int3 nb = nblocks(NXH, G.NY, G.NZ);
    hipLaunchKernelGGL(k_to_split, grid1(nb), dim3(BX, BY), 0, s, nat, sy, sz, G.g, G.NX, G.NY,
                       NXH, red, black, nb);
    return hipGetLastError();
}

template <typename TS>
hipError_t launch_from_split(double *nat, int sy, int sz, Geom G, const TS *red, const TS *black,
                             bool add, const int *done, hipStream_t s)
{
    const int NXH = (G.NX + 1) / 2;
    const int3 nb = nblocks(NXH, G.NY, G.NZ);
    const int total = nb.x * nb.y * nb.z;
    if (done) {
        const dim3 grid((unsigned)(total < kGatedGrid ? total : kGatedGrid));
        if (add)
            hipLaunchKernelGGL((k_from_split_gated<TS, true>), grid, dim3(BX, BY), 0, s, nat, sy,
                               sz, G.g, G.NX, G.NY, NXH, red, black, done, nb);
        else
            hipLaunchKernelGGL((k_from_split_gated<TS, false>), grid, dim3(BX, BY), 0, s, nat, sy,
                               sz, G.g, G.NX, G.NY, NXH, red, black, done, nb);
    } else if (add)
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
    const int3 nb = nblocks(NXH, G.NY, G.NZ);
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
