// af_advect_lds.hip -- advection sweeps (stencil.go:21-119) with LDS-staged
// sources, for gfx950.
//
// The gather of advectComponent reads ~40 values per cell from three source
// fields -- trilinear taps, the diffusion stencil and faceVelocity's
// averages -- most of them also read by neighbouring threads, so a
// gather straight from global memory moves ~14x its input bytes through the
// L1 caches.  Here a workgroup owns a 64 x BY tile of (x, y) and marches
// through a chunk of z planes.  For each plane it stages, for all three
// source fields, the box of rows and columns its faces can touch, in a ring
// of R planes in LDS; every source value is read from global memory once
// per workgroup and the gather is served from LDS.
//
// The staged box holds every value a face reads when its departure point
// moves at most D cells along the sweep axis (D = 2 for the X and Y sweeps,
// 1 for the Z sweep, whose reach costs whole planes); a tap outside the box (a
// larger displacement, or a clamped departure far from the face) is read
// from global memory instead.  Which copy a value comes from never changes
// it, and the arithmetic is body::advect_value's, so results are
// bit-identical to the unstaged sweep.
//
// Box per sweep axis (padded coordinates, tile x in [xs, xs+64), y in
// [ys, ys+BY), current plane z):
//   X sweep: x in [xs-D, xs+64+D], y in [ys-1, ys+BY], z in [z-1, z+1]
//   Y sweep: x in [xs-1, xs+64],   y in [ys-D, ys+BY+D], z in [z-1, z+1]
//   Z sweep: x in [xs-1, xs+64],   y in [ys-1, ys+BY],   z in [z-D, z+D+1]
// (taps at base and base+1 on every axis, the diffusion stencil at +-1 on
// the sweep axis, and faceVelocity's averages one cell below on the other
// axes).  The tiles cover faces with x < g+NX; the U faces of the wall
// column x = g+NX are computed by extra workgroups from global memory.
#include "af_sweep_lds.h"
#include "af_sweep_stage.h"
#include "af_xcd.h"

#include <climits>
#include <cstdint>

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;

// nb = (x-tiles, y-tiles, z-chunks); the z-chunks split each tile column's
// face planes evenly (slds::zchunk).
template <int AX, int BY, int D, bool BC>
__global__ __launch_bounds__(BX *BY) void k_sweep_lds(SweepArgs a, int3 nb, int ncol)
{
    using B = slds::Box<AX, BY, D>;
    __shared__ double lds_mem[B::TOTAL];
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if ((int)blockIdx.x >= ntile) {
        slds::wall_face<AX, BC>(
            a, ((int)blockIdx.x - ntile) * (BX * BY) + threadIdx.y * BX + threadIdx.x, bad);
    } else {
        int bx, by, bz, zs, zend;
        xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
        slds::zchunk(a.G, nb.z, bz, zs, zend);
        auto src_of = [&](int c) -> const DField & { return c == 0 ? a.U : (c == 1 ? a.V : a.W); };
        slds::tile<AX, BY, D, BC>(a, src_of, (slds::lds_t *)lds_mem, bx, by, zs, zend, bad);
    }
    if (bad) *body::gp(a.flag) = 1;
}

// The same tiles with per-field staged boxes (af_sweep_stage.h).
// Three 8-wave workgroups per CU fit the LDS of every axis; the register
// budget for six waves per SIMD (80) is asked for explicitly.
template <int AX, int BY, bool BC>
__global__ __launch_bounds__(BX *BY) __attribute__((amdgpu_waves_per_eu(6))) void k_sweep_stg(
    SweepArgs a, int3 nb, int ncol)
{
    using L = stg::Layout<AX, BY>;
    __shared__ __attribute__((aligned(16))) double lds_mem[L::TOTAL];
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if ((int)blockIdx.x >= ntile) {
        slds::wall_face<AX, BC>(
            a, ((int)blockIdx.x - ntile) * (BX * BY) + threadIdx.y * BX + threadIdx.x, bad);
    } else {
        int bx, by, bz, zs, zend;
        xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
        slds::zchunk(a.G, nb.z, bz, zs, zend);
        auto src_of = [&](int c) -> const DField & { return c == 0 ? a.U : (c == 1 ? a.V : a.W); };
        stg::tile<AX, BY, BC>(a, src_of, (slds::lds_t *)lds_mem, bx, by, zs, zend, bad);
    }
    if (bad) *body::gp(a.flag) = 1;
}

template <int BY, int ZC>
hipError_t launch_stg(int axis, const SweepArgs &a, hipStream_t s)
{
    int3 nb = slds::tiles<BY, ZC>(a.G);
    if (a.zchunks > 0) nb.z = a.zchunks < a.G.NZ + 1 ? a.zchunks : a.G.NZ + 1;
    const int ncol = slds::wall_blocks<BY>(a.G);
    const dim3 grid((unsigned)(nb.x * nb.y * nb.z + ncol));
    const dim3 block(BX, BY);
    if (a.bc) {
        switch (axis) {
        case 0: hipLaunchKernelGGL((k_sweep_stg<0, BY, true>), grid, block, 0, s, a, nb, ncol); break;
        case 1: hipLaunchKernelGGL((k_sweep_stg<1, BY, true>), grid, block, 0, s, a, nb, ncol); break;
        default: hipLaunchKernelGGL((k_sweep_stg<2, BY, true>), grid, block, 0, s, a, nb, ncol); break;
        }
    } else {
        switch (axis) {
        case 0: hipLaunchKernelGGL((k_sweep_stg<0, BY, false>), grid, block, 0, s, a, nb, ncol); break;
        case 1: hipLaunchKernelGGL((k_sweep_stg<1, BY, false>), grid, block, 0, s, a, nb, ncol); break;
        default: hipLaunchKernelGGL((k_sweep_stg<2, BY, false>), grid, block, 0, s, a, nb, ncol); break;
        }
    }
    return hipGetLastError();
}

template <int BY, int D, int ZC>
hipError_t launch_lds(int axis, const SweepArgs &a, hipStream_t s)
{
    int3 nb = slds::tiles<BY, ZC>(a.G);
    if (a.zchunks > 0) nb.z = a.zchunks < a.G.NZ + 1 ? a.zchunks : a.G.NZ + 1;
    const int ncol = slds::wall_blocks<BY>(a.G);
    const dim3 grid((unsigned)(nb.x * nb.y * nb.z + ncol));
    const dim3 block(BX, BY);
    if (a.bc) {
        switch (axis) {
        case 0: hipLaunchKernelGGL((k_sweep_lds<0, BY, D, true>), grid, block, 0, s, a, nb, ncol); break;
        case 1: hipLaunchKernelGGL((k_sweep_lds<1, BY, D, true>), grid, block, 0, s, a, nb, ncol); break;
        default: hipLaunchKernelGGL((k_sweep_lds<2, BY, D, true>), grid, block, 0, s, a, nb, ncol); break;
        }
    } else {
        switch (axis) {
        case 0: hipLaunchKernelGGL((k_sweep_lds<0, BY, D, false>), grid, block, 0, s, a, nb, ncol); break;
        case 1: hipLaunchKernelGGL((k_sweep_lds<1, BY, D, false>), grid, block, 0, s, a, nb, ncol); break;
        default: hipLaunchKernelGGL((k_sweep_lds<2, BY, D, false>), grid, block, 0, s, a, nb, ncol); break;
        }
    }
    return hipGetLastError();
}

// Resident workgroups per CU of variant (BY, D)'s X-sweep kernel, and the
// device's CU count (0 on error).
template <class K>
int kernel_slots(K kernel, int threads, int *cus)
{
    int dev = 0, n = 0;
    *cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess)
        return 0;
    return n;
}

template <int BY, int D>
int lds_slots(int *cus, bool staged)
{
    return staged ? kernel_slots(k_sweep_stg<0, BY, true>, BX * BY, cus)
                  : kernel_slots(k_sweep_lds<0, BY, D, true>, BX * BY, cus);
}

template <int BY, int D, int ZC>
int lds_chunks(const Geom &G, bool staged = false)
{
    int cus = 0;
    const int occ = lds_slots<BY, D>(&cus, staged);
    const int planes = G.NZ + 1;
    const int3 nb = slds::tiles<BY, ZC>(G);
    if (occ <= 0 || cus < kXcds) return nb.z;
    const int64_t per_xcd = (int64_t)occ * (cus / kXcds);
    // A round of tiles costs the chunk's planes plus about one plane of
    // ring prologue; chunks of 6 to 64 planes are considered.
    int best = nb.z;
    int64_t best_cost = INT64_MAX;
    for (int n = (planes + 63) / 64; n <= planes / 6 || n == (planes + 63) / 64; n++) {
        const int64_t tiles = (int64_t)nb.x * nb.y * n;
        const int64_t rounds = ((tiles + kXcds - 1) / kXcds + per_xcd - 1) / per_xcd;
        const int64_t cost = rounds * ((planes + n - 1) / n + 1);
        if (cost < best_cost) {
            best_cost = cost;
            best = n;
        }
    }
    return best;
}

}  // namespace

int sweep_lds_variant_count() { return 6; }

// The z-chunk count per tile column of LDS-staged variant `variant` on grid
// G whose tiles best fill whole rounds of the workgroups each XCD holds at
// once (XCD x runs the x-th eighth of the tiles).  The tiles of a round run
// side by side and take about equally long, so a last round that is mostly
// empty costs nearly a full one.
int sweep_lds_chunks(const Geom &G, int variant)
{
    switch (variant) {
    case 0: return lds_chunks<4, 2, 16>(G);
    case 1: return lds_chunks<4, 2, 32>(G);
    case 2: return lds_chunks<8, 2, 16>(G);
    case 3: return lds_chunks<8, 2, 32>(G);
    case 4: return lds_chunks<8, 2, 16>(G, true);
    default: return lds_chunks<8, 2, 32>(G, true);
    }
}

// The shape's own chunk count: z-chunks of at most ZC planes.
int sweep_lds_default_chunks(const Geom &G, int variant)
{
    const int zc = variant == 1 || variant == 3 || variant == 5 ? 32 : 16;
    return (G.NZ + 1 + zc - 1) / zc;
}

// LDS-staged variants: (rows per workgroup, planes per z-chunk); staged reach
// two cells along x and y, one along z.
hipError_t launch_sweep_lds(int axis, const SweepArgs &a, int variant, hipStream_t s)
{
    switch (variant) {
    case 0: return launch_lds<4, 2, 16>(axis, a, s);
    case 1: return launch_lds<4, 2, 32>(axis, a, s);
    case 2: return launch_lds<8, 2, 16>(axis, a, s);
    case 3: return launch_lds<8, 2, 32>(axis, a, s);
    case 4: return launch_stg<8, 16>(axis, a, s);
    case 5: return launch_stg<8, 32>(axis, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace af
