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
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;

template <int AX, int BY, int D, int ZC>
__global__ __launch_bounds__(BX *BY) void k_sweep_lds(SweepArgs a, int3 nb, int ncol)
{
    using B = slds::Box<AX, BY, D>;
    __shared__ double lds_mem[B::TOTAL];
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if ((int)blockIdx.x >= ntile) {
        slds::wall_face<AX, false>(
            a, ((int)blockIdx.x - ntile) * (BX * BY) + threadIdx.y * BX + threadIdx.x, bad);
    } else {
        int bx, by, bz;
        xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
        auto src_of = [&](int c) -> const DField & { return c == 0 ? a.U : (c == 1 ? a.V : a.W); };
        slds::tile<AX, BY, D, ZC, false>(a, src_of, (slds::lds_t *)lds_mem, bx, by, bz, bad);
    }
    if (bad) *body::gp(a.flag) = 1;
}

template <int BY, int D, int ZC>
hipError_t launch_lds(int axis, const SweepArgs &a, hipStream_t s)
{
    const int3 nb = slds::tiles<BY, ZC>(a.G);
    const int ncol = slds::wall_blocks<BY>(a.G);
    const dim3 grid((unsigned)(nb.x * nb.y * nb.z + ncol));
    const dim3 block(BX, BY);
    switch (axis) {
    case 0: hipLaunchKernelGGL((k_sweep_lds<0, BY, D, ZC>), grid, block, 0, s, a, nb, ncol); break;
    case 1: hipLaunchKernelGGL((k_sweep_lds<1, BY, D, ZC>), grid, block, 0, s, a, nb, ncol); break;
    default: hipLaunchKernelGGL((k_sweep_lds<2, BY, D, ZC>), grid, block, 0, s, a, nb, ncol); break;
    }
    return hipGetLastError();
}

}  // namespace

int sweep_lds_variant_count() { return 4; }

// LDS-staged variants: (rows per workgroup, planes per z-chunk); staged reach
// two cells along x and y, one along z.
hipError_t launch_sweep_lds(int axis, const SweepArgs &a, int variant, hipStream_t s)
{
    switch (variant) {
    case 0: return launch_lds<4, 2, 16>(axis, a, s);
    case 1: return launch_lds<4, 2, 32>(axis, a, s);
    case 2: return launch_lds<8, 2, 16>(axis, a, s);
    case 3: return launch_lds<8, 2, 32>(axis, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace af
