// af_advect.hip -- semi-Lagrangian advection + directional diffusion sweeps
// (advectDiffuseAxis / advectComponent, stencil.go:21-119) for gfx950.
//
// One launch per sweep covers all three components: threads cover x in
// [g, g+NX], y in [g, g+NY] and ZR consecutive z in [g, g+NZ], and compute
// every component face that exists at their coordinates
// (body::sweep_column).  The workgroup is 64 x BYT threads; blocks run in
// y-slab order (a face reads the planes above and below it in six fields,
// more than an XCD's L2 holds for whole planes of a large grid).  The
// engine's tuner picks (BYT, ZR) and the slab height per grid.
#include "af_bodies.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;

template <int AX, int BYT, int ZR>
__global__ __launch_bounds__(BX *BYT) void k_sweep(SweepArgs a, int3 nb)
{
    int bx, by, bz;
    xcd_block3_slab(nb.x, nb.y, nb.z, a.slab, bx, by, bz);
    const int g = a.G.g;
    body::sweep_column<AX, ZR>(a, g + bx * BX + threadIdx.x, g + by * BYT + threadIdx.y,
                               g + bz * ZR);
}

template <int BYT, int ZRT>
hipError_t launch_shape(int axis, SweepArgs a, hipStream_t s)
{
    const int3 nb = make_int3((a.G.NX + 1 + BX - 1) / BX, (a.G.NY + 1 + BYT - 1) / BYT,
                              (a.G.NZ + 1 + ZRT - 1) / ZRT);
    // Slab height in block rows (at least one).
    a.slab = a.slab >= nb.y * BYT ? nb.y : (a.slab + BYT - 1) / BYT;
    if (a.slab < 1) a.slab = 1;
    const dim3 grid((unsigned)(nb.x * nb.y * nb.z));
    const dim3 block(BX, BYT);
    switch (axis) {
    case 0: hipLaunchKernelGGL((k_sweep<0, BYT, ZRT>), grid, block, 0, s, a, nb); break;
    case 1: hipLaunchKernelGGL((k_sweep<1, BYT, ZRT>), grid, block, 0, s, a, nb); break;
    default: hipLaunchKernelGGL((k_sweep<2, BYT, ZRT>), grid, block, 0, s, a, nb); break;
    }
    return hipGetLastError();
}

}  // namespace

int sweep_shape_count() { return 8 + sweep_lds_variant_count(); }

// Shapes the tuner chooses from: 0-7 gather from global memory with
// workgroups of (rows per block, z per thread); 8 and up stage the sources
// in LDS (af_advect_lds.hip).
hipError_t launch_sweep(int axis, const SweepArgs &a, hipStream_t s)
{
    if (a.shape >= 8) return launch_sweep_lds(axis, a, a.shape - 8, s);
    switch (a.shape) {
    case 0: return launch_shape<1, 1>(axis, a, s);
    case 1: return launch_shape<1, 2>(axis, a, s);
    case 2: return launch_shape<2, 1>(axis, a, s);
    case 3: return launch_shape<2, 2>(axis, a, s);
    case 4: return launch_shape<4, 1>(axis, a, s);
    case 5: return launch_shape<4, 2>(axis, a, s);
    case 6: return launch_shape<8, 1>(axis, a, s);
    case 7: return launch_shape<16, 1>(axis, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace af
