// af_advect.hip -- semi-Lagrangian advection + directional diffusion sweeps
// (advectDiffuseAxis / advectComponent, stencil.go:21-119) for gfx950.
//
// Bit-for-bit the reference's IEEE binary64 evaluation, written so that the
// compiler can issue each face's loads together: no branches inside a face
// (the departure clamp is a select; a NaN coordinate only raises a flag that
// is stored once at the end), and every face value of a thread is computed
// before any store, so no store can order the next face's loads behind it.
//
// Two exact specialisations of sampleTrilinear (stencil.go:169-225):
//   * only the velocity component along the sweep axis enters the
//     backtrace (stencil.go:89-96), so faceVelocity's other two averages,
//     which have no other use, are not formed;
//   * on the two axes the backtrace does not move, the coordinate is the
//     face's own index, which lies in [0, extent-2] for every face a sweep
//     writes (x in [g, g+NX] of the NX+2g+1-wide U field, and likewise), so
//     the reference's clamp and x0 >= extent-1 adjustment never fire there:
//     the tap base is the face index and the weight is exactly 0.  All eight
//     taps are still loaded and combined in the reference's order (0*Inf
//     and signed zeros depend on the zero-weight taps).
#include "af_internal.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = 64;

__device__ __forceinline__ int idx(const DField &f, int x, int y, int z)
{
    return x + y * f.sy + z * f.sz;  // Field.Index, grid.go:24-26
}

// The velocity component along sweep axis AX at a face of component C, as
// faceVelocity (stencil.go:128-155) forms it (same four-point order).
template <int AX, int C>
__device__ __forceinline__ double axis_velocity(const SweepArgs &a, int x, int y, int z)
{
    if (AX == C) {
        const DField &f = C == 0 ? a.U : (C == 1 ? a.V : a.W);
        return f.d[idx(f, x, y, z)];
    }
    if (AX == 0) {
        const double *u = a.U.d;
        if (C == 1)
            return 0.25 * (u[idx(a.U, x, y - 1, z)] + u[idx(a.U, x + 1, y - 1, z)] +
                           u[idx(a.U, x, y, z)] + u[idx(a.U, x + 1, y, z)]);
        return 0.25 * (u[idx(a.U, x, y, z - 1)] + u[idx(a.U, x + 1, y, z - 1)] +
                       u[idx(a.U, x, y, z)] + u[idx(a.U, x + 1, y, z)]);
    }
    if (AX == 1) {
        const double *v = a.V.d;
        if (C == 0)
            return 0.25 * (v[idx(a.V, x - 1, y, z)] + v[idx(a.V, x, y, z)] +
                           v[idx(a.V, x - 1, y + 1, z)] + v[idx(a.V, x, y + 1, z)]);
        return 0.25 * (v[idx(a.V, x, y, z - 1)] + v[idx(a.V, x, y + 1, z - 1)] +
                       v[idx(a.V, x, y, z)] + v[idx(a.V, x, y + 1, z)]);
    }
    const double *w = a.W.d;
    if (C == 0)
        return 0.25 * (w[idx(a.W, x - 1, y, z)] + w[idx(a.W, x, y, z)] +
                       w[idx(a.W, x - 1, y, z + 1)] + w[idx(a.W, x, y, z + 1)]);
    return 0.25 * (w[idx(a.W, x, y - 1, z)] + w[idx(a.W, x, y, z)] +
                   w[idx(a.W, x, y - 1, z + 1)] + w[idx(a.W, x, y, z + 1)]);
}

// advectComponent's value at one face (stencil.go:83-115).  `bad` is set
// when the departure coordinate is NaN -- the reference's bounds-panic
// trigger; the value is then NaN and no load leaves the field.
template <int AX, int C>
__device__ __forceinline__ double advect_value(const SweepArgs &a, int x, int y, int z, bool &bad)
{
    const DField &s = C == 0 ? a.U : (C == 1 ? a.V : a.W);
    const double v = axis_velocity<AX, C>(a, x, y, z);
    const int coord = AX == 0 ? x : (AX == 1 ? y : z);
    const int ext = AX == 0 ? s.nx : (AX == 1 ? s.ny : s.nz);
    double p = (double)coord;
    p -= a.dt * v * a.invH;
    const double pmax = (double)(ext - 1);
    p = p < 0 ? 0.0 : (p > pmax ? pmax : p);  // NaN fails both tests
    const bool nan = p != p;
    bad = bad || nan;
    int i0 = nan ? 0 : (int)p;
    i0 = i0 >= ext - 1 ? ext - 2 : i0;
    const double t = p - (double)i0;

    const int x0 = AX == 0 ? i0 : x;
    const int y0 = AX == 1 ? i0 : y;
    const int z0 = AX == 2 ? i0 : z;
    const double tx = AX == 0 ? t : 0.0;
    const double ty = AX == 1 ? t : 0.0;
    const double tz = AX == 2 ? t : 0.0;
    const double *d = s.d;
    const int i000 = idx(s, x0, y0, z0);
    const int i100 = i000 + 1;
    const int i010 = i000 + s.sy;
    const int i110 = i010 + 1;
    const int i001 = i000 + s.sz;
    const int i101 = i001 + 1;
    const int i011 = i001 + s.sy;
    const int i111 = i011 + 1;
    const double c00 = d[i000] * (1 - tx) + d[i100] * tx;
    const double c10 = d[i010] * (1 - tx) + d[i110] * tx;
    const double c01 = d[i001] * (1 - tx) + d[i101] * tx;
    const double c11 = d[i011] * (1 - tx) + d[i111] * tx;
    const double c0 = c00 * (1 - ty) + c10 * ty;
    const double c1 = c01 * (1 - ty) + c11 * ty;
    const double adv = c0 * (1 - tz) + c1 * tz;

    const int i = idx(s, x, y, z);
    const int st = AX == 0 ? 1 : (AX == 1 ? s.sy : s.sz);
    double lap = a.lam * (d[i - st] - 2 * d[i] + d[i + st]);
    lap = a.lam != 0 ? lap : 0.0;  // `var lap float64` stays +0 when lam == 0
    return adv + lap;
}

// One launch per sweep, all three components; threads cover x in [g,g+NX],
// y in [g,g+NY] and ZR consecutive z in [g,g+NZ], and compute every
// component face that exists at their coordinates.  All face values are
// formed before any is stored.  The workgroup is 64 x BYT threads; the
// engine's tuner picks (BYT, ZR).
template <int AX, int BYT, int ZR>
__global__ __launch_bounds__(BX *BYT) void k_sweep(SweepArgs a, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = a.G.g;
    const int x = g + bx * BX + threadIdx.x;
    const int y = g + by * BYT + threadIdx.y;
    if (x > g + a.G.NX || y > g + a.G.NY) return;
    const bool xin = x < g + a.G.NX, yin = y < g + a.G.NY;
    bool bad = false;
    double vu[ZR], vv[ZR], vw[ZR];
    bool du[ZR], dv[ZR], dw[ZR];
#pragma unroll
    for (int r = 0; r < ZR; r++) {
        const int z = g + bz * ZR + r;
        const bool zok = z <= g + a.G.NZ, zin = z < g + a.G.NZ;
        du[r] = yin && zin;
        dv[r] = xin && zin;
        dw[r] = xin && yin && zok;
        vu[r] = vv[r] = vw[r] = 0;
        if (du[r]) vu[r] = advect_value<AX, 0>(a, x, y, z, bad);
        if (dv[r]) vv[r] = advect_value<AX, 1>(a, x, y, z, bad);
        if (dw[r]) vw[r] = advect_value<AX, 2>(a, x, y, z, bad);
    }
#pragma unroll
    for (int r = 0; r < ZR; r++) {
        const int z = g + bz * ZR + r;
        if (du[r]) a.T0.d[idx(a.T0, x, y, z)] = vu[r];
        if (dv[r]) a.T1.d[idx(a.T1, x, y, z)] = vv[r];
        if (dw[r]) a.T2.d[idx(a.T2, x, y, z)] = vw[r];
    }
    if (bad) *a.flag = 1;
}

template <int BYT, int ZRT>
hipError_t launch_shape(int axis, const SweepArgs &a, hipStream_t s)
{
    const int3 nb = make_int3((a.G.NX + 1 + BX - 1) / BX, (a.G.NY + 1 + BYT - 1) / BYT,
                              (a.G.NZ + 1 + ZRT - 1) / ZRT);
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

int sweep_shape_count() { return 8; }

// Workgroup shapes the tuner chooses from: (rows per block, z per thread).
hipError_t launch_sweep(int axis, const SweepArgs &a, hipStream_t s)
{
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
