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
#include "af_bodies.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;

// LDS pointers carry their address space so that the compiler keeps LDS
// reads as LDS instructions: through a generic pointer it may merge an LDS
// read and its global fallback into one flat load, which takes the vector
// memory path the staging is meant to relieve.
typedef __attribute__((address_space(3))) double lds_t;

template <int AX, int BY, int DD>
struct Box {
    // The Z sweep's reach costs whole planes of the ring (2D+2 planes of
    // three fields), the X and Y sweeps' only columns or rows: the Z sweep
    // stages one cell of reach, the others DD.
    static constexpr int D = AX == 2 ? 1 : DD;
    static constexpr int XLO = AX == 0 ? D : 1;
    static constexpr int XW = AX == 0 ? BX + 2 * D + 1 : BX + 2;
    static constexpr int YLO = AX == 1 ? D : 1;
    static constexpr int YW = AX == 1 ? BY + 2 * D + 1 : BY + 2;
    static constexpr int ZLO = AX == 2 ? D : 1;
    static constexpr int R = AX == 2 ? 2 * D + 2 : 3;
    static constexpr int PLANE = XW * YW;  // doubles per staged plane of one field
    static constexpr int FIELD = PLANE * R;
    static constexpr int TOTAL = 3 * FIELD;
};

// Serves a face's values from the LDS ring.  The velocity averages and the
// diffusion stencil of a face inside the tile always lie in the staged box
// (the box is built from them), so they are read from LDS unchecked; the
// trilinear taps are read from LDS when their 2x2x2 cube lies in the box and
// inside the field, else from global memory.
template <class B>
struct LdsLoad {
    const SweepArgs &a;
    const lds_t *lds;
    int ox, oy;  // padded coordinates of the box's first column and row
    int zlo;     // first plane of the current window
    __device__ __forceinline__ const DField &f(int c) const
    {
        return c == 0 ? a.U : (c == 1 ? a.V : a.W);
    }
    __device__ __forceinline__ int slot(int c, int x, int y, int z) const
    {
        return c * B::FIELD + (z % B::R) * B::PLANE + (y - oy) * B::XW + (x - ox);
    }
    __device__ __forceinline__ double one(int c, int x, int y, int z) const
    {
        return lds[slot(c, x, y, z)];
    }
    __device__ __forceinline__ double2 pair(int c, int x, int y, int z) const
    {
        const int k = slot(c, x, y, z);
        return make_double2(lds[k], lds[k + 1]);
    }
    __device__ __forceinline__ void taps(int c, int x, int y, int z, double2 &p00, double2 &p10,
                                         double2 &p01, double2 &p11) const
    {
        const DField &s = f(c);
        // Base and base+1 inside the box on every axis, and inside the field
        // (the box is clipped to the field when staged).
        const bool in = (unsigned)(x - ox) < (unsigned)(B::XW - 1) &&
                        (unsigned)(y - oy) < (unsigned)(B::YW - 1) &&
                        (unsigned)(z - zlo) < (unsigned)(B::R - 1) && x >= 0 && x + 1 < s.nx &&
                        y >= 0 && y + 1 < s.ny && z >= 0 && z + 1 < s.nz;
        if (in) {
            const int k = slot(c, x, y, z);
            const int k1 = slot(c, x, y, z + 1);
            p00 = make_double2(lds[k], lds[k + 1]);
            p10 = make_double2(lds[k + B::XW], lds[k + B::XW + 1]);
            p01 = make_double2(lds[k1], lds[k1 + 1]);
            p11 = make_double2(lds[k1 + B::XW], lds[k1 + B::XW + 1]);
            return;
        }
        const double *d = s.d + body::idx(s, x, y, z);
        p00 = body::ld2(d);
        p10 = body::ld2(d + s.sy);
        p01 = body::ld2(d + s.sz);
        p11 = body::ld2(d + s.sz + s.sy);
    }
};

// Staging of one plane of the three source fields, split over the
// workgroup: thread t owns elements t, t+NT, ... of the flattened
// (field, row, column) box.  Their plane-independent offsets are computed
// once; per plane a thread issues all its global loads into registers
// (`fetch`) and later writes them to the plane's ring slot (`store`), so the
// next plane's loads are in flight while the current plane is computed.
template <class B, int NT>
struct Stager {
    static constexpr int N = (3 * B::PLANE + NT - 1) / NT;
    const double *src[N];  // element's (x, y) address in its field, plane 0
    int sz[N];             // its field's plane stride
    int nz[N];             // its field's plane count
    int off[N];            // LDS offset within a ring slot
    bool ok[N];            // inside the box and inside the field
    double v[N];

    __device__ __forceinline__ void init(const SweepArgs &a, int ox, int oy)
    {
        const int t = threadIdx.y * blockDim.x + threadIdx.x;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const int e = t + k * NT;
            const int c = e / B::PLANE;
            const int r = (e - c * B::PLANE) / B::XW;
            const int col = e - c * B::PLANE - r * B::XW;
            // Member-wise selects keep the descriptors in registers.
            double *d = c == 0 ? a.U.d : (c == 1 ? a.V.d : a.W.d);
            const int fnx = c == 0 ? a.U.nx : (c == 1 ? a.V.nx : a.W.nx);
            const int fny = c == 0 ? a.U.ny : (c == 1 ? a.V.ny : a.W.ny);
            const int fsy = c == 0 ? a.U.sy : (c == 1 ? a.V.sy : a.W.sy);
            const int fsz = c == 0 ? a.U.sz : (c == 1 ? a.V.sz : a.W.sz);
            const int fnz = c == 0 ? a.U.nz : (c == 1 ? a.V.nz : a.W.nz);
            const int x = ox + col, y = oy + r;
            ok[k] = e < 3 * B::PLANE && (unsigned)x < (unsigned)fnx && (unsigned)y < (unsigned)fny;
            src[k] = d + (ok[k] ? x + y * fsy : 0);
            sz[k] = fsz;
            nz[k] = fnz;
            off[k] = c * B::FIELD + r * B::XW + col;
        }
    }
    __device__ __forceinline__ void fetch(int z)
    {
#pragma unroll
        for (int k = 0; k < N; k++)
            if (ok[k] && (unsigned)z < (unsigned)nz[k]) v[k] = src[k][(size_t)z * sz[k]];
    }
    __device__ __forceinline__ void store(lds_t *lds, int z) const
    {
        if (z < 0) return;
        const int base = (z % B::R) * B::PLANE;
#pragma unroll
        for (int k = 0; k < N; k++)
            if (ok[k] && z < nz[k]) lds[base + off[k]] = v[k];
    }
};

// Tiles: nb.x x-tiles of 64 over [g, g+NX), nb.y y-tiles of BY over
// [g, g+NY], nb.z chunks of ZC planes over [g, g+NZ]; then `ncol` blocks of
// BX*BY threads for the U faces at x = g+NX.
template <int AX, int BY, int D, int ZC>
__global__ __launch_bounds__(BX *BY) void k_sweep_lds(SweepArgs a, int3 nb, int ncol)
{
    using B = Box<AX, BY, D>;
    __shared__ double lds_mem[B::TOTAL];
    lds_t *lds = (lds_t *)lds_mem;
    const int g = a.G.g;
    const int xe = g + a.G.NX, ye = g + a.G.NY, ze = g + a.G.NZ;
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if ((int)blockIdx.x >= ntile) {
        // Wall column x = g+NX: U faces only (V and W faces need x < g+NX).
        const int k = ((int)blockIdx.x - ntile) * (BX * BY) + threadIdx.y * BX + threadIdx.x;
        const int y = g + k % a.G.NY, z = g + k / a.G.NY;
        if (k < a.G.NY * a.G.NZ) {
            const body::GlobalLoad ld{a};
            const double v = body::advect_value<AX, 0>(a, ld, xe, y, z, bad);
            a.T0.d[body::idx(a.T0, xe, y, z)] = v;
        }
        if (bad) *a.flag = 1;
        return;
    }
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int xs = g + bx * BX, ys = g + by * BY, zs = g + bz * ZC;
    const int zend = zs + ZC < ze + 1 ? zs + ZC : ze + 1;
    const int ox = xs - B::XLO, oy = ys - B::YLO;
    const int x = xs + threadIdx.x, y = ys + threadIdx.y;
    const bool xin = x < xe, yin = y < ye, yok = y <= ye;
    Stager<B, BX * BY> st;
    st.init(a, ox, oy);
    for (int q = 0; q < B::R - 1; q++) {
        st.fetch(zs - B::ZLO + q);
        st.store(lds, zs - B::ZLO + q);
    }
    st.fetch(zs - B::ZLO + B::R - 1);
    for (int z = zs; z < zend; z++) {
        st.store(lds, z - B::ZLO + B::R - 1);
        __syncthreads();
        if (z + 1 < zend) st.fetch(z + 1 - B::ZLO + B::R - 1);
        const LdsLoad<B> ld{a, lds, ox, oy, z - B::ZLO};
        const bool zin = z < ze;
        double vu = 0, vv = 0, vw = 0;
        const bool du = xin && yin && zin, dv = xin && yok && zin, dw = xin && yin;
        if (du) vu = body::advect_value<AX, 0>(a, ld, x, y, z, bad);
        if (dv) vv = body::advect_value<AX, 1>(a, ld, x, y, z, bad);
        if (dw) vw = body::advect_value<AX, 2>(a, ld, x, y, z, bad);
        if (du) a.T0.d[body::idx(a.T0, x, y, z)] = vu;
        if (dv) a.T1.d[body::idx(a.T1, x, y, z)] = vv;
        if (dw) a.T2.d[body::idx(a.T2, x, y, z)] = vw;
        __syncthreads();
    }
    if (bad) *a.flag = 1;
}

template <int BY, int D, int ZC>
hipError_t launch_lds(int axis, const SweepArgs &a, hipStream_t s)
{
    const int3 nb = make_int3((a.G.NX + BX - 1) / BX, (a.G.NY + 1 + BY - 1) / BY,
                              (a.G.NZ + 1 + ZC - 1) / ZC);
    const int ncol = (a.G.NY * a.G.NZ + BX * BY - 1) / (BX * BY);
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
