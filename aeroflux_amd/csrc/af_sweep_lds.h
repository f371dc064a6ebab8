// af_sweep_lds.h -- the LDS-staged advection sweep's tile, shared by the
// single-engine kernel (af_advect_lds.hip) and the batched persistent kernel
// (af_batch.hip).  The design is described in af_advect_lds.hip.
#pragma once

#include "af_bodies.h"

#pragma clang fp contract(off)

namespace af {
namespace slds {

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
// diffusion stencil of any position of the tile lie in the staged box (the
// box is built from them), so they are read from LDS unchecked; the
// trilinear taps are read from LDS when their 2x2x2 cube lies in the box and
// inside the field, else from global memory (by the caller).
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
    // The eight trilinear taps at base (x, y, z) from LDS when their 2x2x2
    // cube lies in the box and inside the field (returns true); otherwise
    // returns false and reads an arbitrary staged slot, and the caller takes
    // the taps from global memory.  No branch: the three faces of a plane
    // issue their LDS reads together.
    __device__ __forceinline__ bool taps_staged(int c, int x, int y, int z, double2 &p00,
                                                double2 &p10, double2 &p01, double2 &p11) const
    {
        const DField &s = f(c);
        // Base and base+1 inside the box on every axis, and inside the field
        // (the box is clipped to the field when staged).
        const bool in = ((unsigned)(x - ox) < (unsigned)(B::XW - 1)) &
                        ((unsigned)(y - oy) < (unsigned)(B::YW - 1)) &
                        ((unsigned)(z - zlo) < (unsigned)(B::R - 1)) & (x >= 0) &
                        (x + 1 < s.nx) & (y >= 0) & (y + 1 < s.ny) & (z >= 0) & (z + 1 < s.nz);
        const int k = in ? slot(c, x, y, z) : c * B::FIELD;
        const int k1 = in ? slot(c, x, y, z + 1) : c * B::FIELD;
        p00 = make_double2(lds[k], lds[k + 1]);
        p10 = make_double2(lds[k + B::XW], lds[k + B::XW + 1]);
        p01 = make_double2(lds[k1], lds[k1 + 1]);
        p11 = make_double2(lds[k1 + B::XW], lds[k1 + B::XW + 1]);
        return in;
    }
};

// Staging of one plane of the three source fields, split over the
// workgroup: thread t owns elements t, t+NT, ... of one plane's (row,
// column) box, in all three fields.  Field pointers, plane strides and
// extents are workgroup-uniform (scalar registers); per element a thread
// keeps only its offset in each field and its LDS offset.  Per plane a
// thread issues all its global loads into registers (`fetch`) and later
// writes them to the plane's ring slot (`store`), so the next plane's loads
// are in flight while the current plane is computed.
template <class B, int NT>
struct Stager {
    static constexpr int NF = (B::PLANE + NT - 1) / NT;  // elements per field per thread
    static_assert(NT <= B::PLANE, "a lane past the box repeats an element of it");
    const __attribute__((address_space(1))) double *base[3];
    int sz[3], zmax[3];  // plane stride and last plane of each field
    int eoff[NF][3];     // element's offset in each field, plane 0
    int loff[NF];        // element's LDS offset within a field's ring slot
    double v[NF][3];

    // No lane branches: lanes past the box repeat an element of it (same
    // load, same value to the same slot), and elements outside a field
    // (columns, rows or planes beyond the padded field) load a clamped
    // in-field value into a slot no face reads -- LdsLoad reads only cells
    // of the field.
    //
    // src(c) gives source field c's descriptor (c is a compile-time
    // constant here).
    template <class Src>
    __device__ __forceinline__ void init(const Src &src_of, int ox, int oy)
    {
        const int t = threadIdx.y * blockDim.x + threadIdx.x;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const DField &f = src_of(c);
            base[c] = body::gp((const double *)f.d);
            sz[c] = f.sz;
            zmax[c] = f.nz - 1;
        }
#pragma unroll
        for (int j = 0; j < NF; j++) {
            const int e0 = t + j * NT;
            const int e = e0 < B::PLANE ? e0 : e0 - B::PLANE;
            const int r = e / B::XW;
            const int col = e - r * B::XW;
            const int x = ox + col, y = oy + r;
            loff[j] = e;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const DField &f = src_of(c);
                const int xc = x < 0 ? 0 : (x >= f.nx ? f.nx - 1 : x);
                const int yc = y < 0 ? 0 : (y >= f.ny ? f.ny - 1 : y);
                eoff[j][c] = xc + yc * f.sy;
            }
        }
    }
    __device__ __forceinline__ void fetch(int z)
    {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int zc = z < 0 ? 0 : (z > zmax[c] ? zmax[c] : z);
            const auto p = base[c] + (size_t)zc * sz[c];
#pragma unroll
            for (int j = 0; j < NF; j++) v[j][c] = p[eoff[j][c]];
        }
    }
    __device__ __forceinline__ void store(lds_t *lds, int z) const
    {
        if (z < 0) return;
        const int slot = (z % B::R) * B::PLANE;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < NF; j++) lds[c * B::FIELD + slot + loff[j]] = v[j][c];
    }
};

// Tile grid of one sweep: x-tiles of 64 over [g, g+NX), y-tiles of BY
// rows over [g, g+NY], z-chunks of at most ZC planes over [g, g+NZ]; and
// the number of BX*BY-thread blocks that cover the wall column's NY*NZ U
// faces.
template <int BY, int ZC>
__host__ __device__ inline int3 tiles(const Geom &G)
{
    return make_int3((G.NX + BX - 1) / BX, (G.NY + 1 + BY - 1) / BY, (G.NZ + 1 + ZC - 1) / ZC);
}

// Planes [zs, zend) of z-chunk bz when the NZ+1 face planes [g, g+NZ] are
// split into nzc chunks whose sizes differ by at most one.
__device__ __forceinline__ void zchunk(const Geom &G, int nzc, int bz, int &zs, int &zend)
{
    const int n = G.NZ + 1;
    zs = G.g + (int)(((long long)bz * n) / nzc);
    zend = G.g + (int)(((long long)(bz + 1) * n) / nzc);
}
template <int BY>
__host__ __device__ inline int wall_blocks(const Geom &G)
{
    return (G.NY * G.NZ + BX * BY - 1) / (BX * BY);
}

// Wall column x = g+NX: U faces only (V and W faces need x < g+NX), from
// global memory; face k of NY*NZ.
template <int AX, bool BC>
__device__ __forceinline__ void wall_face(const SweepArgs &a, int k, bool &bad)
{
    const int g = a.G.g;
    if (k >= a.G.NY * a.G.NZ) return;
    const int xe = g + a.G.NX, y = g + k % a.G.NY, z = g + k / a.G.NY;
    const body::GlobalLoad ld{a};
    const double v = body::advect_value<AX, 0>(a, ld, xe, y, z, bad);
    if (BC)
        body::store_face_bc<0>(a.G, a.T0, xe, y, z, v);
    else
        body::gp(a.T0.d)[body::idx(a.T0, xe, y, z)] = v;
}

// A face of the current plane formed from LDS (body::advect_value's parts,
// same arithmetic): departure, taps if staged, and diffusion term.
struct Face {
    body::Departure d;
    double adv, lap;
    bool far, nan;  // taps beyond the staged box; NaN departure coordinate
};

template <int AX, int C, class B>
__device__ __forceinline__ Face face(const SweepArgs &a, const LdsLoad<B> &ld, int x, int y, int z)
{
    Face f;
    f.d = body::departure<AX, C>(a, ld, x, y, z, f.nan);
    double2 p00, p10, p01, p11;
    f.far = !ld.taps_staged(C, f.d.x0, f.d.y0, f.d.z0, p00, p10, p01, p11);
    f.adv = body::trilinear(p00, p10, p01, p11, f.d);
    f.lap = body::diffusion<AX, C>(a, ld, x, y, z);
    return f;
}

// The trilinear value from global memory, for a face whose taps lie beyond
// the staged box.
__device__ __forceinline__ double far_taps(const DField &s, const body::Departure &d)
{
    const double *p = s.d + body::idx(s, d.x0, d.y0, d.z0);
    return body::trilinear(body::ld2(p), body::ld2(p + s.sy), body::ld2(p + s.sz),
                           body::ld2(p + s.sz + s.sy), d);
}

// One tile (bx, by) over face planes [zs, zend), computed by the whole
// workgroup (64 x BY threads) (src_of as for Stager::init): the faces of x
// in [g+64bx, g+64bx+64) (x < g+NX), y = g+BY*by+ty (y <= g+NY).  Ends with
// a barrier, so the caller may reuse `lds` at once.
template <int AX, int BY, int DD, bool BC, class Src>
__device__ __forceinline__ void tile(const SweepArgs &a, const Src &src_of, lds_t *lds, int bx,
                                     int by, int zs, int zend, bool &bad)
{
    using B = Box<AX, BY, DD>;
    const int g = a.G.g;
    const int xe = g + a.G.NX, ye = g + a.G.NY, ze = g + a.G.NZ;
    const int xs = g + bx * BX, ys = g + by * BY;
    const int ox = xs - B::XLO, oy = ys - B::YLO;
    const int x = xs + threadIdx.x, y = ys + threadIdx.y;
    const bool xin = x < xe, yin = y < ye, yok = y <= ye;
    Stager<B, BX * BY> st;
    st.init(src_of, ox, oy);
    for (int q = 0; q < B::R - 1; q++) {
        st.fetch(zs - B::ZLO + q);
        st.store(lds, zs - B::ZLO + q);
    }
    st.fetch(zs - B::ZLO + B::R - 1);
    // A plane's results are stored at the start of the next iteration, just
    // before that iteration's prefetch, so the prefetch is the last vector
    // memory operation in flight when the following iteration waits for it:
    // that wait then never includes the write-back of the faces.
    double vu = 0, vv = 0, vw = 0;
    bool du = false, dv = false, dw = false;
    int zp = zs;
    auto put = [&]() {
        if (BC) {
            if (du) body::store_face_bc<0>(a.G, a.T0, x, y, zp, vu);
            if (dv) body::store_face_bc<1>(a.G, a.T1, x, y, zp, vv);
            if (dw) body::store_face_bc<2>(a.G, a.T2, x, y, zp, vw);
        } else {
            if (du) body::gp(a.T0.d)[body::idx(a.T0, x, y, zp)] = vu;
            if (dv) body::gp(a.T1.d)[body::idx(a.T1, x, y, zp)] = vv;
            if (dw) body::gp(a.T2.d)[body::idx(a.T2, x, y, zp)] = vw;
        }
    };
    for (int z = zs; z < zend; z++) {
        st.store(lds, z - B::ZLO + B::R - 1);
        __syncthreads();
        put();
        if (z + 1 < zend) st.fetch(z + 1 - B::ZLO + B::R - 1);
        const LdsLoad<B> ld{a, lds, ox, oy, z - B::ZLO};
        const bool zin = z < ze;
        du = xin && yin && zin;
        dv = xin && yok && zin;
        dw = xin && yin;
        zp = z;
        // Every lane forms each face from LDS without a branch (positions
        // outside the grid read staged slots and are not stored); a face
        // whose taps lie beyond the staged box then takes them from global
        // memory.  Faces go one after the other: all three at once would
        // hold three faces' taps in registers and halve the waves per SIMD.
        Face f = face<AX, 0>(a, ld, x, y, z);
        if (du && f.far) f.adv = far_taps(a.U, f.d);
        bad = bad || (du && f.nan);
        vu = f.adv + f.lap;
        f = face<AX, 1>(a, ld, x, y, z);
        if (dv && f.far) f.adv = far_taps(a.V, f.d);
        bad = bad || (dv && f.nan);
        vv = f.adv + f.lap;
        f = face<AX, 2>(a, ld, x, y, z);
        if (dw && f.far) f.adv = far_taps(a.W, f.d);
        bad = bad || (dw && f.nan);
        vw = f.adv + f.lap;
        __syncthreads();
    }
    put();
}

}  // namespace slds
}  // namespace af
