// af_sweep_stage.h -- the advection sweep's LDS tile with per-field staged
// boxes (stencil.go:44-119), shared by the single-engine kernel
// (af_advect_lds.hip) and the batched kernel (af_batch.hip).
//
// A workgroup of 64 x BY threads owns a 64 x BY tile of (x, y) faces and
// marches through a chunk of z planes, as the tile of af_sweep_lds.h does.
// What differs is how much it stages and how.
//
// * Each source field is staged over the box that its own reads need, not
//   over the union of all three.  Only the field that carries the sweep
//   axis's velocity component (U for the X sweep, V for Y, W for Z) is read
//   at the neighbouring rows and planes of faceVelocity's four-point
//   averages (stencil.go:128-155); the other two are read only at their
//   trilinear taps (base and base+1 on every axis, stencil.go:169-225) and
//   at the diffusion stencil on the sweep axis (stencil.go:97-111).  With
//   the tile at (xs, ys) and the current plane z (reach R along the sweep
//   axis: R = 2 for X and Y, departures within (-1, 1] for Z):
//     X sweep: all fields x in [xs-2, xs+66]; U y in [ys-1, ys+BY],
//              z in [z-1, z+1]; V, W y in [ys, ys+BY], z in [z, z+1]
//     Y sweep: V x in [xs-1, xs+64], U, W x in [xs, xs+64]; all
//              y in [ys-2, ys+BY+2]; V z in [z-1, z+1], U, W z in [z, z+1]
//     Z sweep: W x in [xs-1, xs+64], y in [ys-1, ys+BY]; U, V
//              x in [xs, xs+64], y in [ys, ys+BY]; all z in [z-1, z+1]
//   The boxes take 36.9 / 48.0 / 44.4 KB for BY = 8 (the union boxes took
//   49.7 / 61.8 / 63.4 KB), so three 8-wave workgroups fit a CU's LDS.
// * Rows are staged as 16-byte pairs of x-adjacent values: one global load
//   and one LDS write per pair.  A pair that straddles the field's first or
//   last column is loaded from the nearest in-field pair and its two slots
//   are filled as the per-element clamp would fill them.
// * Only the sweep axis is bounds-checked at the taps: on the other two the
//   tap base is the face's own index, inside every box by construction.
//   A face whose tap base lies beyond the staged reach reads its eight taps
//   from global memory.  Which copy a value comes from never changes it.
//
// Every face is formed with body::advect_value's operations in its order
// (contraction off), so the results are bit-identical to the unstaged
// sweep.
#pragma once

#include "af_bodies.h"
#include "af_sweep_lds.h"

#pragma clang fp contract(off)

namespace af {
namespace stg {

constexpr int BX = body::BX;
using slds::lds_t;

// Staged box of source field F (0 = U, 1 = V, 2 = W) for sweep axis AX:
// columns [xs-XLO, xs+XHI], rows [ys-YLO, ys+YHI], planes [z-ZLO, z+1],
// held in a ring of R planes of XW x YW doubles (XW even).
template <int AX, int BY, int F>
struct FBox {
    static constexpr int XLO = AX == 0 ? 2 : (F == AX ? 1 : 0);
    static constexpr int XHI = AX == 0 ? BX + 2 : BX;
    static constexpr int YLO = AX == 1 ? 2 : (F == AX ? 1 : 0);
    static constexpr int YHI = AX == 1 ? BY + 2 : BY;
    static constexpr int ZLO = AX == 2 ? 1 : (F == AX ? 1 : 0);
    static constexpr int R = ZLO + 2;
    static constexpr int XW = (XHI + XLO + 2) & ~1;
    static constexpr int YW = YHI + YLO + 1;
    static constexpr int PAIRS = XW / 2 * YW;
    static constexpr int PLANE = XW * YW;
    static constexpr int SIZE = PLANE * R;
};

template <int AX, int BY>
struct Layout {
    using B0 = FBox<AX, BY, 0>;
    using B1 = FBox<AX, BY, 1>;
    using B2 = FBox<AX, BY, 2>;
    static constexpr int OFF1 = B0::SIZE;
    static constexpr int OFF2 = B0::SIZE + B1::SIZE;
    static constexpr int TOTAL = B0::SIZE + B1::SIZE + B2::SIZE;  // doubles
};

// Staging of one field's box, split over the workgroup's NT threads: thread
// t owns pairs t, t+NT, ... of a plane (lanes past the box repeat a pair of
// it: same load, same value to the same slots).  The field pointer and
// plane stride are workgroup-uniform; per pair a thread keeps the offset of
// the loaded pair in a plane, its LDS offset and the two edge flags.
template <class B, int NT>
struct FStage {
    static constexpr int NP = (B::PAIRS + NT - 1) / NT;
    static_assert(NP * NT < 2 * B::PAIRS, "a repeated pair must lie in the box");
    const __attribute__((address_space(1))) double *base;
    int sz, zmax;
    int goff[NP];
    int loff[NP];
    bool lo[NP], hi[NP];  // pair starts past the last column / before the first
    double2 v[NP];

    __device__ __forceinline__ void init(const DField &f, int ox, int oy)
    {
        const int t = threadIdx.y * blockDim.x + threadIdx.x;
        base = body::gp((const double *)f.d);
        sz = f.sz;
        zmax = f.nz - 1;
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int e0 = t + j * NT;
            const int e = e0 < B::PAIRS ? e0 : e0 - B::PAIRS;
            const int r = e / (B::XW / 2);
            const int x = ox + 2 * (e - r * (B::XW / 2));
            const int y = oy + r;
            // The pair (x, x+1) is read from (q, q+1), q the nearest in-field
            // pair start; elements outside the field take the clamped
            // column's value, as a per-element clamp would give them.
            const int q = x < 0 ? 0 : (x > f.nx - 2 ? f.nx - 2 : x);
            const int yc = y < 0 ? 0 : (y >= f.ny ? f.ny - 1 : y);
            goff[j] = q + yc * f.sy;
            loff[j] = 2 * e;
            lo[j] = x > q;
            hi[j] = x < q;
        }
    }
    __device__ __forceinline__ void fetch(int z)
    {
        const int zc = z < 0 ? 0 : (z > zmax ? zmax : z);
        const auto p = base + (size_t)zc * sz;
#pragma unroll
        for (int j = 0; j < NP; j++) v[j] = body::ld2((const double *)(p + goff[j]));
    }
    __device__ __forceinline__ void store(lds_t *slot) const
    {
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const double a = lo[j] ? v[j].y : v[j].x;
            const double b = hi[j] ? v[j].x : v[j].y;
            typedef double dbl2v __attribute__((ext_vector_type(2)));
            *(__attribute__((address_space(3))) dbl2v *)(slot + loff[j]) = dbl2v{a, b};
        }
    }
};

// Ring slot offsets (doubles from the LDS base) of planes z-1, z and z+1
// of each field, for the current plane z.
struct Planes {
    int m[3], c[3], p[3];
};

template <int AX, int BY>
__device__ __forceinline__ Planes planes(int z)
{
    using L = Layout<AX, BY>;
    using B0 = typename L::B0;
    using B1 = typename L::B1;
    using B2 = typename L::B2;
    const int k3 = z % 3, k2 = z & 1;
    const int m3 = k3 == 0 ? 2 : k3 - 1, p3 = k3 == 2 ? 0 : k3 + 1;
    Planes P;
    // R = 3 rings hold planes z-1, z, z+1 in slots m3, k3, p3; R = 2 rings
    // hold z, z+1 in slots k2, 1-k2 (plane z-1 is not staged there).
    P.m[0] = B0::R == 3 ? m3 * B0::PLANE : 0;
    P.c[0] = (B0::R == 3 ? k3 : k2) * B0::PLANE;
    P.p[0] = (B0::R == 3 ? p3 : 1 - k2) * B0::PLANE;
    P.m[1] = L::OFF1 + (B1::R == 3 ? m3 * B1::PLANE : 0);
    P.c[1] = L::OFF1 + (B1::R == 3 ? k3 : k2) * B1::PLANE;
    P.p[1] = L::OFF1 + (B1::R == 3 ? p3 : 1 - k2) * B1::PLANE;
    P.m[2] = L::OFF2 + (B2::R == 3 ? m3 * B2::PLANE : 0);
    P.c[2] = L::OFF2 + (B2::R == 3 ? k3 : k2) * B2::PLANE;
    P.p[2] = L::OFF2 + (B2::R == 3 ? p3 : 1 - k2) * B2::PLANE;
    return P;
}

// Offset of thread (tx, ty)'s own face within a plane of field F's box.
template <int AX, int BY, int F>
__device__ __forceinline__ int own(int tx, int ty)
{
    using B = FBox<AX, BY, F>;
    return (ty + B::YLO) * B::XW + tx + B::XLO;
}

// The velocity component along the sweep axis at a face of component C:
// faceVelocity's value (stencil.go:128-155), same four-point order.
template <int AX, int BY, int C>
__device__ __forceinline__ double velocity(const lds_t *L, const Planes &P, int r)
{
    constexpr int XW = FBox<AX, BY, AX>::XW;
    const lds_t *c = L + P.c[AX] + r;
    if (C == AX) return c[0];
    if (AX == 0) {
        if (C == 1) return 0.25 * (c[-XW] + c[1 - XW] + c[0] + c[1]);
        const lds_t *m = L + P.m[AX] + r;
        return 0.25 * (m[0] + m[1] + c[0] + c[1]);
    }
    if (AX == 1) {
        if (C == 0) return 0.25 * (c[-1] + c[0] + c[XW - 1] + c[XW]);
        const lds_t *m = L + P.m[AX] + r;
        return 0.25 * (m[0] + m[XW] + c[0] + c[XW]);
    }
    const lds_t *p = L + P.p[AX] + r;
    if (C == 0) return 0.25 * (c[-1] + c[0] + p[-1] + p[0]);
    return 0.25 * (c[-XW] + c[0] + p[-XW] + p[0]);
}

// advectComponent's value at face (x, y, z) of component C from the staged
// boxes: body::departure, body::trilinear and body::diffusion's operations
// in their order.  `coord` is the face's coordinate on the sweep axis.
// `ok` marks a face that is stored: only then may its taps come from global
// memory (a staged reach exceeded) and its NaN departure raise `bad`.
template <int AX, int BY, int C>
__device__ __forceinline__ double face(const SweepArgs &a, const DField &s, const lds_t *L,
                                       const Planes &P, int rc, int rv, int tx, int ty, int xs,
                                       int ys, int x, int y, int z, double coord, bool ok,
                                       bool &bad)
{
    using B = FBox<AX, BY, C>;
    // Departure point (body::departure).
    const double v = velocity<AX, BY, C>(L, P, rv);
    const int ext = AX == 0 ? s.nx : (AX == 1 ? s.ny : s.nz);
    double p = coord;
    p -= a.dt * v * a.invH;
    const bool nan = p != p;
    // The reference's clamp (p < 0 -> 0, p > max -> max) for every p that
    // is not NaN: p is never -0 (an integer minus a double is +0 when they
    // are equal), and a NaN keeps its value (and base 0) as there.
    const double pc = fmin(fmax(p, 0.0), (double)(ext - 1));
    int i0 = (int)pc;
    i0 = i0 < ext - 2 ? i0 : ext - 2;
    const double t = (nan ? p : pc) - (double)i0;
    body::Departure d;
    d.x0 = AX == 0 ? i0 : x;
    d.y0 = AX == 1 ? i0 : y;
    d.z0 = AX == 2 ? i0 : z;
    d.tx = AX == 0 ? t : 0.0;
    d.ty = AX == 1 ? t : 0.0;
    d.tz = AX == 2 ? t : 0.0;
    // Taps: from the staged box when the base lies in its reach.
    int o0, o1;
    bool in;
    if (AX == 0) {
        const int col = i0 - (xs - B::XLO);
        in = (unsigned)col < (unsigned)(B::XW - 1);
        o0 = P.c[C] + (ty + B::YLO) * B::XW + (in ? col : 0);
        o1 = o0 + (P.p[C] - P.c[C]);
    } else if (AX == 1) {
        const int row = i0 - (ys - B::YLO);
        in = (unsigned)row < (unsigned)(B::YW - 1);
        o0 = P.c[C] + (in ? row : 0) * B::XW + tx + B::XLO;
        o1 = o0 + (P.p[C] - P.c[C]);
    } else {
        const int dz = i0 - (z - 1);
        in = (unsigned)dz < 2u;
        o0 = (dz == 0 ? P.m[C] : P.c[C]) + rc;
        o1 = (dz == 0 ? P.c[C] : P.p[C]) + rc;
    }
    const double2 p00 = make_double2(L[o0], L[o0 + 1]);
    const double2 p10 = make_double2(L[o0 + B::XW], L[o0 + B::XW + 1]);
    const double2 p01 = make_double2(L[o1], L[o1 + 1]);
    const double2 p11 = make_double2(L[o1 + B::XW], L[o1 + B::XW + 1]);
    double adv = body::trilinear(p00, p10, p01, p11, d);
    if (ok && !in) adv = slds::far_taps(s, d);
    bad = bad || (ok && nan);
    // Diffusion (body::diffusion); `var lap float64` stays +0 when lam == 0.
    double lap = 0.0;
    if (a.lam != 0) {
        const lds_t *c = L + P.c[C] + rc;
        double dm, dp;
        if (AX == 0) {
            dm = c[-1];
            dp = c[1];
        } else if (AX == 1) {
            dm = c[-B::XW];
            dp = c[B::XW];
        } else {
            dm = L[P.m[C] + rc];
            dp = L[P.p[C] + rc];
        }
        lap = a.lam * (dm - 2 * c[0] + dp);
    }
    return adv + lap;
}

// One tile (bx, by) over face planes [zs, zend) (src_of(c) gives source
// field c's descriptor): the faces of x in [g+64bx, g+64bx+64) (x < g+NX),
// y = g+BY*by+ty (y <= g+NY).  Ends with a barrier, so the caller may
// reuse `lds` at once.  With BC the stores also apply the velocity BC
// (body::store_face_bc).
template <int AX, int BY, bool BC, class Src>
__device__ __forceinline__ void tile(const SweepArgs &a, const Src &src_of, lds_t *lds, int bx,
                                     int by, int zs, int zend, bool &bad)
{
    using L = Layout<AX, BY>;
    using B0 = typename L::B0;
    using B1 = typename L::B1;
    using B2 = typename L::B2;
    constexpr int NT = BX * BY;
    const int g = a.G.g;
    const int xe = g + a.G.NX, ye = g + a.G.NY, ze = g + a.G.NZ;
    const int xs = g + bx * BX, ys = g + by * BY;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = xs + tx, y = ys + ty;
    const bool xin = x < xe, yin = y < ye, yok = y <= ye;
    const DField &SU = src_of(0), &SV = src_of(1), &SW = src_of(2);
    FStage<B0, NT> s0;
    FStage<B1, NT> s1;
    FStage<B2, NT> s2;
    s0.init(SU, xs - B0::XLO, ys - B0::YLO);
    s1.init(SV, xs - B1::XLO, ys - B1::YLO);
    s2.init(SW, xs - B2::XLO, ys - B2::YLO);
    // Prologue: planes [zs - ZLO, zs] of each field, then plane zs+1 into
    // registers.
    {
        const Planes P = planes<AX, BY>(zs);
        if (B0::ZLO) { s0.fetch(zs - 1); s0.store(lds + P.m[0]); }
        if (B1::ZLO) { s1.fetch(zs - 1); s1.store(lds + P.m[1]); }
        if (B2::ZLO) { s2.fetch(zs - 1); s2.store(lds + P.m[2]); }
        s0.fetch(zs); s1.fetch(zs); s2.fetch(zs);
        s0.store(lds + P.c[0]); s1.store(lds + P.c[1]); s2.store(lds + P.c[2]);
        s0.fetch(zs + 1); s1.fetch(zs + 1); s2.fetch(zs + 1);
    }
    const int r0 = own<AX, BY, 0>(tx, ty), r1 = own<AX, BY, 1>(tx, ty),
              r2 = own<AX, BY, 2>(tx, ty);
    const int rv = AX == 0 ? r0 : (AX == 1 ? r1 : r2);
    const double cx = (double)x, cy = (double)y;
    // Each face is stored as soon as it is formed, after the next plane's
    // prefetch was issued.  (Vector memory operations complete in issue
    // order, so a store issued before the prefetch would have to complete
    // before the next iteration's wait for the prefetch returns.)
    auto put = [&](int c, const DField &T, int z, double v) {
        if (BC) {
            if (c == 0) body::store_face_bc<0>(a.G, T, x, y, z, v);
            if (c == 1) body::store_face_bc<1>(a.G, T, x, y, z, v);
            if (c == 2) body::store_face_bc<2>(a.G, T, x, y, z, v);
        } else {
            body::gp(T.d)[body::idx(T, x, y, z)] = v;
        }
    };
    for (int z = zs; z < zend; z++) {
        const Planes P = planes<AX, BY>(z);
        s0.store(lds + P.p[0]);
        s1.store(lds + P.p[1]);
        s2.store(lds + P.p[2]);
        __syncthreads();
        if (z + 1 < zend) {
            s0.fetch(z + 2);
            s1.fetch(z + 2);
            s2.fetch(z + 2);
        }
        const bool zin = z < ze;
        const bool du = xin && yin && zin, dv = xin && yok && zin, dw = xin && yin;
        const double coord = AX == 0 ? cx : (AX == 1 ? cy : (double)z);
        const double vu =
            face<AX, BY, 0>(a, SU, lds, P, r0, rv, tx, ty, xs, ys, x, y, z, coord, du, bad);
        if (du) put(0, a.T0, z, vu);
        const double vv =
            face<AX, BY, 1>(a, SV, lds, P, r1, rv, tx, ty, xs, ys, x, y, z, coord, dv, bad);
        if (dv) put(1, a.T1, z, vv);
        const double vw =
            face<AX, BY, 2>(a, SW, lds, P, r2, rv, tx, ty, xs, ys, x, y, z, coord, dw, bad);
        if (dw) put(2, a.T2, z, vw);
        __syncthreads();
    }
}

}  // namespace stg
}  // namespace af
