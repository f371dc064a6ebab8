// af_bodies.h -- per-thread bodies of the step's stencil kernels.
//
// Each body computes what one thread of one workgroup tile computes, given
// the tile's coordinates.  The single-engine kernels (af_kernels.hip,
// af_advect.hip, af_rbgs_split.hip) and the batched persistent kernels
// (af_batch.hip) both call them, so a batched engine runs exactly the
// arithmetic of a single one.  Every body reproduces the reference's IEEE
// binary64 operation order; contraction is off for every file that includes
// this header.
#pragma once

#include "af_internal.h"
#include "af_rbgs_split.h"

#pragma clang fp contract(off)

namespace af {
namespace body {

constexpr int BX = 64;  // lanes along x

// Field memory as global memory.  A descriptor read from device memory (the
// batched slots' table) carries generic pointers, and loads and stores
// through generic pointers become flat instructions, which count against
// the LDS wait counter as well: a kernel that stages through LDS would then
// wait for its in-flight global prefetch at every LDS read.  Every field
// access in these bodies goes through gp(), so it is a global instruction
// whichever way the descriptor arrived.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gp(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ int idx(const DField &f, int x, int y, int z)
{
    return x + y * f.sy + z * f.sz;  // Field.Index, grid.go:24-26
}

// ------------------------------------------------------------ velocity BC
//
// applyVelocityBC (stencil.go:246-315) as one gather.  For a component whose
// normal axis is n, the reference writes exactly the boundary cells of the
// box  c_n in [g, g+N_n],  c_a in [g-1, g+N_a] (a != n):
//   * wall faces c_n in {g, g+N_n} with interior tangential coordinates get 0;
//   * every other written cell gets -(value of the cell one step inward along
//     one ghost axis), applied axis by axis.
// Unrolling that chain, the final value of any written cell is the value of
// its "base" cell (ghost coordinates clamped inward) negated once per ghost
// coordinate, where the base is +0 if it lies on a wall face and the
// untouched interior value otherwise.  Negation is exact, so this equals the
// serial loops bit for bit, signed zeros included (one reflection of a wall
// zero gives -0, two give +0).  The base cells are never written by this
// kernel, so all writes are independent.

// Boundary of an Lx*Ly*Lz box, enumerated once per cell:
//   family Z: z in {0, Lz-1}, all x, y          (2*Lx*Ly cells)
//   family Y: y in {0, Ly-1}, 0 < z < Lz-1      (2*Lx*(Lz-2))
//   family X: x in {0, Lx-1}, 0 < y,z < L-1     (2*(Ly-2)*(Lz-2))
__host__ __device__ inline int64_t shell_count(int Lx, int Ly, int Lz)
{
    return 2LL * Lx * Ly + 2LL * Lx * (Lz - 2) + 2LL * (Ly - 2) * (Lz - 2);
}

__device__ __forceinline__ void shell_decode(int64_t k64, int Lx, int Ly, int Lz, int &x, int &y,
                                             int &z)
{
    // Shell sizes are far below 2^31 (fields are indexed with 32-bit
    // integers), so 32-bit unsigned division suffices; 64-bit division is a
    // long instruction sequence that dominated this kernel.
    unsigned k = (unsigned)k64;
    const unsigned ux = (unsigned)Lx, uy = (unsigned)Ly;
    const unsigned nz = 2u * ux * uy;
    if (k < nz) {
        const unsigned r = k / ux;
        x = (int)(k - r * ux);
        const unsigned q = r / uy;
        y = (int)(r - q * uy);
        z = q == 0 ? 0 : Lz - 1;
        return;
    }
    k -= nz;
    const unsigned uz2 = (unsigned)(Lz - 2);
    const unsigned ny = 2u * ux * uz2;
    if (k < ny) {
        const unsigned r = k / ux;
        x = (int)(k - r * ux);
        const unsigned q = r / uz2;
        z = 1 + (int)(r - q * uz2);
        y = q == 0 ? 0 : Ly - 1;
        return;
    }
    k -= ny;
    const unsigned uy2 = (unsigned)(Ly - 2);
    const unsigned r = k / uy2;
    y = 1 + (int)(k - r * uy2);
    const unsigned q = r / uz2;
    z = 1 + (int)(r - q * uz2);
    x = q == 0 ? 0 : Lx - 1;
}

// Velocity-BC shell sizes of U, V, W (counts of cells written).
__host__ __device__ inline void velocity_bc_counts(const Geom &G, int64_t &cu, int64_t &cv,
                                                   int64_t &cw)
{
    cu = shell_count(G.NX + 1, G.NY + 2, G.NZ + 2);
    cv = shell_count(G.NX + 2, G.NY + 1, G.NZ + 2);
    cw = shell_count(G.NX + 2, G.NY + 2, G.NZ + 1);
}

// Writes the final BC value of shell cell k (k < cu+cv+cw).
__device__ __forceinline__ void velocity_bc_cell(const DField &U, const DField &V, const DField &W,
                                                 const Geom &G, int64_t k, int64_t cu, int64_t cv,
                                                 int64_t cw)
{
    const int g = G.g;
    int n;
    DField f;
    int Lx, Ly, Lz, ox, oy, oz;  // box extents and origin
    if (k < cu) {
        n = 0; f = U;
        Lx = G.NX + 1; Ly = G.NY + 2; Lz = G.NZ + 2; ox = g; oy = g - 1; oz = g - 1;
    } else if (k < cu + cv) {
        k -= cu; n = 1; f = V;
        Lx = G.NX + 2; Ly = G.NY + 1; Lz = G.NZ + 2; ox = g - 1; oy = g; oz = g - 1;
    } else if (k < cu + cv + cw) {
        k -= cu + cv; n = 2; f = W;
        Lx = G.NX + 2; Ly = G.NY + 2; Lz = G.NZ + 1; ox = g - 1; oy = g - 1; oz = g;
    } else {
        return;
    }
    int bx, by, bz;
    shell_decode(k, Lx, Ly, Lz, bx, by, bz);
    int c[3] = {ox + bx, oy + by, oz + bz};
    const int x = c[0], y = c[1], z = c[2];
    const int N[3] = {G.NX, G.NY, G.NZ};
    const bool wall = (c[n] == g) || (c[n] == g + N[n]);
    int flips = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a == n) continue;
        if (c[a] == g - 1) {
            c[a] = g;
            flips++;
        } else if (c[a] == g + N[a]) {
            c[a] = g + N[a] - 1;
            flips++;
        }
    }
    double v = wall ? 0.0 : gp(f.d)[idx(f, c[0], c[1], c[2])];
    if (flips & 1) v = -v;  // two exact negations are the identity
    gp(f.d)[idx(f, x, y, z)] = v;
}

// -------------------------------------------------------------- advection
//
// advectComponent (stencil.go:44-119), written so that each face's loads can
// be issued together: no branches inside a face (the departure clamp is a
// select; a NaN coordinate only sets `bad`), and a thread forms every face
// value before storing any.  Two exact specialisations of sampleTrilinear
// (stencil.go:169-225):
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

// Two x-adjacent values p[0], p[1] in one 16-byte load (8-byte aligned; the
// hardware accepts dword-aligned vector loads).  Every trilinear tap pair
// and most four-point averages are x-adjacent, which halves the load
// instructions of a sweep.
typedef double dbl2_a8 __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ double2 ld2(const double *p)
{
    typedef __attribute__((address_space(1))) const dbl2_a8 gdbl2;
    const dbl2_a8 v = *(gdbl2 *)p;
    return make_double2(v.x, v.y);
}

// p[0] = a, p[1] = b in one 16-byte store (8-byte aligned).
__device__ __forceinline__ void st2(double *p, double a, double b)
{
    typedef __attribute__((address_space(1))) dbl2_a8 gdbl2;
    dbl2_a8 v;
    v.x = a;
    v.y = b;
    *(gdbl2 *)p = v;
}

// Where a sweep's values come from.  A loader returns component c's value
// at padded coordinates (x, y, z) (`one`) or the x-adjacent pair at x, x+1
// (`pair`).  GlobalLoad reads the fields in HBM; the LDS-staged sweep
// (af_advect_lds.hip) supplies a loader that serves staged values from LDS.
// Both feed the same arithmetic below, so both give the same bits.
struct GlobalLoad {
    const SweepArgs &a;
    __device__ __forceinline__ const DField &f(int c) const
    {
        return c == 0 ? a.U : (c == 1 ? a.V : a.W);
    }
    __device__ __forceinline__ double one(int c, int x, int y, int z) const
    {
        const DField &s = f(c);
        return gp(s.d)[idx(s, x, y, z)];
    }
    __device__ __forceinline__ double2 pair(int c, int x, int y, int z) const
    {
        const DField &s = f(c);
        return ld2(s.d + idx(s, x, y, z));
    }
    // The eight trilinear taps at base (x, y, z), as four x-adjacent pairs.
    __device__ __forceinline__ void taps(int c, int x, int y, int z, double2 &p00, double2 &p10,
                                         double2 &p01, double2 &p11) const
    {
        const DField &s = f(c);
        const double *d = s.d + idx(s, x, y, z);
        p00 = ld2(d);
        p10 = ld2(d + s.sy);
        p01 = ld2(d + s.sz);
        p11 = ld2(d + s.sz + s.sy);
    }
};

// The velocity component along sweep axis AX at a face of component C, as
// faceVelocity (stencil.go:128-155) forms it (same four-point order).
template <int AX, int C, class L>
__device__ __forceinline__ double axis_velocity(const L &ld, int x, int y, int z)
{
    if (AX == C) return ld.one(C, x, y, z);
    if (AX == 0) {
        const double2 lo = C == 1 ? ld.pair(0, x, y - 1, z) : ld.pair(0, x, y, z - 1);
        const double2 hi = ld.pair(0, x, y, z);
        return 0.25 * (lo.x + lo.y + hi.x + hi.y);
    }
    if (AX == 1) {
        if (C == 0) {
            const double2 lo = ld.pair(1, x - 1, y, z);
            const double2 hi = ld.pair(1, x - 1, y + 1, z);
            return 0.25 * (lo.x + lo.y + hi.x + hi.y);
        }
        return 0.25 * (ld.one(1, x, y, z - 1) + ld.one(1, x, y + 1, z - 1) + ld.one(1, x, y, z) +
                       ld.one(1, x, y + 1, z));
    }
    if (C == 0) {
        const double2 lo = ld.pair(2, x - 1, y, z);
        const double2 hi = ld.pair(2, x - 1, y, z + 1);
        return 0.25 * (lo.x + lo.y + hi.x + hi.y);
    }
    return 0.25 * (ld.one(2, x, y - 1, z) + ld.one(2, x, y, z) + ld.one(2, x, y - 1, z + 1) +
                   ld.one(2, x, y, z + 1));
}

// advectComponent's value at one face (stencil.go:83-115), in three parts
// that the LDS-staged sweep also calls separately: the departure point
// (tap base and weights), the trilinear combination of the eight taps, and
// the diffusion term.  `bad` is set when the departure coordinate is NaN --
// the reference's bounds-panic trigger; the value is then NaN and no load
// leaves the field.
struct Departure {
    int x0, y0, z0;      // tap base
    double tx, ty, tz;   // weights (0 on the axes the backtrace does not move)
};

template <int AX, int C, class L>
__device__ __forceinline__ Departure departure(const SweepArgs &a, const L &ld, int x, int y,
                                               int z, bool &nan)
{
    const DField &s = C == 0 ? a.U : (C == 1 ? a.V : a.W);
    const double v = axis_velocity<AX, C>(ld, x, y, z);
    const int coord = AX == 0 ? x : (AX == 1 ? y : z);
    const int ext = AX == 0 ? s.nx : (AX == 1 ? s.ny : s.nz);
    double p = (double)coord;
    p -= a.dt * v * a.invH;
    const double pmax = (double)(ext - 1);
    p = p < 0 ? 0.0 : (p > pmax ? pmax : p);  // NaN fails both tests
    nan = p != p;
    int i0 = nan ? 0 : (int)p;
    i0 = i0 >= ext - 1 ? ext - 2 : i0;
    const double t = p - (double)i0;
    Departure d;
    d.x0 = AX == 0 ? i0 : x;
    d.y0 = AX == 1 ? i0 : y;
    d.z0 = AX == 2 ? i0 : z;
    d.tx = AX == 0 ? t : 0.0;
    d.ty = AX == 1 ? t : 0.0;
    d.tz = AX == 2 ? t : 0.0;
    return d;
}

// Taps as (d000, d100), (d010, d110), (d001, d101), (d011, d111).
__device__ __forceinline__ double trilinear(const double2 &p00, const double2 &p10,
                                            const double2 &p01, const double2 &p11,
                                            const Departure &d)
{
    const double tx = d.tx, ty = d.ty, tz = d.tz;
    const double c00 = p00.x * (1 - tx) + p00.y * tx;
    const double c10 = p10.x * (1 - tx) + p10.y * tx;
    const double c01 = p01.x * (1 - tx) + p01.y * tx;
    const double c11 = p11.x * (1 - tx) + p11.y * tx;
    const double c0 = c00 * (1 - ty) + c10 * ty;
    const double c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

template <int AX, int C, class L>
__device__ __forceinline__ double diffusion(const SweepArgs &a, const L &ld, int x, int y, int z)
{
    double dm, dc, dp;  // src at i - st, i, i + st
    if (AX == 0) {
        const double2 q = ld.pair(C, x - 1, y, z);
        dm = q.x;
        dc = q.y;
        dp = ld.one(C, x + 1, y, z);
    } else if (AX == 1) {
        dm = ld.one(C, x, y - 1, z);
        dc = ld.one(C, x, y, z);
        dp = ld.one(C, x, y + 1, z);
    } else {
        dm = ld.one(C, x, y, z - 1);
        dc = ld.one(C, x, y, z);
        dp = ld.one(C, x, y, z + 1);
    }
    double lap = a.lam * (dm - 2 * dc + dp);
    return a.lam != 0 ? lap : 0.0;  // `var lap float64` stays +0 when lam == 0
}

template <int AX, int C, class L>
__device__ __forceinline__ double advect_value(const SweepArgs &a, const L &ld, int x, int y, int z,
                                               bool &bad)
{
    bool nan;
    const Departure d = departure<AX, C>(a, ld, x, y, z, nan);
    bad = bad || nan;
    double2 p00, p10, p01, p11;
    ld.taps(C, d.x0, d.y0, d.z0, p00, p10, p01, p11);
    const double adv = trilinear(p00, p10, p01, p11, d);
    return adv + diffusion<AX, C>(a, ld, x, y, z);
}

// Stores the swept value v of component C at face (x, y, z) of T and (with
// a.bc set in store_face) applies applyVelocityBC (stencil.go:246-315) to every
// cell that takes its value from this face (see the velocity BC above): a
// wall face -- normal coordinate g or g+N -- stores +0, and each ghost cell
// whose base is this face stores the stored value negated once per ghost
// coordinate.  A sweep writes every base cell of the BC box exactly once
// (its faces are the box's cells with interior tangential coordinates), so
// with a.bc its output holds what the BC applied after it would leave, and
// that BC launch is skipped.  Without a.bc it is the plain store of
// advectDiffuseAxis.
template <int C>
__device__ __forceinline__ double face_bc_value(const Geom &G, int x, int y, int z, double v)
{
    const int g = G.g;
    const int c[3] = {x, y, z};
    const int N[3] = {G.NX, G.NY, G.NZ};
    return (c[C] == g || c[C] == g + N[C]) ? 0.0 : v;
}

// The ghost cells whose base is face (x, y, z) of component C, given the
// value v the face stores (see store_face_bc).
template <int C>
__device__ __forceinline__ void face_bc_ghosts(const Geom &G, const DField &T, int x, int y, int z,
                                               double v)
{
    const auto t = gp(T.d);
    const int i = idx(T, x, y, z);
    const int g = G.g;
    const int c[3] = {x, y, z};
    const int N[3] = {G.NX, G.NY, G.NZ};
    constexpr int A = C == 0 ? 1 : 0;   // the two tangential axes
    constexpr int Bt = C == 2 ? 1 : 2;
    const int sa = A == 0 ? 1 : T.sy;
    const int sb = Bt == 1 ? T.sy : T.sz;
    const bool alo = c[A] == g, ahi = c[A] == g + N[A] - 1;
    const bool blo = c[Bt] == g, bhi = c[Bt] == g + N[Bt] - 1;
    if (!(alo || ahi || blo || bhi)) return;
    const double nv = -v;  // one ghost coordinate; two give v back (negation is exact)
    if (alo) t[i - sa] = nv;
    if (ahi) t[i + sa] = nv;
    if (blo) t[i - sb] = nv;
    if (bhi) t[i + sb] = nv;
    if (alo && blo) t[i - sa - sb] = v;
    if (alo && bhi) t[i - sa + sb] = v;
    if (ahi && blo) t[i + sa - sb] = v;
    if (ahi && bhi) t[i + sa + sb] = v;
}

template <int C>
__device__ __forceinline__ void store_face_bc(const Geom &G, const DField &T, int x, int y, int z,
                                              double v)
{
    v = face_bc_value<C>(G, x, y, z, v);
    gp(T.d)[idx(T, x, y, z)] = v;
    face_bc_ghosts<C>(G, T, x, y, z, v);
}

template <int C>
__device__ __forceinline__ void store_face(const SweepArgs &a, const DField &T, int x, int y,
                                           int z, double v)
{
    if (a.bc)
        store_face_bc<C>(a.G, T, x, y, z, v);
    else
        gp(T.d)[idx(T, x, y, z)] = v;
}

// One thread of advectDiffuseAxis (stencil.go:21-25): the U, V and W faces
// at (x, y, z0 .. z0+ZR-1) that exist (x <= g+NX, y <= g+NY, z <= g+NZ).
template <int AX, int ZR>
__device__ __forceinline__ void sweep_column(const SweepArgs &a, int x, int y, int z0)
{
    const int g = a.G.g;
    if (x > g + a.G.NX || y > g + a.G.NY) return;
    const bool xin = x < g + a.G.NX, yin = y < g + a.G.NY;
    bool bad = false;
    double vu[ZR], vv[ZR], vw[ZR];
    bool du[ZR], dv[ZR], dw[ZR];
#pragma unroll
    for (int r = 0; r < ZR; r++) {
        const int z = z0 + r;
        const bool zok = z <= g + a.G.NZ, zin = z < g + a.G.NZ;
        du[r] = yin && zin;
        dv[r] = xin && zin;
        dw[r] = xin && yin && zok;
        vu[r] = vv[r] = vw[r] = 0;
        const GlobalLoad ld{a};
        if (du[r]) vu[r] = advect_value<AX, 0>(a, ld, x, y, z, bad);
        if (dv[r]) vv[r] = advect_value<AX, 1>(a, ld, x, y, z, bad);
        if (dw[r]) vw[r] = advect_value<AX, 2>(a, ld, x, y, z, bad);
    }
#pragma unroll
    for (int r = 0; r < ZR; r++) {
        const int z = z0 + r;
        if (du[r]) store_face<0>(a, a.T0, x, y, z, vu[r]);
        if (dv[r]) store_face<1>(a, a.T1, x, y, z, vv[r]);
        if (dw[r]) store_face<2>(a, a.T2, x, y, z, vw[r]);
    }
    if (bad) *gp(a.flag) = 1;
}

// ------------------------------------------------------------- projection

// computeDivergence, project.go:42-62, interior cell (x, y, z).  Returns
// false (writing nothing) outside the interior; else stores the value to D
// (unless store is false) and returns it in `out`.
__device__ __forceinline__ bool divergence_cell(const DField &U, const DField &V, const DField &W,
                                                const DField &D, const Geom &G, double scale,
                                                int x, int y, int z, double &out,
                                                bool store = true)
{
    const int g = G.g;
    if (x >= g + G.NX || y >= g + G.NY) return false;
    const auto u = gp(U.d), v = gp(V.d), w = gp(W.d);
    const double du = u[idx(U, x + 1, y, z)] - u[idx(U, x, y, z)];
    const double dv = v[idx(V, x, y + 1, z)] - v[idx(V, x, y, z)];
    const double dw = w[idx(W, x, y, z + 1)] - w[idx(W, x, y, z)];
    out = scale * (du + dv + dw);
    if (store) gp(D.d)[idx(D, x, y, z)] = out;
    return true;
}

// divergence_cell for the cell pair x = 2k, 2k+1 of interior row (yi, zi)
// (interior coordinates), with the same values: U's faces 2k .. 2k+2 and the
// x-adjacent V and W faces as 16-byte pairs.  Returns false outside the
// interior; the second value is set, not meaningful, when cell 2k+1 is
// outside it (odd NX).  The last face read, U's 2k+2 or a V/W pair's second
// face, is inside the padded row.
__device__ __forceinline__ bool divergence_pair(const DField &U, const DField &V, const DField &W,
                                                const Geom &G, double scale, int k, int yi, int zi,
                                                double &d0, double &d1)
{
    const int g = G.g;
    if (2 * k >= G.NX || yi >= G.NY) return false;
    const int x = g + 2 * k, y = g + yi, z = g + zi;
    const int iu = idx(U, x, y, z), iv = idx(V, x, y, z), iw = idx(W, x, y, z);
    const double2 ua = ld2(U.d + iu), ub = ld2(U.d + iu + 1);
    const double2 v0 = ld2(V.d + iv), v1 = ld2(V.d + iv + V.sy);
    const double2 w0 = ld2(W.d + iw), w1 = ld2(W.d + iw + W.sz);
    const double du0 = ua.y - ua.x, du1 = ub.y - ub.x;
    const double dv0 = v1.x - v0.x, dv1 = v1.y - v0.y;
    const double dw0 = w1.x - w0.x, dw1 = w1.y - w0.y;
    d0 = scale * (du0 + dv0 + dw0);
    d1 = scale * (du1 + dv1 + dw1);
    return true;
}

// Size of P's ghost shell of thickness g.
__host__ __device__ inline int64_t pressure_shell_count(int Lx, int Ly, int Lz, int t)
{
    return 2LL * t * Lx * Ly + 2LL * t * Lx * (Lz - 2 * t) + 2LL * t * (Ly - 2 * t) * (Lz - 2 * t);
}

// applyPressureBC, project.go:139-163, as one gather over the whole ghost
// shell, cell k.  The three serial loops copy each first-layer ghost from the
// cell with its coordinates clamped into the interior (the edge and corner
// copies read ghosts written by the earlier loops, which already hold those
// clamped values).  Cells beyond the first layer (ghost > 1) are never
// written by the reference after the solve zeroes P; zero_outer reproduces
// that zero.
__device__ __forceinline__ void pressure_bc_cell(const DField &P, const Geom &G, int64_t k,
                                                 bool zero_outer)
{
    const int g = G.g;
    int x, y, z;
    const int Lx = P.nx, Ly = P.ny, Lz = P.nz, t = g;
    // 32-bit unsigned division: the shell is far smaller than 2^31 cells.
    const unsigned ux = (unsigned)Lx, uy = (unsigned)Ly;
    const unsigned uyi = (unsigned)(Ly - 2 * t), uzi = (unsigned)(Lz - 2 * t);
    unsigned r = (unsigned)k;
    const unsigned fz = 2u * t * ux * uy;
    if (r < fz) {
        unsigned q = r / ux;
        x = (int)(r - q * ux);
        r = q;
        q = r / uy;
        y = (int)(r - q * uy);
        const int s = (int)q;
        z = s < t ? s : Lz - 2 * t + s;
    } else {
        r -= fz;
        const unsigned fy = 2u * t * ux * uzi;
        if (r < fy) {
            unsigned q = r / ux;
            x = (int)(r - q * ux);
            r = q;
            q = r / uzi;
            z = t + (int)(r - q * uzi);
            const int s = (int)q;
            y = s < t ? s : Ly - 2 * t + s;
        } else {
            r -= fy;
            unsigned q = r / uyi;
            y = t + (int)(r - q * uyi);
            r = q;
            q = r / uzi;
            z = t + (int)(r - q * uzi);
            const int s = (int)q;
            x = s < t ? s : Lx - 2 * t + s;
        }
    }
    const bool layer1 = x >= g - 1 && x <= g + G.NX && y >= g - 1 && y <= g + G.NY &&
                        z >= g - 1 && z <= g + G.NZ;
    double v = 0;
    if (layer1) {
        const int cx = x < g ? g : (x >= g + G.NX ? g + G.NX - 1 : x);
        const int cy = y < g ? g : (y >= g + G.NY ? g + G.NY - 1 : y);
        const int cz = z < g ? g : (z >= g + G.NZ ? g + G.NZ - 1 : z);
        v = gp(P.d)[idx(P, cx, cy, cz)];
    } else if (!zero_outer) {
        return;
    }
    gp(P.d)[idx(P, x, y, z)] = v;
}

// subtractGradient, project.go:172-205.  The thread of interior column
// (x, y) at z updates U, V (z < g+NZ) and W; the last column, row and z
// position also update the wall faces U(g+NX), V(g+NY), W(g+NZ).  With BC
// the updated faces go through store_face_bc: the velocity BC that follows
// the gradient in project (project.go:33) is then already applied (the
// faces a thread updates are exactly the BC's base cells).
template <bool BC, int C>
__device__ __forceinline__ void put_grad(const Geom &G, const DField &T, int x, int y, int z,
                                         double v)
{
    if (BC)
        store_face_bc<C>(G, T, x, y, z, v);
    else
        gp(T.d)[idx(T, x, y, z)] = v;
}

template <bool BC>
__device__ __forceinline__ void gradient_cell(const DField &U, const DField &V, const DField &W,
                                              const DField &P, const Geom &G, double scale, int x,
                                              int y, int z)
{
    const int g = G.g;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const auto p = gp((const double *)P.d);
    const auto u = gp(U.d), v = gp(V.d), w = gp(W.d);
    const int ip = idx(P, x, y, z);
    if (z < g + G.NZ) {
        const int iu = idx(U, x, y, z);
        put_grad<BC, 0>(G, U, x, y, z, u[iu] - scale * (p[ip] - p[ip - 1]));
        if (x == g + G.NX - 1)
            put_grad<BC, 0>(G, U, x + 1, y, z, u[iu + 1] - scale * (p[ip + 1] - p[ip]));
        const int iv = idx(V, x, y, z);
        put_grad<BC, 1>(G, V, x, y, z, v[iv] - scale * (p[ip] - p[ip - P.sy]));
        if (y == g + G.NY - 1)
            put_grad<BC, 1>(G, V, x, y + 1, z, v[iv + V.sy] - scale * (p[ip + P.sy] - p[ip]));
    }
    const int iw = idx(W, x, y, z);
    put_grad<BC, 2>(G, W, x, y, z, w[iw] - scale * (p[ip] - p[ip - P.sz]));
}

// Faces (x, y, z) and (x+1, y, z) of component C of T: with BC, the wall
// value and the ghosts each face is the base of, as store_face_bc; the two
// values in one 16-byte store, or the first alone when `two` is false.
template <bool BC, int C>
__device__ __forceinline__ void put_pair(const Geom &G, const DField &T, int x, int y, int z,
                                         double v0, double v1, bool two)
{
    if (BC) {
        v0 = face_bc_value<C>(G, x, y, z, v0);
        v1 = face_bc_value<C>(G, x + 1, y, z, v1);
    }
    double *d = T.d + idx(T, x, y, z);
    if (two)
        st2(d, v0, v1);
    else
        gp(T.d)[(int)(d - T.d)] = v0;
    if (BC) {
        face_bc_ghosts<C>(G, T, x, y, z, v0);
        if (two) face_bc_ghosts<C>(G, T, x + 1, y, z, v1);
    }
}

// gradient_cell for the cell pair x = 2k, 2k+1 of row yi (interior
// coordinates) at positions z0 .. z0+ZG-1 (padded; g+NZ is W's wall face),
// with the same values: P, U, V and W read as x-adjacent 16-byte pairs (P's
// x-neighbour of cell 2k as one more value), the faces written by put_pair.
// Every position's loads are issued before the first wait and every result
// is held in a register before the first store.  A position past the last
// (z > g+NZ) loads from the last and stores nothing.  The wall faces of the
// last column and row are updated afterwards, position by position.
template <bool BC, int ZG>
__device__ __forceinline__ void gradient_pair(const DField &U, const DField &V, const DField &W,
                                              const DField &P, const Geom &G, double scale, int k,
                                              int yi, int z0)
{
    const int g = G.g, NX = G.NX, NY = G.NY, NXH = (NX + 1) / 2;
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const bool two = x0 + 1 < NX;
    const int x = g + x0, y = g + yi;
    const int zl = g + G.NZ;  // last position (W's wall face)
    const auto p = gp((const double *)P.d);
    double2 pc[ZG], py[ZG], pz[ZG], uu[ZG], vv[ZG], ww[ZG];
    double px[ZG];
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        const int z = z0 + r < zl ? z0 + r : zl;
        const int ip = idx(P, x, y, z);
        // x0 + 1 + g is inside every padded row, and z = g+NZ is a ghost
        // plane of P, U and V: every pair is in range.
        pc[r] = ld2(P.d + ip);
        px[r] = p[ip - 1];
        py[r] = ld2(P.d + ip - P.sy);
        pz[r] = ld2(P.d + ip - P.sz);
        uu[r] = ld2(U.d + idx(U, x, y, z));
        vv[r] = ld2(V.d + idx(V, x, y, z));
        ww[r] = ld2(W.d + idx(W, x, y, z));
    }
    double ou0[ZG], ou1[ZG], ov0[ZG], ov1[ZG], ow0[ZG], ow1[ZG];
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        ou0[r] = uu[r].x - scale * (pc[r].x - px[r]);
        ou1[r] = uu[r].y - scale * (pc[r].y - pc[r].x);
        ov0[r] = vv[r].x - scale * (pc[r].x - py[r].x);
        ov1[r] = vv[r].y - scale * (pc[r].y - py[r].y);
        ow0[r] = ww[r].x - scale * (pc[r].x - pz[r].x);
        ow1[r] = ww[r].y - scale * (pc[r].y - pz[r].y);
        asm volatile("" : "+v"(ou0[r]), "+v"(ou1[r]), "+v"(ov0[r]), "+v"(ov1[r]), "+v"(ow0[r]),
                     "+v"(ow1[r]));
    }
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        const int z = z0 + r;
        if (z < zl) {
            put_pair<BC, 0>(G, U, x, y, z, ou0[r], ou1[r], two);
            put_pair<BC, 1>(G, V, x, y, z, ov0[r], ov1[r], two);
        }
        if (z <= zl) put_pair<BC, 2>(G, W, x, y, z, ow0[r], ow1[r], two);
    }
    // The wall faces of the last column (U at x = g+NX) and row (V at
    // y = g+NY), as gradient_cell forms them.
    const int xl = two ? x + 1 : x;
    const bool lastx = xl == g + NX - 1, lasty = y == g + NY - 1;
    if (lastx || lasty) {
        const auto u = gp(U.d), v = gp(V.d);
        for (int r = 0; r < ZG; r++) {
            const int z = z0 + r;
            if (z >= zl) break;
            if (lastx) {
                const int ip = idx(P, xl, y, z);
                put_grad<BC, 0>(G, U, xl + 1, y, z,
                                u[idx(U, xl, y, z) + 1] - scale * (p[ip + 1] - p[ip]));
            }
            if (lasty) {
                for (int c = 0; c < (two ? 2 : 1); c++) {
                    const int ip = idx(P, x + c, y, z);
                    put_grad<BC, 1>(G, V, x + c, y + 1, z,
                                    v[idx(V, x + c, y, z) + V.sy] - scale * (p[ip + P.sy] - p[ip]));
                }
            }
        }
    }
}

// gradient_pair<true, ZG> with P read from the colour-split solution (red,
// black: interior cells, NXH = ceil(NX/2) entries per row) instead of the
// natural layout: P's interior cells are the split entries, and the only
// ghosts the gradient reads are first-layer face ghosts, which
// applyPressureBC (project.go:139-163) fills with the interior cell beside
// them -- here the split entry at the clamped coordinate.  So the values,
// and the bits, are those of merging the solution into P, applying the
// pressure BC and then gradient_pair.  TS = float reads the mixed solve's
// FP32 correction, whose widening is exact: the values merging it into P
// would give.
//
// A thread owns the cell pair x = 2k, 2k+1 of row yi (interior
// coordinates) at positions z0 .. z0+ZG-1 (padded; g+NZ is W's wall face).
// With Pout it also writes the pair's P values to Pout's interior (the
// merge; Pout's ghosts are left to applyPressureBC).  Cell 2k has the row's colour and
// 2k+1 the other, so each P value the pair needs is entry k (or k-1 for
// cell 2k-1) of a colour array chosen per row and plane, uniform across a
// wave: the pair's own values, the row below and the plane below (the
// clamped coordinate where that is outside the interior, as applyPressureBC
// fills the ghost).  U, V and W are read and written as 16-byte pairs; a
// wall face's +0 and the ghosts whose base is a face are written as
// store_face_bc does.
template <int ZG, typename TS>
__device__ __forceinline__ void gradient_pair_split(const DField &U, const DField &V,
                                                    const DField &W, const TS *red,
                                                    const TS *black, const Geom &G, double scale,
                                                    int k, int yi, int z0,
                                                    const DField *Pout = nullptr)
{
    const int g = G.g, NX = G.NX, NY = G.NY, NZ = G.NZ, NXH = (NX + 1) / 2;
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const bool two = x0 + 1 < NX;
    const int zl = g + NZ;  // last position (W's wall face)
    const int km = k > 0 ? k - 1 : 0;
    const int ym = yi > 0 ? yi - 1 : 0;
    const auto sr = gp(red), sb = gp(black);
    // Entry kk of row yy, plane zz, in the array of cell 2kk's colour there
    // (second false) or of cell 2kk+1's (second true).
    const auto at = [&](int zz, int yy, int kk, bool second) -> double {
        const bool b = ((yy + zz + 3 * g) & 1) ^ (int)second;
        const int i = (zz * NY + yy) * NXH + kk;
        return (double)(b ? sb[i] : sr[i]);
    };
    double pc0[ZG], pc1[ZG], px0[ZG], py0[ZG], py1[ZG], pz0[ZG], pz1[ZG];
    double2 uu[ZG], vv[ZG], ww[ZG];
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        const int z = z0 + r < zl ? z0 + r : zl;
        const int zi = z - g;
        const int zc = zi < NZ ? zi : NZ - 1;                    // P's plane, clamped
        const int zm = zi - 1 < 0 ? 0 : (zi - 1 < NZ ? zi - 1 : NZ - 1);
        pc0[r] = at(zc, yi, k, false);
        pc1[r] = at(zc, yi, k, true);
        // Cell 2k-1 is the second cell of entry k-1; at k == 0 the clamped
        // neighbour is cell 0 itself.
        px0[r] = k > 0 ? at(zc, yi, km, true) : pc0[r];
        // The row and plane below (their colours follow from their parity);
        // a clamped one is the pair's own row or plane.
        py0[r] = yi > 0 ? at(zc, ym, k, false) : pc0[r];
        py1[r] = yi > 0 ? at(zc, ym, k, true) : pc1[r];
        pz0[r] = zm != zc ? at(zm, yi, k, false) : pc0[r];
        pz1[r] = zm != zc ? at(zm, yi, k, true) : pc1[r];
        // z = g+NZ is a ghost plane of U and V, and x0 + 1 + g is inside
        // every padded row: in range.
        uu[r] = ld2(U.d + idx(U, g + x0, g + yi, z));
        vv[r] = ld2(V.d + idx(V, g + x0, g + yi, z));
        ww[r] = ld2(W.d + idx(W, g + x0, g + yi, z));
    }
    double ou0[ZG], ou1[ZG], ov0[ZG], ov1[ZG], ow0[ZG], ow1[ZG];
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        ou0[r] = uu[r].x - scale * (pc0[r] - px0[r]);
        ou1[r] = uu[r].y - scale * (pc1[r] - pc0[r]);
        ov0[r] = vv[r].x - scale * (pc0[r] - py0[r]);
        ov1[r] = vv[r].y - scale * (pc1[r] - py1[r]);
        ow0[r] = ww[r].x - scale * (pc0[r] - pz0[r]);
        ow1[r] = ww[r].y - scale * (pc1[r] - pz1[r]);
        asm volatile("" : "+v"(ou0[r]), "+v"(ou1[r]), "+v"(ov0[r]), "+v"(ov1[r]), "+v"(ow0[r]),
                     "+v"(ow1[r]));
    }
    const int x = g + x0, y = g + yi;
#pragma unroll
    for (int r = 0; r < ZG; r++) {
        const int z = z0 + r;
        if (z < zl) {
            put_pair<true, 0>(G, U, x, y, z, ou0[r], ou1[r], two);
            put_pair<true, 1>(G, V, x, y, z, ov0[r], ov1[r], two);
            // The merge of the solution into P's interior (the values
            // from_split_pair stores), for a caller that keeps P.
            if (Pout) put_pair<false, 0>(G, *Pout, x, y, z, pc0[r], pc1[r], two);
        }
        if (z <= zl) put_pair<true, 2>(G, W, x, y, z, ow0[r], ow1[r], two);
    }
    // The wall faces of the last column (U at x = g+NX) and row (V at
    // y = g+NY), from the last cell of the pair.
    const int xl = two ? x + 1 : x;
    const bool lastx = xl == g + NX - 1, lasty = y == g + NY - 1;
    if (lastx || lasty) {
        const auto u = gp(U.d), v = gp(V.d);
        for (int r = 0; r < ZG; r++) {
            const int z = z0 + r;
            if (z >= zl) break;
            if (lastx) {
                const double c = two ? pc1[r] : pc0[r];
                // P beyond the last column is the clamped cell itself.
                store_face_bc<0>(G, U, xl + 1, y, z, u[idx(U, xl, y, z) + 1] - scale * (c - c));
            }
            if (lasty) {
                store_face_bc<1>(G, V, x, y + 1, z,
                                 v[idx(V, x, y, z) + V.sy] - scale * (pc0[r] - pc0[r]));
                if (two)
                    store_face_bc<1>(G, V, x + 1, y + 1, z,
                                     v[idx(V, x + 1, y, z) + V.sy] - scale * (pc1[r] - pc1[r]));
            }
        }
    }
}

// ---------------------------------------------------- colour-split layout
//
// A colour pass of gaussSeidelSweep (project.go:103-130) writes every cell of
// one colour from its six neighbours -- all of the other colour -- and its
// divergence; it never reads the cell's own old value (except where the
// Neumann mirror substitutes it for a ghost neighbour).  The split layout
// stores interior cells only: row (y', z') of colour c holds the cells
// x' = x0 + 2k (x0 = first x' of colour c in that row) at index
// ((z'*NY + y') * NXH + k), NXH = ceil(NX/2).  For a cell of colour c at x',
// its x-neighbours are entries (x'-1)>>1 and (x'+1)>>1 of the other-colour
// row, and its y/z neighbours are entry x'>>1 of the adjacent other-colour
// rows -- the same k.

template <typename T>
__device__ __forceinline__ T relax(T sum, T rhs, double h2);

template <>
__device__ __forceinline__ double relax<double>(double sum, double rhs, double h2)
{
    return (sum - h2 * rhs) / 6.0;
}

template <>
__device__ __forceinline__ float relax<float>(float sum, float rhs, double)
{
    return (sum + rhs) * (1.0f / 6.0f);
}

// Parity of interior cell (x', y', z') in padded coordinates: red = 0.
__device__ __forceinline__ int parity(int xi, int yi, int zi, int g)
{
    return (xi + yi + zi + 3 * g) & 1;
}

// One thread of a colour pass: cell k of rows (yi, zb .. zb+ZR-1).
// MODE 0: normal pass.  MODE 1: first pass of a solve from zero (every
// neighbour is +0).  MODE 2: second pass of a solve from zero (the cell's own
// old value, used for mirrored ghosts, is +0).
//
// The thread walks its planes in z.  Entry k of the other colour's rows
// (yi, z) is the z-neighbour of this colour's entry k in rows (yi, z-1) and
// (yi, z+1), and one of the two x-neighbours of entry k in row (yi, z) (x'
// = 2k + x0 has neighbours x'-1 and x'+1, one of which is 2k or 2k+1, entry
// k).  So that column of the other colour is read once per plane into a
// register queue, and a plane needs four more loads: the other x-neighbour,
// the two y-neighbours and the right-hand side.  The planes go in chunks of
// CH: a chunk's loads are all issued, then its results formed, then stored,
// so CH planes' loads are in flight together (in FP64 also the next chunk's,
// see below).  CH holds 32 bytes of results per thread (8 FP32 or 4 FP64
// planes).
//
// Every neighbour value, wall substitution and the order of the sum are
// those of gaussSeidelSweep (project.go:103-130) after applyPressureBC
// (project.go:139-163): a neighbour outside the interior is replaced by the
// cell's own old value (the Neumann mirror).  Out-of-range stand-in indices
// keep every load inside the arrays.
// What a thread's entry k of row yi is, for one parity x0 of the row (the
// first x of this colour): the cell x' = x0 + 2k, or for a row whose last
// entry has no cell of this colour (odd NX) a clamped in-range stand-in that
// is computed and not stored.  Rows of successive planes alternate parity.
struct SplitCol {
    int off;            // the cell's entry in its plane: yi * NXH + (x' >> 1)
    int xo;             // the x-neighbour entry other than k (k itself if both are k)
    unsigned char wx;   // wall bits 0..3: x-, x+, y-, y+ outside the interior
    bool st, live;      // stored; a cell of this colour within reach
    bool xk0, xk1;      // the x- / x+ neighbour is entry k
};

__device__ __forceinline__ SplitCol split_col(int NX, int NXH, int k, int kq, int yi, int x0,
                                              bool ylo, bool yhi)
{
    SplitCol c;
    const int xr = x0 + 2 * k;
    const int xi = xr < NX ? xr : (xr - 2 >= 0 ? xr - 2 : xr);
    c.live = xi < NX;  // false when no cell of this colour is within reach
    const int xs = c.live ? xi : 0;
    c.off = yi * NXH + (xs >> 1);
    c.st = xr < NX && c.live;
    const bool xlo = xs == 0, xhi = xs == NX - 1;
    c.wx = (unsigned char)(xlo | (xhi << 1) | (ylo << 2) | (yhi << 3));
    // A wall cell reads an in-range stand-in for its outside neighbour (the
    // value is replaced by the cell's own).
    const int xm1 = xlo ? (xs + 1 < NX ? xs + 1 : xs) : xs - 1;
    const int xp1 = xhi ? (xs > 0 ? xs - 1 : xs) : xs + 1;
    const int km = xm1 >> 1, kp = xp1 >> 1;
    c.xk0 = km == kq;
    c.xk1 = kp == kq;
    c.xo = yi * NXH + (c.xk0 ? kp : km);
    return c;
}

template <typename T, int MODE, int ZR>
__device__ __forceinline__ void split_pass_column(const SplitArgs<T> &a, int color, int k, int yi,
                                                  int zb)
{
    const int NX = a.NX, NY = a.NY, NZ = a.NZ, NXH = a.NXH;
    if (yi >= NY || zb >= NZ) return;
    const auto cur = gp(color ? a.black : a.red);
    const auto oth = gp((const T *)(color ? a.red : a.black));
    const auto rhs = gp(color ? a.rhs_black : a.rhs_red);
    const int P = NY * NXH;
    constexpr int CH = (int)(ZR * sizeof(T)) > 32 ? 32 / (int)sizeof(T) : ZR;
    static_assert(ZR % CH == 0 && CH % 2 == 0, "whole, even chunks of planes per thread");
    const bool ylo = yi == 0, yhi = yi == NY - 1;
    // Entry k of the other colour in plane zb-1 (zm) and zb (zc): the queue.
    // Planes outside the interior are read as an in-range stand-in (the
    // value is replaced by the wall's own value).  Lanes past the row's end
    // (k >= NXH: the launch covers whole 64-lane rows) store nothing and
    // read entry NXH-1.
    const int kq = k < NXH ? k : NXH - 1;
    const int col = yi * NXH + kq;
    // The column's two parities: plane zb + r has the first for even r.
    const int x0 = (color + yi + zb + 3 * a.g) & 1;
    const SplitCol ca = split_col(NX, NXH, k, kq, yi, x0, ylo, yhi);
    const SplitCol cb = split_col(NX, NXH, k, kq, yi, x0 ^ 1, ylo, yhi);
    T zm = T(0), zc = T(0);
    if (MODE != 1) {
        zm = oth[(zb > 0 ? zb - 1 : zb) * P + col];
        zc = oth[zb * P + col];
    }
#pragma unroll
    for (int zc0 = zb; zc0 < zb + ZR; zc0 += CH) {
        // Phase 1: every plane's index and its loads, all in flight together.
        // A plane past the last (zr >= NZ) is computed on the last plane's
        // indices and not stored.
        T nq[CH], nx[CH], ny0[CH], ny1[CH], rh[CH], cw[CH];
        int at[CH];
        unsigned char wall[CH];  // bit 0..5: x-, x+, y-, y+, z-, z+ outside the interior
#pragma unroll
        for (int r = 0; r < CH; r++) {
            const SplitCol &cc = (r & 1) ? cb : ca;
            const int zr = zc0 + r;
            const int zi = zr < NZ ? zr : NZ - 1;
            const int i = zi * P + cc.off;
            at[r] = (zr < NZ && cc.st) ? i : -1;
            wall[r] = (unsigned char)(cc.wx | ((zi == 0) << 4) | ((zi == NZ - 1) << 5));
            // A wall cell's own old value (the Neumann mirror's stand-in for
            // its outside neighbours); its entry is i, which is k except for
            // a clamped stand-in cell (not stored).
            cw[r] = T(0);
            if (MODE == 0 && wall[r] && cc.live) cw[r] = cur[i];
            if (MODE != 1) {
                // The x-neighbour that is not entry k, the y-neighbours, and
                // entry k of plane zr+1 (the queue's next value).
                nx[r] = oth[zi * P + cc.xo];
                ny0[r] = oth[ylo ? i : i - NXH];
                ny1[r] = oth[yhi ? i : i + NXH];
                const int zn = zr + 1 < NZ ? zr + 1 : NZ - 1;
                nq[r] = oth[zn * P + col];
            }
            rh[r] = rhs[i];
        }
        // Phase 2: a wall cell's own old value stands in for its outside
        // neighbours (the Neumann mirror applied just before the pass).
        T out[CH];
#pragma unroll
        for (int r = 0; r < CH; r++) {
            const SplitCol &cc = (r & 1) ? cb : ca;
            T sum;
            if (MODE == 1) {
                sum = T(0) + T(0) + T(0) + T(0) + T(0) + T(0);
            } else {
                // Entry k of this plane (zc) and of the planes below (zm) and
                // above (nq) are the queue.
                const T c = cw[r];
                const unsigned w = wall[r];
                const T xa = cc.xk0 ? zc : nx[r];
                const T xb = cc.xk1 ? zc : nx[r];
                const T xm = (w & 1) ? c : xa;
                const T xp = (w & 2) ? c : xb;
                const T ym = (w & 4) ? c : ny0[r];
                const T yp = (w & 8) ? c : ny1[r];
                const T zmv = (w & 16) ? c : zm;
                const T zpv = (w & 32) ? c : nq[r];
                sum = xm + xp + ym + yp + zmv + zpv;
                zm = zc;
                zc = nq[r];
            }
            out[r] = relax<T>(sum, rh[r], a.h2);
        }
        // Every result is in a register before the first (conditional) store.
        // Otherwise each plane's arithmetic is sunk into its store's branch and
        // every branch then waits for all outstanding memory operations,
        // earlier stores included, which serialises the stores.
#pragma unroll
        for (int r = 0; r < CH; r++) asm volatile("" : "+v"(out[r]));
#pragma unroll
        for (int r = 0; r < CH; r++)
            if (at[r] >= 0) cur[at[r]] = out[r];
    }
}

// The same pass with its chunks software-pipelined over two register
// buffers: chunk c+1's loads are issued before chunk c's arithmetic and
// stores, so a wave keeps the next chunk's loads in flight while it finishes
// one.  The values and their order are split_pass_column's.  Used for the
// single engine's FP64 pass (8 planes, two chunks: 75 -> 128 VGPRs, 32.2 ->
// 31.5 us at 256^3) and the batched FP64 pass with CHP = 2 (four chunks of
// two planes, 92 VGPRs: 87 -> 86 us on the 96 engines; with the default
// four-plane chunks 142 VGPRs and 94 us).  CHP: planes per chunk (0: 32
// bytes of results per thread).  Measured and not used: FP32 passes
// pipelined (16 or 32 planes: 148-153 VGPRs, 19.5-19.9 us against 17.8 us
// for 8 planes; 8 planes in chunks of 4 or 2: no change).
template <typename T, int MODE, int ZR, int CHP = 0>
__device__ __forceinline__ void split_pass_column_pipelined(const SplitArgs<T> &a, int color, int k,
                                                            int yi, int zb)
{
    const int NX = a.NX, NY = a.NY, NZ = a.NZ, NXH = a.NXH;
    if (yi >= NY || zb >= NZ) return;
    const auto cur = gp(color ? a.black : a.red);
    const auto oth = gp((const T *)(color ? a.red : a.black));
    const auto rhs = gp(color ? a.rhs_black : a.rhs_red);
    const int P = NY * NXH;
    constexpr int CH = CHP > 0 ? CHP : ((int)(ZR * sizeof(T)) > 32 ? 32 / (int)sizeof(T) : ZR);
    static_assert(ZR % CH == 0 && CH % 2 == 0, "whole, even chunks of planes per thread");
    const bool ylo = yi == 0, yhi = yi == NY - 1;
    // Entry k of the other colour in plane zb-1 (zm) and zb (zc): the queue.
    // Planes outside the interior are read as an in-range stand-in (the
    // value is replaced by the wall's own value).  Lanes past the row's end
    // (k >= NXH: the launch covers whole 64-lane rows) store nothing and
    // read entry NXH-1.
    const int kq = k < NXH ? k : NXH - 1;
    const int col = yi * NXH + kq;
    // The column's two parities: plane zb + r has the first for even r.
    const int x0 = (color + yi + zb + 3 * a.g) & 1;
    const SplitCol ca = split_col(NX, NXH, k, kq, yi, x0, ylo, yhi);
    const SplitCol cb = split_col(NX, NXH, k, kq, yi, x0 ^ 1, ylo, yhi);
    T zm = T(0), zc = T(0);
    if (MODE != 1) {
        zm = oth[(zb > 0 ? zb - 1 : zb) * P + col];
        zc = oth[zb * P + col];
    }
    // One chunk's loads: every plane's index and its loads, all in flight
    // together.  A plane past the last (zr >= NZ) is computed on the last
    // plane's indices and not stored.
    auto load = [&](T *nq, T *nx, T *ny0, T *ny1, T *rh, T *cw, int *at, unsigned char *wall,
                    int zc0) {
#pragma unroll
        for (int r = 0; r < CH; r++) {
            const SplitCol &cc = (r & 1) ? cb : ca;
            const int zr = zc0 + r;
            const int zi = zr < NZ ? zr : NZ - 1;
            const int i = zi * P + cc.off;
            at[r] = (zr < NZ && cc.st) ? i : -1;
            // bit 0..5: x-, x+, y-, y+, z-, z+ outside the interior
            wall[r] = (unsigned char)(cc.wx | ((zi == 0) << 4) | ((zi == NZ - 1) << 5));
            // A wall cell's own old value (the Neumann mirror's stand-in for
            // its outside neighbours); its entry is i, which is k except for
            // a clamped stand-in cell (not stored).
            cw[r] = T(0);
            if (MODE == 0 && wall[r] && cc.live) cw[r] = cur[i];
            if (MODE != 1) {
                // The x-neighbour that is not entry k, the y-neighbours, and
                // entry k of plane zr+1 (the queue's next value).
                nx[r] = oth[zi * P + cc.xo];
                ny0[r] = oth[ylo ? i : i - NXH];
                ny1[r] = oth[yhi ? i : i + NXH];
                const int zn = zr + 1 < NZ ? zr + 1 : NZ - 1;
                nq[r] = oth[zn * P + col];
            }
            rh[r] = rhs[i];
        }
    };
    // One chunk's results and stores.  A wall cell's own old value stands in
    // for its outside neighbours (the Neumann mirror applied just before the
    // pass).
    auto finish = [&](const T *nq, const T *nx, const T *ny0, const T *ny1, const T *rh,
                      const T *cw, const int *at, const unsigned char *wall) {
        T out[CH];
#pragma unroll
        for (int r = 0; r < CH; r++) {
            const SplitCol &cc = (r & 1) ? cb : ca;
            T sum;
            if (MODE == 1) {
                sum = T(0) + T(0) + T(0) + T(0) + T(0) + T(0);
            } else {
                // Entry k of this plane (zc) and of the planes below (zm) and
                // above (nq) are the queue.
                const T c = cw[r];
                const unsigned w = wall[r];
                const T xa = cc.xk0 ? zc : nx[r];
                const T xb = cc.xk1 ? zc : nx[r];
                const T xm = (w & 1) ? c : xa;
                const T xp = (w & 2) ? c : xb;
                const T ym = (w & 4) ? c : ny0[r];
                const T yp = (w & 8) ? c : ny1[r];
                const T zmv = (w & 16) ? c : zm;
                const T zpv = (w & 32) ? c : nq[r];
                sum = xm + xp + ym + yp + zmv + zpv;
                zm = zc;
                zc = nq[r];
            }
            out[r] = relax<T>(sum, rh[r], a.h2);
        }
        // Every result is in a register before the first (conditional)
        // store.  Otherwise each plane's arithmetic is sunk into its store's
        // branch and every branch then waits for all outstanding memory
        // operations, earlier stores included, which serialises the stores.
#pragma unroll
        for (int r = 0; r < CH; r++) asm volatile("" : "+v"(out[r]));
#pragma unroll
        for (int r = 0; r < CH; r++)
            if (at[r] >= 0) cur[at[r]] = out[r];
    };
    constexpr int NCH = ZR / CH;
    T qa[CH], xa_[CH], y0a[CH], y1a[CH], ra[CH], wa[CH];
    T qb[CH], xb_[CH], y0b[CH], y1b[CH], rb[CH], wb[CH];
    int ata[CH], atb[CH];
    unsigned char fa[CH], fb[CH];
    load(qa, xa_, y0a, y1a, ra, wa, ata, fa, zb);
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        if (c + 1 < NCH) {
            if (c & 1)
                load(qa, xa_, y0a, y1a, ra, wa, ata, fa, zb + (c + 1) * CH);
            else
                load(qb, xb_, y0b, y1b, rb, wb, atb, fb, zb + (c + 1) * CH);
        }
        if (c & 1)
            finish(qb, xb_, y0b, y1b, rb, wb, atb, fb);
        else
            finish(qa, xa_, y0a, y1a, ra, wa, ata, fa);
    }
}

// Split -> interior cell of the FP64 natural-layout P.  ADD: P += split (the
// correction of the mixed-precision solve); else P = split.
template <typename TS, bool ADD>
__device__ __forceinline__ void from_split_cell(double *nat, int sy, int sz, int g, int NX, int NY,
                                                int NXH, const TS *red, const TS *black, int xi,
                                                int yi, int zi)
{
    if (xi >= NX || yi >= NY) return;
    const int i = (zi * NY + yi) * NXH + (xi >> 1);
    const double v = (double)(parity(xi, yi, zi, g) ? gp(black)[i] : gp(red)[i]);
    const int n = (xi + g) + (yi + g) * sy + (zi + g) * sz;
    gp(nat)[n] = ADD ? gp(nat)[n] + v : v;
}

// Split conversions for the cell pair x = 2k, 2k+1 of interior row
// (yi, zi): entry k of both colour arrays (cell 2k has the row's parity,
// the same for every pair of the row) and the two x-adjacent natural values
// in one 16-byte access.  When the pair's second cell is outside the
// interior (odd NX), the natural row still holds it (a ghost, read and not
// used) and only the first cell is written.
__device__ __forceinline__ void to_split_pair(const double *nat, int sy, int sz, int g, int NX,
                                              int NY, int NXH, double *red, double *black, int k,
                                              int yi, int zi)
{
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const double2 v = ld2(nat + (x0 + g) + (yi + g) * sy + (zi + g) * sz);
    const int i = (zi * NY + yi) * NXH + k;
    const bool b0 = parity(0, yi, zi, g);
    gp(b0 ? black : red)[i] = v.x;
    if (x0 + 1 < NX) gp(b0 ? red : black)[i] = v.y;
}

__device__ __forceinline__ void rhs32_pair(const double *div, int sy, int sz, int g, int NX,
                                           int NY, int NXH, double h2, float *red, float *black,
                                           int k, int yi, int zi)
{
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const double2 v = ld2(div + (x0 + g) + (yi + g) * sy + (zi + g) * sz);
    const int i = (zi * NY + yi) * NXH + k;
    const bool b0 = parity(0, yi, zi, g);
    gp(b0 ? black : red)[i] = (float)(-(h2 * v.x));
    if (x0 + 1 < NX) gp(b0 ? red : black)[i] = (float)(-(h2 * v.y));
}

template <typename TS, bool ADD>
__device__ __forceinline__ void from_split_pair(double *nat, int sy, int sz, int g, int NX, int NY,
                                                int NXH, const TS *red, const TS *black, int k,
                                                int yi, int zi)
{
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const int i = (zi * NY + yi) * NXH + k;
    const bool b0 = parity(0, yi, zi, g);
    const double v0 = (double)gp(b0 ? black : red)[i];
    const double v1 = (double)gp(b0 ? red : black)[i];
    double *d = nat + (x0 + g) + (yi + g) * sy + (zi + g) * sz;
    if (x0 + 1 < NX) {
        double2 o = make_double2(0.0, 0.0);
        if (ADD) o = ld2(d);
        st2(d, ADD ? o.x + v0 : v0, ADD ? o.y + v1 : v1);
    } else {
        const int n = (int)(d - nat);
        gp(nat)[n] = ADD ? gp(nat)[n] + v0 : v0;
    }
}

// from_split_pair for the planes z0 .. z0+NZC-1 (interior coordinates) of
// one pair, every plane's loads issued before the first store.
template <typename TS, bool ADD, int NZC>
__device__ __forceinline__ void from_split_pairs(double *nat, int sy, int sz, int g, int NX, int NY,
                                                 int NZ, int NXH, const TS *red, const TS *black,
                                                 int k, int yi, int z0)
{
    if (k >= NXH || yi >= NY) return;
    const int x0 = 2 * k;
    const bool two = x0 + 1 < NX;
    double v0[NZC], v1[NZC], o0[NZC], o1[NZC];
#pragma unroll
    for (int r = 0; r < NZC; r++) {
        const int zi = z0 + r < NZ ? z0 + r : NZ - 1;
        const int i = (zi * NY + yi) * NXH + k;
        const bool b0 = parity(0, yi, zi, g);
        v0[r] = (double)gp(b0 ? black : red)[i];
        v1[r] = (double)gp(b0 ? red : black)[i];
        o0[r] = o1[r] = 0.0;
        if (ADD) {
            const double2 o = ld2(nat + (x0 + g) + (yi + g) * sy + (zi + g) * sz);
            o0[r] = o.x;
            o1[r] = o.y;
        }
    }
#pragma unroll
    for (int r = 0; r < NZC; r++) {
        if (ADD) {
            v0[r] = o0[r] + v0[r];
            v1[r] = o1[r] + v1[r];
        }
        asm volatile("" : "+v"(v0[r]), "+v"(v1[r]));
    }
#pragma unroll
    for (int r = 0; r < NZC; r++) {
        const int zi = z0 + r;
        if (zi >= NZ) break;
        double *d = nat + (x0 + g) + (yi + g) * sy + (zi + g) * sz;
        if (two)
            st2(d, v0[r], v1[r]);
        else
            gp(nat)[(int)(d - nat)] = v0[r];
    }
}

// FP64 residual of the mixed-precision solve along interior column (xi, yi):
// r = (sum p~ - H^2 div) - 6 p with the Neumann mirror fused (p~ = p for a
// ghost neighbour).  Stores r in FP32 on the split layout and returns the
// column's max |r|.  `div` shares P's layout.
__device__ __forceinline__ double residual_column(const DField &P, const double *div, float *r_red,
                                                  float *r_black, const Geom &G, double h2, int xi,
                                                  int yi)
{
    double m = 0;
    if (xi >= G.NX || yi >= G.NY) return m;
    const int g = G.g;
    const int NXH = (G.NX + 1) / 2;
    const auto p = gp((const double *)P.d);
    const auto dv = gp(div);
    const auto rr = gp(r_red), rb = gp(r_black);
    const int sy = P.sy, sz = P.sz;
    const int x = xi + g, y = yi + g;
    for (int zi = 0; zi < G.NZ; zi++) {
        const int z = zi + g;
        const int i = x + y * sy + z * sz;
        const double c = p[i];
        const double xm = (xi == 0) ? c : p[i - 1];
        const double xp = (xi == G.NX - 1) ? c : p[i + 1];
        const double ym = (yi == 0) ? c : p[i - sy];
        const double yp = (yi == G.NY - 1) ? c : p[i + sy];
        const double zm = (zi == 0) ? c : p[i - sz];
        const double zp = (zi == G.NZ - 1) ? c : p[i + sz];
        const double s = ((((xm + xp) + ym) + yp) + zm) + zp;
        const double res = (s - h2 * dv[i]) - 6.0 * c;
        const int si = (zi * G.NY + yi) * NXH + (xi >> 1);
        if ((x + y + z) & 1) rb[si] = (float)res; else rr[si] = (float)res;
        const double a = fabs(res);
        m = a > m ? a : m;
    }
    return m;
}

// The mixed solve's first-cycle close for cell pair x = 2k, 2k+1 of rows
// (yi, z0 .. z0+kMergeZ-1): P's interior takes the FP32 correction e
// (p = e, exact widening), and the FP64 residual of that p,
// r = (sum p~ - H^2 div) - 6 p with the Neumann mirror fused, is formed from
// the same e values (a neighbour's p is its e), stored in FP32 on the split
// layout for the next cycle.  Returns the max |r| over the cells stored.
// The values and their operation order are those of the merge
// (from_split_pair) followed by residual_column, so the results are the
// same bits; the close moves 24 B per cell (e 4, Div 8, P 8, r 4) instead of
// the 32 B of the two, and a thread issues all of its loads before the
// first wait.  SP / SR: store P / r (the mixed solve's read-only check of
// p = e stores neither and moves 12 B per cell, e and Div).
//
// The pair holds one cell of each colour (cell 2k has the row's parity, so
// the colour array of each cell is uniform across a row).  Its own values,
// its x-neighbours 2k-1 and 2k+2 (entries k-1 and k+1 of the other and the
// same colour), its y-neighbours (entry k of both colours in rows y-1, y+1)
// and its z-neighbours (the pair's own values in the planes beside it) are
// whole loads per lane, and Div and P are read and written as 16-byte pairs.
// `div` shares P's layout.
constexpr int kMergeZ = 4;

template <bool SP = true, bool SR = true>
__device__ __forceinline__ double merge_residual_pair(const DField &P, const double *div,
                                                      const float *e_red, const float *e_black,
                                                      float *r_red, float *r_black, const Geom &G,
                                                      double h2, int k, int yi, int z0)
{
    constexpr int ZC = kMergeZ;
    const int NX = G.NX, NY = G.NY, NZ = G.NZ, NXH = (NX + 1) / 2, g = G.g;
    double m = 0;
    if (k >= NXH || yi >= NY || z0 >= NZ) return m;
    const int x0 = 2 * k;
    const bool two = x0 + 1 < NX;  // the pair's second cell exists
    // Colour arrays of cell x0 (parity of the row) and x0+1 in plane z.
    const auto arr = [&](int z, bool second) -> const float * {
        return (parity(0, yi, z, g) ^ (int)second) ? e_black : e_red;
    };
    const bool xlo = x0 == 0;
    const bool xhi0 = x0 == NX - 1, xhi1 = x0 + 1 == NX - 1;
    const bool ylo = yi == 0, yhi = yi == NY - 1;
    // In-range stand-ins for outside neighbours (the mirror replaces them).
    const int km = k > 0 ? k - 1 : 0, kp = k + 1 < NXH ? k + 1 : NXH - 1;
    const int ym1 = ylo ? yi : yi - 1, yp1 = yhi ? yi : yi + 1;
    // Phase 1: every load in flight.  Own values of planes z0-1 .. z0+ZC
    // (clamped into the grid; a clamped plane is only read where the
    // mirror replaces it, or not stored).
    double a0[ZC + 2], a1[ZC + 2];
#pragma unroll
    for (int r = 0; r < ZC + 2; r++) {
        int z = z0 - 1 + r;
        z = z < 0 ? 0 : (z >= NZ ? NZ - 1 : z);
        const int i = (z * NY + yi) * NXH + k;
        a0[r] = (double)gp(arr(z, false))[i];
        a1[r] = (double)gp(arr(z, true))[i];
    }
    double xm0[ZC], xp1[ZC], ym0[ZC], ym1v[ZC], yp0[ZC], yp1v[ZC];
    double2 dv[ZC];
#pragma unroll
    for (int r = 0; r < ZC; r++) {
        const int z = (z0 + r) < NZ ? z0 + r : NZ - 1;
        const int row = z * NY;
        const float *c0 = arr(z, false), *c1 = arr(z, true);
        xm0[r] = (double)gp(c1)[(row + yi) * NXH + km];    // cell 2k-1
        xp1[r] = (double)gp(c0)[(row + yi) * NXH + kp];    // cell 2k+2
        ym0[r] = (double)gp(c1)[(row + ym1) * NXH + k];    // rows y+-1 flip colour
        ym1v[r] = (double)gp(c0)[(row + ym1) * NXH + k];
        yp0[r] = (double)gp(c1)[(row + yp1) * NXH + k];
        yp1v[r] = (double)gp(c0)[(row + yp1) * NXH + k];
        // x0 + 1 + g < NX + 2g: the pair's second value is in the row
        // even when that cell is outside the interior.
        dv[r] = ld2(div + (x0 + g) + (yi + g) * P.sy + (z + g) * P.sz);
    }
    // Phase 2: results in registers, in residual_column's order.
    double res0[ZC], res1[ZC];
#pragma unroll
    for (int r = 0; r < ZC; r++) {
        const int zi = z0 + r;
        const bool zlo = zi == 0, zhi = zi >= NZ - 1;
        const double c = a0[r + 1], d = a1[r + 1];
        const double s0 = (((((xlo ? c : xm0[r]) + (xhi0 ? c : d)) + (ylo ? c : ym0[r])) +
                            (yhi ? c : yp0[r])) +
                           (zlo ? c : a0[r])) +
                          (zhi ? c : a0[r + 2]);
        const double s1 = (((((c) + (xhi1 ? d : xp1[r])) + (ylo ? d : ym1v[r])) +
                            (yhi ? d : yp1v[r])) +
                           (zlo ? d : a1[r])) +
                          (zhi ? d : a1[r + 2]);
        res0[r] = (s0 - h2 * dv[r].x) - 6.0 * c;
        res1[r] = (s1 - h2 * dv[r].y) - 6.0 * d;
        const double f0 = fabs(res0[r]), f1 = fabs(res1[r]);
        m = (zi < NZ && f0 > m) ? f0 : m;
        m = (zi < NZ && two && f1 > m) ? f1 : m;
    }
#pragma unroll
    for (int r = 0; r < ZC; r++) asm volatile("" : "+v"(res0[r]), "+v"(res1[r]));
    // Phase 3: stores.
    const auto p = gp(P.d);
#pragma unroll
    for (int r = 0; r < ZC; r++) {
        const int zi = z0 + r;
        if (zi >= NZ) break;
        if (SP) {
            double *pp = P.d + (x0 + g) + (yi + g) * P.sy + (zi + g) * P.sz;
            if (two)
                st2(pp, a0[r + 1], a1[r + 1]);
            else
                p[pp - P.d] = a0[r + 1];
        }
        if (SR) {
            const int i = (zi * NY + yi) * NXH + k;
            const bool black0 = parity(0, yi, zi, g);
            gp(black0 ? r_black : r_red)[i] = (float)res0[r];
            if (two) gp(black0 ? r_red : r_black)[i] = (float)res1[r];
        }
    }
    return m;
}

// Max over a 256-thread workgroup (threads numbered x-fastest), fixed tree.
__device__ __forceinline__ double block_max256(double v, double *sh)
{
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t + s] > sh[t] ? sh[t + s] : sh[t];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();  // sh is reused by the caller's next item
    return r;
}

}  // namespace body
}  // namespace af
