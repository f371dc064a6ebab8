// af_kernels.hip -- FP64 kernels of the AeroFlux step data plane for gfx950.
//
// Every kernel reproduces the reference's IEEE-754 binary64 operation order
// exactly (the file is compiled with -ffp-contract=off and additionally pins
// contraction off below), so results are bit-identical to the Go path's
// unfused evaluation.  Thread mappings put x (the unit-stride axis) on the
// 64 lanes of a wavefront so every row access is coalesced.  The per-thread
// bodies shared with the batched kernels are in af_bodies.h.
#include "af_bodies.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

using body::idx;
constexpr int BX = body::BX;  // lanes along x
constexpr int BY = 4;         // rows along y per block

// Block-count triple of a 3-D stencil launch; grids are launched 1-D in
// XCD-aware order and decoded with xcd_block3.
inline int3 nblocks(int nx, int ny, int nz)
{
    return make_int3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
}
inline dim3 grid1(int3 nb) { return dim3((unsigned)(nb.x * nb.y * nb.z)); }

// applyVelocityBC, stencil.go:246-315: one thread per written shell cell.
__global__ __launch_bounds__(256) void k_velocity_bc(DField U, DField V, DField W, Geom G,
                                                     int64_t cu, int64_t cv, int64_t cw)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    body::velocity_bc_cell(U, V, W, G, k, cu, cv, cw);
}

// computeDivergence, project.go:42-62.  SPLIT 1 also writes the value to
// the colour-split layout of the FP64 solve, SPLIT 2 writes the FP32
// right-hand side -(H^2 div) of the mixed-precision solve there: the same
// values k_to_split / k_rhs32_split would form from Div, without reading it
// back.  SPLIT 3 is SPLIT 1 plus the solve's first red pass: the solve
// starts from P == 0 (project.go:73-80), so every neighbour a red cell reads
// in the first pass, and the cell's own value that stands in for a mirrored
// ghost, is +0, and the cell's new value is relax(+0 + ... + +0, div) -- the
// colour pass's own arithmetic (body::split_pass_column, MODE 1) on the
// value just formed.  SPLIT 2 likewise writes the first FP32 red pass of
// the mixed solve's first cycle (from e == 0) into `first` (the
// correction's red array).  DIV false skips the store to Div, for a
// projection whose Div is overwritten before anything reads it.
//
// Each thread forms the cell pair x = 2k, 2k+1 of its row (body::
// divergence_pair), one cell of each colour: Div is written as a 16-byte
// pair and the split arrays take entry k of both colours.
template <int SPLIT, bool DIV>
__global__ __launch_bounds__(BX *BY) void k_divergence(DField U, DField V, DField W, DField D,
                                                       Geom G, double scale, int3 nb, void *red,
                                                       void *black, double h2, void *first)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = G.g;
    const int k = bx * BX + threadIdx.x, yi = by * BY + threadIdx.y, zi = bz;
    double v0, v1;
    if (!body::divergence_pair(U, V, W, G, scale, k, yi, zi, v0, v1)) return;
    const bool two = 2 * k + 1 < G.NX;
    if (DIV) {
        double *d = D.d + body::idx(D, g + 2 * k, g + yi, g + zi);
        if (two)
            body::st2(d, v0, v1);
        else
            body::gp(D.d)[(int)(d - D.d)] = v0;
    }
    if (SPLIT == 0) return;
    const int i = (zi * G.NY + yi) * ((G.NX + 1) / 2) + k;
    // Cell 2k has the row's colour; cell 2k+1 the other.
    const bool b0 = body::parity(0, yi, zi, g);
    void *dst0 = b0 ? black : red, *dst1 = b0 ? red : black;
    const double vr = b0 ? v1 : v0;  // the red cell's value
    const bool has_red = !b0 || two;
    if (SPLIT != 2) {
        body::gp(static_cast<double *>(dst0))[i] = v0;
        if (two) body::gp(static_cast<double *>(dst1))[i] = v1;
        if (SPLIT >= 3 && has_red) {
            const double zero = 0.0;
            body::gp(static_cast<double *>(first))[i] =
                body::relax<double>(zero + zero + zero + zero + zero + zero, vr, h2);
        }
    } else {
        const float r0 = (float)(-(h2 * v0)), r1 = (float)(-(h2 * v1));
        body::gp(static_cast<float *>(dst0))[i] = r0;
        if (two) body::gp(static_cast<float *>(dst1))[i] = r1;
        if (has_red) {
            const float zero = 0.0f;
            body::gp(static_cast<float *>(first))[i] =
                body::relax<float>(zero + zero + zero + zero + zero + zero, b0 ? r1 : r0, 0.0);
        }
    }
}

// gaussSeidelSweep, project.go:103-130: one lane per cell of the colour.
// With FUSED, a neighbour in the ghost layer is replaced by the cell's own
// value: applyPressureBC (project.go:139-163) runs right before every colour
// pass and sets that ghost to exactly this cell's current value, so the
// substitution is bitwise identical and the ghost fill can be skipped.
template <bool FUSED>
__global__ __launch_bounds__(BX *BY) void k_rbgs(DField P, DField D, Geom G, double h2, int color,
                                                     int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = G.g;
    const int y = g + by * BY + threadIdx.y;
    const int z = g + bz;
    if (y >= g + G.NY) return;
    const int start = g + ((((g + y + z) & 1) != color) ? 1 : 0);
    const int x = start + 2 * (int)(bx * BX + threadIdx.x);
    if (x >= g + G.NX) return;
    double *p = P.d;
    const int i = idx(P, x, y, z);
    const int sy = P.sy, sz = P.sz;
    double xm, xp, ym, yp, zm, zp;
    if (FUSED) {
        const double c = p[i];
        xm = (x == g) ? c : p[i - 1];
        xp = (x == g + G.NX - 1) ? c : p[i + 1];
        ym = (y == g) ? c : p[i - sy];
        yp = (y == g + G.NY - 1) ? c : p[i + sy];
        zm = (z == g) ? c : p[i - sz];
        zp = (z == g + G.NZ - 1) ? c : p[i + sz];
    } else {
        xm = p[i - 1]; xp = p[i + 1];
        ym = p[i - sy]; yp = p[i + sy];
        zm = p[i - sz]; zp = p[i + sz];
    }
    const double sum = xm + xp + ym + yp + zm + zp;
    p[i] = (sum - h2 * D.d[i]) / 6.0;
}

// First red pass of solvePressureRedBlack from P == 0 (project.go:81-90):
// every neighbour, ghosts included, is +0, so a red cell becomes
// ((((0+0)+0)+0)+0)+0 - h2*div, over 6, and a black cell stays +0.  Writing
// both colours here also performs the zeroing of the interior.
__global__ __launch_bounds__(BX *BY) void k_rbgs_first(DField P, DField D, Geom G, double h2, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = G.g;
    const int x = g + bx * BX + threadIdx.x;
    const int y = g + by * BY + threadIdx.y;
    const int z = g + bz;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const int i = idx(P, x, y, z);
    double v = 0;
    if (((x + y + z) & 1) == 0) {
        const double sum = 0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0;
        v = (sum - h2 * D.d[i]) / 6.0;
    }
    P.d[i] = v;
}

// applyPressureBC, project.go:139-163: one thread per ghost-shell cell.
// A gated launch (gate != nullptr) runs only when *gate == run_if.
__global__ __launch_bounds__(256) void k_pressure_bc(DField P, Geom G, int64_t count,
                                                     bool zero_outer, const int *gate, int run_if)
{
    if (gate && *gate != run_if) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    body::pressure_bc_cell(P, G, k, zero_outer);
}

__global__ __launch_bounds__(256) void k_zero(double *d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

// subtractGradient, project.go:172-205, one thread per interior cell pair
// (x = 2k, 2k+1) and GZ consecutive z positions (BC: with the velocity BC
// that follows it applied, body::gradient_pair).
constexpr int GZ = 4;

template <bool BC>
__global__ __launch_bounds__(BX *BY) void k_gradient(DField U, DField V, DField W, DField P,
                                                     Geom G, double scale, int3 nb,
                                                     const int *gate, int run_if)
{
    if (gate && *gate != run_if) return;
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    body::gradient_pair<BC, GZ>(U, V, W, P, G, scale, bx * BX + threadIdx.x,
                                by * BY + threadIdx.y, G.g + bz * GZ);
}

// The same with P read from the colour-split solution, a cell pair per
// thread (body::gradient_pair_split).
template <typename TS, bool MERGE>
__global__ __launch_bounds__(BX *BY) void k_gradient_split(DField U, DField V, DField W,
                                                           const TS *red, const TS *black, Geom G,
                                                           double scale, int3 nb, DField P,
                                                           const int *gate, int run_if)
{
    if (gate && *gate != run_if) return;
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    body::gradient_pair_split<GZ, TS>(U, V, W, red, black, G, scale, bx * BX + threadIdx.x,
                                      by * BY + threadIdx.y, G.g + bz * GZ,
                                      MERGE ? &P : nullptr);
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// ------------------------------------------------------------- reductions
//
// Fixed-shape trees, no atomics.  Max is exact in any order; the reference's
// `if a > mx` never admits NaN, which the combine below preserves.

constexpr int RB = 256;  // threads per reduction block

__device__ __forceinline__ double max_keep(double m, double a) { return a > m ? a : m; }

template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = MAX ? max_keep(sh[t], sh[t + s]) : sh[t] + sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// MaxDivergence, project.go:210-232.  Block (bx, by) owns a 64 x 4 column of
// (x, y) and walks every z, so loads stay coalesced along x.
__global__ __launch_bounds__(RB) void k_maxdiv_partial(DField U, DField V, DField W, Geom G,
                                                       double invH, double *partials)
{
    __shared__ double sh[RB];
    const int g = G.g;
    const int x = g + blockIdx.x * BX + (threadIdx.x % BX);
    const int y = g + blockIdx.y * BY + (threadIdx.x / BX);
    double m = 0;
    if (x < g + G.NX && y < g + G.NY) {
        for (int z = g; z < g + G.NZ; z++) {
            const double d = invH * ((U.d[idx(U, x + 1, y, z)] - U.d[idx(U, x, y, z)]) +
                                     (V.d[idx(V, x, y + 1, z)] - V.d[idx(V, x, y, z)]) +
                                     (W.d[idx(W, x, y, z + 1)] - W.d[idx(W, x, y, z)]));
            m = max_keep(m, fabs(d));
        }
    }
    const double r = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = r;
}

// KineticEnergy, project.go:238-258: each lane sums one row serially in the
// reference's term order; rows are combined by a fixed pairwise tree whose
// shape depends only on the grid, never on the launch.
__global__ __launch_bounds__(RB) void k_ke_partial(DField U, DField V, DField W, Geom G,
                                                   double *partials)
{
    __shared__ double sh[RB];
    const int g = G.g;
    const int rows = G.NY * G.NZ;
    double s = 0;
    const int row = blockIdx.x * RB + threadIdx.x;
    if (row < rows) {
        const int y = g + row % G.NY, z = g + row / G.NY;
        for (int x = g; x < g + G.NX; x++) {
            const double uc = 0.5 * (U.d[idx(U, x, y, z)] + U.d[idx(U, x + 1, y, z)]);
            const double vc = 0.5 * (V.d[idx(V, x, y, z)] + V.d[idx(V, x, y + 1, z)]);
            const double wc = 0.5 * (W.d[idx(W, x, y, z)] + W.d[idx(W, x, y, z + 1)]);
            s += uc * uc + vc * vc + wc * wc;
        }
    }
    const double r = block_reduce<false>(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// MaxVelocity, stencil.go:319-331: max |.| over the whole padded arrays.
__global__ __launch_bounds__(RB) void k_maxabs_partial(const double *u, int64_t nu,
                                                       const double *v, int64_t nv,
                                                       const double *w, int64_t nw,
                                                       double *partials, int64_t chunk)
{
    __shared__ double sh[RB];
    double m = 0;
    const int64_t total = nu + nv + nw;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < total ? lo + chunk : total;
    for (int64_t k = lo + threadIdx.x; k < hi; k += RB) {
        const double val = k < nu ? u[k] : (k < nu + nv ? v[k - nu] : w[k - nu - nv]);
        m = max_keep(m, fabs(val));
    }
    const double r = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <bool MAX>
__global__ __launch_bounds__(RB) void k_reduce_final(const double *partials, int n, double *out)
{
    __shared__ double sh[RB];
    // Fixed assignment: lane t folds partials t, t+RB, t+2RB, ... in order.
    double v = 0;
    for (int k = threadIdx.x; k < n; k += RB)
        v = MAX ? max_keep(v, partials[k]) : v + partials[k];
    const double r = block_reduce<MAX>(v, sh);
    if (threadIdx.x == 0) *out = r;
}

}  // namespace

// ---------------------------------------------------------------- launchers

hipError_t launch_velocity_bc(DField U, DField V, DField W, Geom G, hipStream_t s)
{
    int64_t cu, cv, cw;
    body::velocity_bc_counts(G, cu, cv, cw);
    const int64_t n = cu + cv + cw;
    hipLaunchKernelGGL(k_velocity_bc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, V, W,
                       G, cu, cv, cw);
    return hipGetLastError();
}

hipError_t launch_divergence(DField U, DField V, DField W, DField D, Geom G, double scale,
                             hipStream_t s, int split, void *red, void *black, double h2,
                             void *first, bool store_div)
{
    const int3 nb = nblocks((G.NX + 1) / 2, G.NY, G.NZ);
    if (split < 0 || split > 3 || (split >= 2 && !first) || (!store_div && split < 2))
        return hipErrorInvalidValue;
#define AF_DIV_LAUNCH(S, DV)                                                                      \
    hipLaunchKernelGGL((k_divergence<S, DV>), grid1(nb), dim3(BX, BY), 0, s, U, V, W, D, G, scale, \
                       nb, red, black, h2, first)
    if (split == 0)
        AF_DIV_LAUNCH(0, true);
    else if (split == 1)
        AF_DIV_LAUNCH(1, true);
    else if (split == 2 && store_div)
        AF_DIV_LAUNCH(2, true);
    else if (split == 2)
        AF_DIV_LAUNCH(2, false);
    else if (store_div)
        AF_DIV_LAUNCH(3, true);
    else
        AF_DIV_LAUNCH(3, false);
#undef AF_DIV_LAUNCH
    return hipGetLastError();
}

hipError_t launch_rbgs(DField P, DField D, Geom G, double h2, int color, bool fused_bc,
                       hipStream_t s)
{
    const int3 nb = nblocks((G.NX + 1) / 2, G.NY, G.NZ);
    if (fused_bc)
        hipLaunchKernelGGL(k_rbgs<true>, grid1(nb), dim3(BX, BY), 0, s, P, D, G, h2, color, nb);
    else
        hipLaunchKernelGGL(k_rbgs<false>, grid1(nb), dim3(BX, BY), 0, s, P, D, G, h2, color, nb);
    return hipGetLastError();
}

hipError_t launch_rbgs_first(DField P, DField D, Geom G, double h2, hipStream_t s)
{
    const int3 nb = nblocks(G.NX, G.NY, G.NZ);
    hipLaunchKernelGGL(k_rbgs_first, grid1(nb), dim3(BX, BY), 0, s, P, D, G, h2, nb);
    return hipGetLastError();
}

hipError_t launch_pressure_bc(DField P, Geom G, bool zero_outer, hipStream_t s, const int *gate,
                              int run_if)
{
    const int64_t n = body::pressure_shell_count(P.nx, P.ny, P.nz, G.g);
    hipLaunchKernelGGL(k_pressure_bc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, G, n,
                       zero_outer, gate, run_if);
    return hipGetLastError();
}

hipError_t launch_zero(double *d, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

hipError_t launch_gradient(DField U, DField V, DField W, DField P, Geom G, double scale,
                           hipStream_t s, bool velocity_bc, const int *gate, int run_if)
{
    // A thread per cell pair (NXH entries per row); z = g .. g+NZ.
    const int3 nb = nblocks((G.NX + 1) / 2, G.NY, (G.NZ + 1 + GZ - 1) / GZ);
    if (velocity_bc)
        hipLaunchKernelGGL(k_gradient<true>, grid1(nb), dim3(BX, BY), 0, s, U, V, W, P, G, scale,
                           nb, gate, run_if);
    else
        hipLaunchKernelGGL(k_gradient<false>, grid1(nb), dim3(BX, BY), 0, s, U, V, W, P, G, scale,
                           nb, gate, run_if);
    return hipGetLastError();
}

template <typename TS>
hipError_t launch_gradient_split(DField U, DField V, DField W, const TS *red, const TS *black,
                                 Geom G, double scale, hipStream_t s, const DField *P,
                                 const int *gate, int run_if)
{
    // A thread per cell pair (NXH entries per row); z = g .. g+NZ.
    const int3 nb = nblocks((G.NX + 1) / 2, G.NY, (G.NZ + 1 + GZ - 1) / GZ);
    if (P)
        hipLaunchKernelGGL((k_gradient_split<TS, true>), grid1(nb), dim3(BX, BY), 0, s, U, V, W,
                           red, black, G, scale, nb, *P, gate, run_if);
    else
        hipLaunchKernelGGL((k_gradient_split<TS, false>), grid1(nb), dim3(BX, BY), 0, s, U, V, W,
                           red, black, G, scale, nb, DField{}, gate, run_if);
    return hipGetLastError();
}
template hipError_t launch_gradient_split<double>(DField, DField, DField, const double *,
                                                  const double *, Geom, double, hipStream_t,
                                                  const DField *, const int *, int);
template hipError_t launch_gradient_split<float>(DField, DField, DField, const float *,
                                                 const float *, Geom, double, hipStream_t,
                                                 const DField *, const int *, int);

hipError_t launch_set_int(int *p, int v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

int reduce_partials_needed(Geom G)
{
    const int a = (G.NY * G.NZ + RB - 1) / RB;
    const int b = ((G.NX + BX - 1) / BX) * ((G.NY + BY - 1) / BY);
    int m = a > b ? a : b;
    return m > 1024 ? m : 1024;
}

int max_divergence_partials(Geom G)
{
    return ((G.NX + BX - 1) / BX) * ((G.NY + BY - 1) / BY);
}

hipError_t launch_max_divergence(DField U, DField V, DField W, Geom G, double invH,
                                 double *partials, double *out, hipStream_t s)
{
    const dim3 grid((G.NX + BX - 1) / BX, (G.NY + BY - 1) / BY);
    hipLaunchKernelGGL(k_maxdiv_partial, grid, dim3(RB), 0, s, U, V, W, G, invH, partials);
    hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(RB), 0, s, partials,
                       (int)(grid.x * grid.y), out);
    return hipGetLastError();
}

hipError_t launch_kinetic_energy(DField U, DField V, DField W, Geom G, double *partials,
                                 double *out, hipStream_t s)
{
    const int nb = (G.NY * G.NZ + RB - 1) / RB;
    hipLaunchKernelGGL(k_ke_partial, dim3(nb), dim3(RB), 0, s, U, V, W, G, partials);
    hipLaunchKernelGGL(k_reduce_final<false>, dim3(1), dim3(RB), 0, s, partials, nb, out);
    return hipGetLastError();
}

hipError_t launch_max_abs3(const double *u, int64_t nu, const double *v, int64_t nv,
                           const double *w, int64_t nw, double *partials, double *out,
                           hipStream_t s)
{
    const int nb = 1024;
    const int64_t total = nu + nv + nw;
    const int64_t chunk = (total + nb - 1) / nb;
    hipLaunchKernelGGL(k_maxabs_partial, dim3(nb), dim3(RB), 0, s, u, nu, v, nv, w, nw, partials,
                       chunk);
    hipLaunchKernelGGL(k_reduce_final<true>, dim3(1), dim3(RB), 0, s, partials, nb, out);
    return hipGetLastError();
}

}  // namespace af
