// af_kernels.hip -- FP64 kernels of the AeroFlux step data plane for gfx950.
//
// Every kernel reproduces the reference's IEEE-754 binary64 operation order
// exactly (the file is compiled with -ffp-contract=off and additionally pins
// contraction off below), so results are bit-identical to the Go path's
// unfused evaluation.  Thread mappings put x (the unit-stride axis) on the
// 64 lanes of a wavefront so every row access is coalesced.
#include "af_internal.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = 64;  // lanes along x
constexpr int BY = 4;   // rows along y per block

// Block-count triple of a 3-D stencil launch; grids are launched 1-D in
// XCD-aware order and decoded with xcd_block3.
inline int3 nblocks(int nx, int ny, int nz)
{
    return make_int3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
}
inline dim3 grid1(int3 nb) { return dim3((unsigned)(nb.x * nb.y * nb.z)); }

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
__device__ __forceinline__ int64_t shell_count(int Lx, int Ly, int Lz)
{
    return 2LL * Lx * Ly + 2LL * Lx * (Lz - 2) + 2LL * (Ly - 2) * (Lz - 2);
}

__host__ __device__ inline int64_t shell_count_h(int Lx, int Ly, int Lz)
{
    return 2LL * Lx * Ly + 2LL * Lx * (Lz - 2) + 2LL * (Ly - 2) * (Lz - 2);
}

__device__ __forceinline__ void shell_decode(int64_t k, int Lx, int Ly, int Lz, int &x, int &y,
                                             int &z)
{
    int64_t nz = 2LL * Lx * Ly;
    if (k < nz) {
        x = (int)(k % Lx);
        int64_t r = k / Lx;
        y = (int)(r % Ly);
        z = (r / Ly) == 0 ? 0 : Lz - 1;
        return;
    }
    k -= nz;
    int64_t ny = 2LL * Lx * (Lz - 2);
    if (k < ny) {
        x = (int)(k % Lx);
        int64_t r = k / Lx;
        z = 1 + (int)(r % (Lz - 2));
        y = (r / (Lz - 2)) == 0 ? 0 : Ly - 1;
        return;
    }
    k -= ny;
    y = 1 + (int)(k % (Ly - 2));
    int64_t r = k / (Ly - 2);
    z = 1 + (int)(r % (Lz - 2));
    x = (r / (Lz - 2)) == 0 ? 0 : Lx - 1;
}

// Final BC value of box-boundary cell (x,y,z) of a component with normal
// axis n.  Coordinates are padded field coordinates.
__device__ __forceinline__ double vel_bc_value(const DField &f, int n, const Geom &G, int x, int y,
                                               int z)
{
    const int g = G.g;
    int c[3] = {x, y, z};
    const int N[3] = {G.NX, G.NY, G.NZ};
    bool wall = (c[n] == g) || (c[n] == g + N[n]);
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
    double v = wall ? 0.0 : f.d[idx(f, c[0], c[1], c[2])];
    if (flips & 1) v = -v;  // two exact negations are the identity
    return v;
}

__global__ __launch_bounds__(256) void k_velocity_bc(DField U, DField V, DField W, Geom G,
                                                     int64_t cu, int64_t cv, int64_t cw)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
    int x = ox + bx, y = oy + by, z = oz + bz;
    f.d[idx(f, x, y, z)] = vel_bc_value(f, n, G, x, y, z);
}

// -------------------------------------------------------------- advection

// The advection sweeps live in af_advect.hip.

// ------------------------------------------------------------- projection

// computeDivergence, project.go:42-62.
__global__ __launch_bounds__(BX *BY) void k_divergence(DField U, DField V, DField W, DField D,
                                                       Geom G, double scale, int3 nb)
{
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = G.g;
    const int x = g + bx * BX + threadIdx.x;
    const int y = g + by * BY + threadIdx.y;
    const int z = g + bz;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const double du = U.d[idx(U, x + 1, y, z)] - U.d[idx(U, x, y, z)];
    const double dv = V.d[idx(V, x, y + 1, z)] - V.d[idx(V, x, y, z)];
    const double dw = W.d[idx(W, x, y, z + 1)] - W.d[idx(W, x, y, z)];
    D.d[idx(D, x, y, z)] = scale * (du + dv + dw);
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

// applyPressureBC, project.go:139-163, as one gather over the whole ghost
// shell.  The three serial loops copy each first-layer ghost from the cell
// with its coordinates clamped into the interior (the edge and corner
// copies read ghosts written by the earlier loops, which already hold those
// clamped values).  Cells beyond the first layer (ghost > 1) are never
// written by the reference after the solve zeroes P; zero_outer reproduces
// that zero.
__global__ __launch_bounds__(256) void k_pressure_bc(DField P, Geom G, int64_t count,
                                                     bool zero_outer)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int g = G.g;
    int x, y, z;
    // Shell of thickness g over the padded box: decode per family.
    const int Lx = P.nx, Ly = P.ny, Lz = P.nz, t = g;
    int64_t r = k;
    const int64_t fz = 2LL * t * Lx * Ly;
    if (r < fz) {
        x = (int)(r % Lx); r /= Lx;
        y = (int)(r % Ly); r /= Ly;
        const int s = (int)r;
        z = s < t ? s : Lz - 2 * t + s;
    } else {
        r -= fz;
        const int64_t fy = 2LL * t * Lx * (Lz - 2 * t);
        if (r < fy) {
            x = (int)(r % Lx); r /= Lx;
            z = t + (int)(r % (Lz - 2 * t)); r /= (Lz - 2 * t);
            const int s = (int)r;
            y = s < t ? s : Ly - 2 * t + s;
        } else {
            r -= fy;
            y = t + (int)(r % (Ly - 2 * t)); r /= (Ly - 2 * t);
            z = t + (int)(r % (Lz - 2 * t)); r /= (Lz - 2 * t);
            const int s = (int)r;
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
        v = P.d[idx(P, cx, cy, cz)];
    } else if (!zero_outer) {
        return;
    }
    P.d[idx(P, x, y, z)] = v;
}

__global__ __launch_bounds__(256) void k_zero(double *d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

// subtractGradient, project.go:172-205, one thread per union-box coordinate.
__global__ __launch_bounds__(BX *BY) void k_gradient(DField U, DField V, DField W, DField P,
                                                     Geom G, double scale, int3 nb)
{
    // A thread owns interior column (x, y) for one z; the last column, row
    // and z position also update the wall faces U(g+NX), V(g+NY), W(g+NZ).
    int bx, by, bz;
    xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
    const int g = G.g;
    const int x = g + bx * BX + threadIdx.x;
    const int y = g + by * BY + threadIdx.y;
    const int z = g + bz;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const double *p = P.d;
    const int ip = idx(P, x, y, z);
    if (z < g + G.NZ) {
        const int iu = idx(U, x, y, z);
        U.d[iu] -= scale * (p[ip] - p[ip - 1]);
        if (x == g + G.NX - 1) U.d[iu + 1] -= scale * (p[ip + 1] - p[ip]);
        const int iv = idx(V, x, y, z);
        V.d[iv] -= scale * (p[ip] - p[ip - P.sy]);
        if (y == g + G.NY - 1) V.d[iv + V.sy] -= scale * (p[ip + P.sy] - p[ip]);
    }
    const int iw = idx(W, x, y, z);
    W.d[iw] -= scale * (p[ip] - p[ip - P.sz]);
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
    const int64_t cu = shell_count_h(G.NX + 1, G.NY + 2, G.NZ + 2);
    const int64_t cv = shell_count_h(G.NX + 2, G.NY + 1, G.NZ + 2);
    const int64_t cw = shell_count_h(G.NX + 2, G.NY + 2, G.NZ + 1);
    const int64_t n = cu + cv + cw;
    hipLaunchKernelGGL(k_velocity_bc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, V, W,
                       G, cu, cv, cw);
    return hipGetLastError();
}

hipError_t launch_divergence(DField U, DField V, DField W, DField D, Geom G, double scale,
                             hipStream_t s)
{
    const int3 nb = nblocks(G.NX, G.NY, G.NZ);
    hipLaunchKernelGGL(k_divergence, grid1(nb), dim3(BX, BY), 0, s, U, V, W, D, G, scale, nb);
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

hipError_t launch_pressure_bc(DField P, Geom G, bool zero_outer, hipStream_t s)
{
    const int t = G.g;
    const int64_t Lx = P.nx, Ly = P.ny, Lz = P.nz;
    const int64_t n = 2LL * t * Lx * Ly + 2LL * t * Lx * (Lz - 2 * t) +
                      2LL * t * (Ly - 2 * t) * (Lz - 2 * t);
    hipLaunchKernelGGL(k_pressure_bc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, G, n,
                       zero_outer);
    return hipGetLastError();
}

hipError_t launch_zero(double *d, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

hipError_t launch_gradient(DField U, DField V, DField W, DField P, Geom G, double scale,
                           hipStream_t s)
{
    const int3 nb = nblocks(G.NX, G.NY, G.NZ + 1);
    hipLaunchKernelGGL(k_gradient, grid1(nb), dim3(BX, BY), 0, s, U, V, W, P, G, scale, nb);
    return hipGetLastError();
}

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
