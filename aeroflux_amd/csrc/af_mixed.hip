// af_mixed.hip -- mixed-precision pressure solve for gfx950.
//
// The FP64 solve relaxes  A p = b,  (A p)_i = 6 p_i - sum_nb p~_nb,
// b_i = -H^2 div_i, with the Neumann mirror p~_nb = p_i for a neighbour in
// the ghost layer (project.go:103-163).  Here each cycle smooths the
// correction equation  A e = r  with FP32 red-black passes from e = 0 (half
// the bytes of an FP64 pass), then applies  p += e  and recomputes the
// residual  r = b - A p  in FP64.  Because Gauss-Seidel is affine, a cycle
// continues the same iteration the FP64 path runs; FP64 residuals keep the
// FP32 rounding from accumulating.
//
// After the gradient subtract, the discrete divergence of every cell is
// exactly  -(DT/(rho H^2)) r_i  (up to rounding), so the FP64 residual's
// max is the post-projection max|div|: cycles stop once it reaches the
// target.  Cycles after the target is met run as empty launches (a device
// flag), which keeps the whole solve capturable in one graph.
#include "af_mixed.h"
#include "af_rbgs_tb.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int RB = 256;

__device__ __forceinline__ int idx(const DField &f, int x, int y, int z)
{
    return x + y * f.sy + z * f.sz;
}

inline dim3 grid3(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

// First red pass of a cycle from e == 0: e = r/6 on red cells, 0 on black.
__global__ __launch_bounds__(BX *BY) void k_e_first(DField P, float *__restrict__ e,
                                                    const float *__restrict__ r, Geom G,
                                                    const int *done, int check_done)
{
    if (check_done && *done) return;
    const int g = G.g;
    const int x = g + blockIdx.x * BX + threadIdx.x;
    const int y = g + blockIdx.y * BY + threadIdx.y;
    const int z = g + blockIdx.z;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const int i = idx(P, x, y, z);
    // (0 + r) keeps the sign of zero that a pass over all-zero neighbours gives.
    e[i] = (((x + y + z) & 1) == 0) ? (0.0f + r[i]) * (1.0f / 6.0f) : 0.0f;
}

// FP32 red-black pass on the correction, Neumann mirror fused.
__global__ __launch_bounds__(BX *BY) void k_e_pass(DField P, float *__restrict__ e,
                                                   const float *__restrict__ r, Geom G, int color,
                                                   const int *done, int check_done)
{
    if (check_done && *done) return;
    const int g = G.g;
    const int y = g + blockIdx.y * BY + threadIdx.y;
    const int z = g + blockIdx.z;
    if (y >= g + G.NY) return;
    const int start = g + ((((g + y + z) & 1) != color) ? 1 : 0);
    const int x = start + 2 * (int)(blockIdx.x * BX + threadIdx.x);
    if (x >= g + G.NX) return;
    const int i = idx(P, x, y, z);
    const int sy = P.sy, sz = P.sz;
    const float c = e[i];
    const float xm = (x == g) ? c : e[i - 1];
    const float xp = (x == g + G.NX - 1) ? c : e[i + 1];
    const float ym = (y == g) ? c : e[i - sy];
    const float yp = (y == g + G.NY - 1) ? c : e[i + sy];
    const float zm = (z == g) ? c : e[i - sz];
    const float zp = (z == g + G.NZ - 1) ? c : e[i + sz];
    const float s = ((((xm + xp) + ym) + yp) + zm) + zp;
    e[i] = (s + r[i]) * (1.0f / 6.0f);
}

__device__ __forceinline__ double block_max(double v, double *sh)
{
    const int t = threadIdx.y * BX + threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t + s] > sh[t] ? sh[t + s] : sh[t];
        __syncthreads();
    }
    return sh[0];
}

// p (+)= e in FP64, interior only.  FIRST: p was zero, so p = e exactly
// (this is also the solve's zeroing of the interior).
template <bool FIRST>
__global__ __launch_bounds__(BX *BY) void k_correct(DField P, const float *__restrict__ e, Geom G,
                                                    const int *done, int check_done)
{
    if (check_done && *done) return;
    const int g = G.g;
    const int x = g + blockIdx.x * BX + threadIdx.x;
    const int y = g + blockIdx.y * BY + threadIdx.y;
    const int z = g + blockIdx.z;
    if (x >= g + G.NX || y >= g + G.NY) return;
    const int i = idx(P, x, y, z);
    P.d[i] = FIRST ? (double)e[i] : P.d[i] + (double)e[i];
}

// FP64 residual r = (sum p~ - H^2 div) - 6 p with the Neumann mirror fused;
// stores r in FP32 for the next cycle and a per-block max|r|.  The loop over
// z keeps the grid at one block per (x,y) tile so partials stay few.
__global__ __launch_bounds__(BX *BY) void k_residual(DField P, const double *__restrict__ div,
                                                     float *__restrict__ r, Geom G, double h2,
                                                     double *partials, const int *done,
                                                     int check_done)
{
    __shared__ double sh[RB];
    if (check_done && *done) return;  // uniform across the block
    const int g = G.g;
    const int x = g + blockIdx.x * BX + threadIdx.x;
    const int y = g + blockIdx.y * BY + threadIdx.y;
    double m = 0;
    if (x < g + G.NX && y < g + G.NY) {
        const double *p = P.d;
        const int sy = P.sy, sz = P.sz;
        for (int z = g; z < g + G.NZ; z++) {
            const int i = idx(P, x, y, z);
            const double c = p[i];
            const double xm = (x == g) ? c : p[i - 1];
            const double xp = (x == g + G.NX - 1) ? c : p[i + 1];
            const double ym = (y == g) ? c : p[i - sy];
            const double yp = (y == g + G.NY - 1) ? c : p[i + sy];
            const double zm = (z == g) ? c : p[i - sz];
            const double zp = (z == g + G.NZ - 1) ? c : p[i + sz];
            const double s = ((((xm + xp) + ym) + yp) + zm) + zp;
            const double res = (s - h2 * div[i]) - 6.0 * c;
            r[i] = (float)res;
            const double a = fabs(res);
            m = a > m ? a : m;
        }
    }
    const double b = block_max(m, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = b;
}

// Folds the partials in a fixed order, converts to max|div| and raises
// `done` once it is within target.
__global__ __launch_bounds__(RB) void k_res_final(const double *partials, int n, double res_to_div,
                                                  double target, int *done, double *last,
                                                  int *cycles, int check_done)
{
    __shared__ double sh[RB];
    if (check_done && *done) return;
    double v = 0;
    for (int k = threadIdx.x; k < n; k += RB) v = partials[k] > v ? partials[k] : v;
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t + s] > sh[t] ? sh[t + s] : sh[t];
        __syncthreads();
    }
    if (t == 0) {
        const double d = res_to_div * sh[0];
        *last = d;
        *cycles = *cycles + 1;
        if (d <= target) *done = 1;
    }
}

__global__ void k_mixed_reset(int *done, int *cycles)
{
    *done = 0;
    *cycles = 0;
}

__global__ void k_mixed_count(int *cycles) { *cycles = *cycles + 1; }

// `passes` FP32 red-black passes (red first) on A e = r from e = 0.
// Returns (through *out) the buffer index holding the result.
hipError_t smooth(MixedState &m, const MixedCtx &c, int passes, bool gated, hipStream_t s,
                  int *out)
{
    const Geom &G = c.G;
    const int *done = gated ? m.done : nullptr;
    if (c.tile.k > 0) {
        TbArgs<float> a;
        a.rhs = m.r;
        a.NX = G.NX; a.NY = G.NY; a.NZ = G.NZ; a.g = G.g;
        a.sy = c.P.sy; a.sz = c.P.sz;
        a.h2 = 0;
        a.color0 = 0;
        a.cz = c.tile.cz;
        a.done = done;
        int cur = 0;
        int done_passes = 0;
        bool first = true;
        while (done_passes < passes) {
            const int k = (passes - done_passes) < c.tile.k ? (passes - done_passes) : c.tile.k;
            a.src = m.e[cur];
            a.dst = m.e[first ? cur : 1 - cur];
            hipError_t err = launch_rbgs_tb<float>(a, k, c.tile.rows, first, s);
            if (err != hipSuccess) return err;
            if (!first) cur = 1 - cur;
            first = false;
            done_passes += k;
        }
        *out = cur;
        return hipSuccess;
    }
    const dim3 blk(BX, BY);
    const dim3 gfull = grid3(G.NX, G.NY, G.NZ);
    const dim3 ghalf = grid3((G.NX + 1) / 2, G.NY, G.NZ);
    hipLaunchKernelGGL(k_e_first, gfull, blk, 0, s, c.P, m.e[0], m.r, G, m.done, gated ? 1 : 0);
    for (int p = 1; p < passes; p++)
        hipLaunchKernelGGL(k_e_pass, ghalf, blk, 0, s, c.P, m.e[0], m.r, G, p & 1, m.done,
                           gated ? 1 : 0);
    *out = 0;
    return hipGetLastError();
}

}  // namespace

hipError_t init_mixed(MixedState &m, Geom G, DField P, hipStream_t s)
{
    const size_t n = (size_t)P.nx * P.ny * P.nz;
    m.np = ((G.NX + BX - 1) / BX) * ((G.NY + BY - 1) / BY);
    hipError_t err;
    for (int b = 0; b < 2; b++) {
        if ((err = hipMalloc(&m.e[b], n * sizeof(float))) != hipSuccess) return err;
        if ((err = hipMemsetAsync(m.e[b], 0, n * sizeof(float), s)) != hipSuccess) return err;
    }
    if ((err = hipMalloc(&m.r, n * sizeof(float))) != hipSuccess) return err;
    if ((err = hipMalloc(&m.partials, (size_t)m.np * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMalloc(&m.done, 2 * sizeof(int))) != hipSuccess) return err;
    if ((err = hipMalloc(&m.last_res, sizeof(double))) != hipSuccess) return err;
    m.cycles = m.done + 1;
    if ((err = hipMemsetAsync(m.r, 0, n * sizeof(float), s)) != hipSuccess) return err;
    if ((err = hipMemsetAsync(m.done, 0, 2 * sizeof(int), s)) != hipSuccess) return err;
    return hipMemsetAsync(m.last_res, 0, sizeof(double), s);
}

void release_mixed(MixedState &m)
{
    for (auto &b : m.e)
        if (b) (void)hipFree(b);
    if (m.r) (void)hipFree(m.r);
    if (m.partials) (void)hipFree(m.partials);
    if (m.done) (void)hipFree(m.done);
    if (m.last_res) (void)hipFree(m.last_res);
    m = MixedState();
}

hipError_t enqueue_mixed_solve(MixedState &m, const MixedCtx &c, hipStream_t s)
{
    const Geom &G = c.G;
    const dim3 blk(BX, BY);
    const dim3 gfull = grid3(G.NX, G.NY, G.NZ);
    const dim3 gcol((G.NX + BX - 1) / BX, (G.NY + BY - 1) / BY);
    const bool check = c.target > 0;
    hipError_t err;
    int buf = 0;

    hipLaunchKernelGGL(k_mixed_reset, dim3(1), dim3(1), 0, s, m.done, m.cycles);
    // Cycle 0: p = 0, r = b (written by the divergence kernel).
    if ((err = smooth(m, c, 2 * c.sweeps, false, s, &buf)) != hipSuccess) return err;
    hipLaunchKernelGGL(k_correct<true>, gfull, blk, 0, s, c.P, m.e[buf], G, m.done, 0);
    if (!check) {
        hipLaunchKernelGGL(k_mixed_count, dim3(1), dim3(1), 0, s, m.cycles);
    } else {
        hipLaunchKernelGGL(k_residual, gcol, blk, 0, s, c.P, c.D.d, m.r, G, c.h2, m.partials,
                           m.done, 0);
        hipLaunchKernelGGL(k_res_final, dim3(1), dim3(RB), 0, s, m.partials, m.np, c.res_to_div,
                           c.target, m.done, m.last_res, m.cycles, 0);
        for (int cyc = 1; cyc < c.cycles; cyc++) {
            if ((err = smooth(m, c, 2 * c.fix_sweeps, true, s, &buf)) != hipSuccess) return err;
            hipLaunchKernelGGL(k_correct<false>, gfull, blk, 0, s, c.P, m.e[buf], G, m.done, 1);
            hipLaunchKernelGGL(k_residual, gcol, blk, 0, s, c.P, c.D.d, m.r, G, c.h2, m.partials,
                               m.done, 1);
            hipLaunchKernelGGL(k_res_final, dim3(1), dim3(RB), 0, s, m.partials, m.np,
                               c.res_to_div, c.target, m.done, m.last_res, m.cycles, 1);
        }
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    // Final Neumann fill of the FP64 pressure (project.go:98).
    return launch_pressure_bc(c.P, G, G.g > 1, s);
}

}  // namespace af
