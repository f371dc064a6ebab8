// af_mixed.hip -- mixed-precision pressure solve for gfx950.
//
// The FP64 solve relaxes  A p = b,  (A p)_i = 6 p_i - sum_nb p~_nb,
// b_i = -H^2 div_i, with the Neumann mirror p~_nb = p_i for a neighbour in
// the ghost layer (project.go:103-163).  Here each cycle smooths the
// correction equation  A e = r  with FP32 red-black passes from e = 0 on the
// colour-split layout (6 B per cell per pass, a quarter of an interleaved
// FP64 pass), then applies  p += e  and recomputes the residual  r = b - A p
// in FP64.  Because Gauss-Seidel is affine, a cycle continues the same
// iteration the FP64 path runs; FP64 residuals keep the FP32 rounding from
// accumulating.
//
// After the gradient subtract, the discrete divergence of every cell is
// exactly  -(DT/(rho H^2)) r_i  (up to rounding), so the FP64 residual's
// max is the post-projection max|div|: cycles stop once it reaches the
// target.  Cycles after the target is met run as empty launches (a device
// flag), which keeps the whole solve capturable in one graph.
//
// A checked solve's first check only reads: the FP64 residual of p = e
// (exact widening) from the FP32 correction and Div, 12 B per cell, so a
// solve that meets its target at once leaves the solution in the split
// arrays for the gradient to read, as an unchecked one does.  A solve that
// misses it continues the same FP32 iteration by a few sweep pairs (the
// right-hand side is untouched) and checks again; only then do FP64
// defect-correction cycles follow, which merge p = e into P, store the
// residual as the next right-hand side, smooth a fresh correction and add
// it, and raise `merged` so that the caller reads P instead.
#include "af_bodies.h"
#include "af_mixed.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int RB = 256;

// FP64 residual r = (sum p~ - H^2 div) - 6 p with the Neumann mirror fused;
// stores r in FP32 (split layout) for the next cycle and a per-block max|r|.
// The loop over z keeps one block per (x,y) tile so partials stay few.
__global__ __launch_bounds__(BX *BY) void k_residual(DField P, const double *__restrict__ div,
                                                     float *r_red, float *r_black, Geom G,
                                                     double h2, double *partials, const int *done,
                                                     int check_done)
{
    __shared__ double sh[RB];
    if (check_done && *done) return;  // uniform across the block
    const double m = body::residual_column(P, div, r_red, r_black, G, h2,
                                           blockIdx.x * BX + threadIdx.x,
                                           blockIdx.y * BY + threadIdx.y);
    const double b = body::block_max256(m, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = b;
}

// The FP64 residual of p = e in one pass (body::merge_residual_pair), with a
// per-workgroup max|r|: SP also writes e into P's interior, SR stores the
// residual in FP32 as the next cycle's right-hand side; the check stores
// neither.  A workgroup owns 64 x 4 split entries of (x, y) and ZC planes.
// GATED: skipped once the target is met; a fixed grid of workgroups strides
// over the tiles, so a skipped launch dispatches few workgroups.
constexpr int ZC = body::kMergeZ;
constexpr int kGatedCloseGrid = 2048;

template <bool SP, bool SR, bool GATED>
__global__ __launch_bounds__(BX *BY) void k_merge_residual(DField P, const double *__restrict__ div,
                                                           const float *__restrict__ e_red,
                                                           const float *__restrict__ e_black,
                                                           float *r_red, float *r_black, Geom G,
                                                           double h2, double *partials, int3 nb,
                                                           const int *done)
{
    __shared__ double sh[BY];
    if (GATED && *done) return;  // uniform across the launch
    const int total = GATED ? nb.x * nb.y * nb.z : 1;
    for (int l = GATED ? (int)blockIdx.x : 0; l < total; l += GATED ? (int)gridDim.x : 1) {
        int bx, by, bz;
        if (GATED) {
            bx = l % nb.x;
            const int r = l / nb.x;
            by = r % nb.y;
            bz = r / nb.y;
        } else {
            xcd_block3(nb.x, nb.y, nb.z, bx, by, bz);
        }
        double m = body::merge_residual_pair<SP, SR>(P, div, e_red, e_black, r_red, r_black, G, h2,
                                                     bx * BX + threadIdx.x, by * BY + threadIdx.y,
                                                     bz * ZC);
        // Max over the workgroup: across each 64-lane wave by shuffles, then
        // over the four waves (max is exact in any order).
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        if (threadIdx.x == 0) sh[threadIdx.y] = m;
        __syncthreads();
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            double b = sh[0];
            for (int w = 1; w < BY; w++) b = sh[w] > b ? sh[w] : b;
            partials[(bz * nb.y + by) * nb.x + bx] = b;
        }
        __syncthreads();  // sh is rewritten by the next tile
    }
}

template <bool SP, bool SR, bool GATED>
void launch_close(MixedState &m, const MixedCtx &c, hipStream_t s)
{
    const Geom &G = c.G;
    const int3 nb = make_int3(((G.NX + 1) / 2 + BX - 1) / BX, (G.NY + BY - 1) / BY,
                              (G.NZ + ZC - 1) / ZC);
    const int total = nb.x * nb.y * nb.z;
    const int grid = GATED ? (total < kGatedCloseGrid ? total : kGatedCloseGrid) : total;
    hipLaunchKernelGGL((k_merge_residual<SP, SR, GATED>), dim3((unsigned)grid), dim3(BX, BY), 0,
                       s, c.P, c.D.d, m.e_red, m.e_black, m.r_red, m.r_black, G, c.h2,
                       m.partials, nb, m.done);
}

// A defect-correction cycle is about to merge the solution into P: raised
// unless the target is already met.
__global__ void k_mark_merged(const int *done, int *merged)
{
    if (!*done) *merged = 1;
}

// k_res_final over many partials: 1024 lanes each fold a fixed stride of
// them (loads issued in batches of 8), then a fixed tree.
constexpr int RF = 1024;

__global__ __launch_bounds__(RF) void k_res_final_wide(const double *partials, int n,
                                                       double res_to_div, double target, int *done,
                                                       double *last, int *cycles, int check_done)
{
    __shared__ double sh[RF];
    if (check_done && *done) return;
    const int t = threadIdx.x;
    double v = 0;
    for (int k0 = t; k0 < n; k0 += 8 * RF) {
        double q[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + j * RF;
            q[j] = k < n ? partials[k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) v = q[j] > v ? q[j] : v;
    }
    sh[t] = v;
    __syncthreads();
    for (int s = RF / 2; s > 0; s >>= 1) {
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

// cycles0: 1 for a one-cycle solve (its one cycle counted here), 0 for a
// checked one (each check counts).
__global__ void k_mixed_reset(int *done, int *cycles, int *merged, int cycles0)
{
    *done = 0;
    *cycles = cycles0;
    *merged = 0;
}

// `passes` FP32 red-black passes (red first) on A e = r from e = 0; with
// first_done the first red pass's values are already in e_red and the
// passes start at the first black one; with cont they continue from the
// e the arrays hold.
hipError_t smooth(MixedState &m, const MixedCtx &c, int passes, bool gated, hipStream_t s,
                  bool first_done = false, bool cont = false)
{
    SplitArgs<float> a;
    a.red = m.e_red;
    a.black = m.e_black;
    a.rhs_red = m.r_red;
    a.rhs_black = m.r_black;
    a.NX = c.G.NX; a.NY = c.G.NY; a.NZ = c.G.NZ; a.NXH = (c.G.NX + 1) / 2; a.g = c.G.g;
    a.h2 = 0;
    a.done = gated ? m.done : nullptr;
    a.slab = c.slab;
    for (int p = first_done ? 1 : 0; p < passes; p++) {
        const int mode = cont ? 0 : (p == 0 ? 1 : (p == 1 ? 2 : 0));
        hipError_t err = launch_split_pass<float>(a, p & 1, mode, s);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

}  // namespace

hipError_t init_mixed(MixedState &m, Geom G, hipStream_t s)
{
    const size_t n = (size_t)split_count(G);
    m.np = ((G.NX + BX - 1) / BX) * ((G.NY + BY - 1) / BY);
    m.np_merge = (((G.NX + 1) / 2 + BX - 1) / BX) * ((G.NY + BY - 1) / BY) *
                 ((G.NZ + ZC - 1) / ZC);
    hipError_t err;
    float **bufs[4] = {&m.e_red, &m.e_black, &m.r_red, &m.r_black};
    for (float **b : bufs) {
        if ((err = hipMalloc(b, n * sizeof(float))) != hipSuccess) return err;
        if ((err = hipMemsetAsync(*b, 0, n * sizeof(float), s)) != hipSuccess) return err;
    }
    const int np_max = m.np > m.np_merge ? m.np : m.np_merge;
    if ((err = hipMalloc(&m.partials, (size_t)np_max * sizeof(double))) != hipSuccess)
        return err;
    if ((err = hipMalloc(&m.done, 3 * sizeof(int))) != hipSuccess) return err;
    if ((err = hipMalloc(&m.last_res, sizeof(double))) != hipSuccess) return err;
    m.cycles = m.done + 1;
    m.merged = m.done + 2;
    if ((err = hipMemsetAsync(m.done, 0, 3 * sizeof(int), s)) != hipSuccess) return err;
    return hipMemsetAsync(m.last_res, 0, sizeof(double), s);
}

void release_mixed(MixedState &m)
{
    float *bufs[4] = {m.e_red, m.e_black, m.r_red, m.r_black};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    if (m.partials) (void)hipFree(m.partials);
    if (m.done) (void)hipFree(m.done);
    if (m.last_res) (void)hipFree(m.last_res);
    m = MixedState();
}

hipError_t enqueue_mixed_solve(MixedState &m, const MixedCtx &c, hipStream_t s)
{
    const Geom &G = c.G;
    const dim3 blk(BX, BY);
    const dim3 gcol((G.NX + BX - 1) / BX, (G.NY + BY - 1) / BY);
    const bool check = c.target > 0;
    hipError_t err;

    hipLaunchKernelGGL(k_mixed_reset, dim3(1), dim3(1), 0, s, m.done, m.cycles, m.merged,
                       check ? 0 : 1);
    // Cycle 0: p = 0, r = b = -(H^2 div).
    if (!c.rhs_ready &&
        (err = launch_rhs32_split(c.D.d, c.D.sy, c.D.sz, G, c.h2, m.r_red, m.r_black, s)) !=
            hipSuccess)
        return err;
    if ((err = smooth(m, c, 2 * c.sweeps, false, s, c.rhs_ready)) != hipSuccess) return err;
    if (!check) {
        if (c.merge && (err = launch_from_split<float>(c.P.d, c.P.sy, c.P.sz, G, m.e_red,
                                                       m.e_black, false, nullptr, s)) != hipSuccess)
            return err;
        if ((err = hipGetLastError()) != hipSuccess) return err;
        // Final Neumann fill of the FP64 pressure (project.go:98).
        return c.merge ? launch_pressure_bc(c.P, G, G.g > 1, s) : hipSuccess;
    }
    // The check: FP64 residual of p = e, read-only.
    launch_close<false, false, false>(m, c, s);
    hipLaunchKernelGGL(k_res_final_wide, dim3(1), dim3(RF), 0, s, m.partials, m.np_merge,
                       c.res_to_div, c.target, m.done, m.last_res, m.cycles, 0);
    if (c.cycles > 1) {
        // The FP32 iteration continued on the same right-hand side.
        if ((err = smooth(m, c, 2 * c.fix_sweeps, true, s, false, true)) != hipSuccess) return err;
        launch_close<false, false, true>(m, c, s);
        hipLaunchKernelGGL(k_res_final_wide, dim3(1), dim3(RF), 0, s, m.partials, m.np_merge,
                           c.res_to_div, c.target, m.done, m.last_res, m.cycles, 1);
    }
    if (c.cycles > 2) {
        // Defect correction: P = e and r = b - A P (FP32) as the next
        // right-hand side, then cycles of a fresh correction added to P.
        hipLaunchKernelGGL(k_mark_merged, dim3(1), dim3(1), 0, s, m.done, m.merged);
        launch_close<true, true, true>(m, c, s);
        for (int cyc = 2; cyc < c.cycles; cyc++) {
            if ((err = smooth(m, c, 2 * c.fix_sweeps, true, s)) != hipSuccess) return err;
            if ((err = launch_from_split<float>(c.P.d, c.P.sy, c.P.sz, G, m.e_red, m.e_black, true,
                                                m.done, s)) != hipSuccess)
                return err;
            hipLaunchKernelGGL(k_residual, gcol, blk, 0, s, c.P, c.D.d, m.r_red, m.r_black, G,
                               c.h2, m.partials, m.done, 1);
            hipLaunchKernelGGL(k_res_final, dim3(1), dim3(RB), 0, s, m.partials, m.np,
                               c.res_to_div, c.target, m.done, m.last_res, m.cycles, 1);
        }
    }
    // P = e unless a defect-correction cycle merged the solution already.
    if (c.merge && (err = launch_from_split<float>(c.P.d, c.P.sy, c.P.sz, G, m.e_red, m.e_black,
                                                   false, m.merged, s)) != hipSuccess)
        return err;
    if ((err = hipGetLastError()) != hipSuccess) return err;
    // Final Neumann fill of the FP64 pressure (project.go:98); without merge
    // only where the defect correction wrote P.
    return c.merge ? launch_pressure_bc(c.P, G, G.g > 1, s)
                   : launch_pressure_bc(c.P, G, G.g > 1, s, m.merged, 1);
}

}  // namespace af
