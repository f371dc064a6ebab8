// af_batch.hip -- many engines of uneven sizes stepped together on one GPU.
//
// The reference steps one Engine per Step call (aeroflux.go:348-409).  Small
// grids leave most of an MI355X idle and pay a launch per stage each, so a
// batch keeps many engines ("slots") resident in one HBM pool and steps them
// together:
//   * every slot is New's arena (grid.go:79-134, same offsets) plus its
//     colour-split solve buffers, carved out of one pool allocation;
//   * each stage of the Strang palindrome is ONE persistent launch over a
//     work-list of tiles of every resident slot; the work-list is a prefix
//     table per stage kind, so membership changes rewrite a few kilobytes
//     and never the launch sequence;
//   * the whole step is captured once as a hipGraph and replayed;
//   * workgroups take contiguous ranges of the work-list per XCD (workgroup
//     b runs on XCD b % 8), so each XCD's L2 sees neighbouring tiles;
//   * the per-thread arithmetic is af_bodies.h, shared with the single
//     engine, so a slot's results are bit-identical to af_step's.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <deque>
#include <thread>
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <set>
#include <vector>

#include "../../include/aeroflux.h"
#include "af_bodies.h"
#include "af_host.h"
#include "af_sweep_lds.h"
#include "af_sweep_stage.h"
#include "af_ic.h"
#include "af_internal.h"
#include "af_snapshot.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int BX = body::BX;
constexpr int BY = 4;       // rows per workgroup of the 64x4 stages
// z-planes per thread of a colour pass: 8 FP64 (two chunks of 4 planes'
// loads in flight) or 16 FP32 (two chunks of 8).  Measured on the 96 uneven
// engines: FP32 items of 16 planes instead of 8 took the pass from 58.6 to
// 54.6 us (fewer items, less per-item set-up); FP64 items of 16 were 2%
// slower.  With the z-marching pass body, 16 FP64 / 32 FP32 planes were
// slower again (88.7 -> 90.7 us, 48.1 -> 51.9 us).
template <typename T>
constexpr int pass_planes()
{
    return sizeof(T) == 4 ? 16 : 8;
}
constexpr int CHUNK = 256;  // cells per item of the shell stages
constexpr int CZ = 4;       // z-planes per item of K_CELL (divergence, merge)
constexpr int XCDS = 8;
// Serving policy 2: interior cells one batched step advances by default.
constexpr int64_t kDefaultStepCells = (int64_t)32 << 20;

// Default solve group size: bytes of split arrays (both colours, solution
// and right-hand side) per group of slots whose solve runs together.  About
// the 256 MB last-level cache, so a group's passes reread it from there.
constexpr int64_t kDefaultSolveGroupBytes = (int64_t)224 << 20;

// Stage kinds: each has its own tile decomposition and prefix table.
enum Kind {
    K_VBC,    // velocity-BC shell cells, CHUNK per item
    K_SWEEP,  // advection tiles (64 x rows x z-per-thread) of the union box
    K_CELL,   // 256-point plane tiles of CZ interior z-planes
    K_GRAD,   // 256-point tiles of one plane, over NZ + 1 planes (measured: more
              // planes per item made the gradient slower; four with every load
              // in flight before the first wait gained nothing, and a cell
              // pair per thread over 4 planes, body::gradient_pair, was 4-8%
              // slower on the 96 engines)
    K_PASS,   // colour-pass tiles of the split layout (64 x 4 x pass_planes)
    K_PBC,    // pressure ghost-shell cells, CHUNK per item
    K_COL,    // 64 x 4 interior columns (z loop inside)
    K_ENG,    // one item per slot
    K_LDS,    // LDS-staged advection tiles (64 x rows x z-chunk) + wall blocks
    K_PAIR,   // 256-entry tiles of the split layout's planes (a cell pair per
              // thread), kMergeZ z-planes per item (merges, the fused close)
    K_COUNT
};

// Sweep shapes of the global gather: workgroups of BX x rows threads over
// flattened plane tiles, z positions per thread.
constexpr int kShapes = 8;
constexpr int kShapeRows[kShapes] = {1, 1, 2, 2, 4, 4, 8, 16};
constexpr int kShapeZ[kShapes] = {1, 2, 1, 2, 1, 2, 1, 1};

// LDS-staged sweep variants: rows per workgroup and z-planes per chunk.
// 0-3 stage one box of the three fields' union (af_sweep_lds.h: reach two
// cells along x and y, one along z); 4-7 stage each field's own box
// (af_sweep_stage.h).
constexpr int kLdsVariants = 8;

// A sweep plan: the gather shape for every slot, and the LDS-staged variant
// (-1: none) for the slots lds_slot accepts, as one integer.
inline int make_plan(int gather, int lds) { return gather + kShapes * (lds + 1); }
inline int plan_gather(int plan) { return plan % kShapes; }
inline int plan_lds(int plan) { return plan / kShapes - 1; }

// Slots whose x extent fills their 64-wide tiles to at least 3/4 (48, 64,
// 96, 128, 192, ...) take the LDS-staged sweep; narrower ones keep the
// flattened-plane gather, which leaves no lane idle.
inline bool lds_slot(const Geom &G)
{
    const int tiles = (G.NX + BX - 1) / BX;
    return 4 * G.NX >= 3 * BX * tiles;
}

template <int LBY, int ZC>
inline int64_t lds_items_t(const Geom &G)
{
    const int3 nb = slds::tiles<LBY, ZC>(G);
    return (int64_t)nb.x * nb.y * nb.z + slds::wall_blocks<LBY>(G);
}

inline int64_t lds_items(const Geom &G, int v)
{
    switch (v) {
    case 0: case 6: return lds_items_t<4, 16>(G);
    case 1: case 7: return lds_items_t<4, 32>(G);
    case 2: case 4: return lds_items_t<8, 16>(G);
    default: return lds_items_t<8, 32>(G);
    }
}

// Device descriptor of one slot.
struct BSlot {
    DField f[AF_FIELD_COUNT];  // P, U, V, W, Div, Tmp0, Tmp1, Tmp2 (unswapped)
    Geom G;
    double dt_half, invH, lam_half, h2, div_scale, grad_scale, res_to_div, target;
    void *s_red, *s_black;  // FP64: P split; mixed: FP32 correction e
    void *d_red, *d_black;  // FP64: Div split; mixed: FP32 right-hand side / residual
    double *partials;       // residual and reduction partials
    double *out;            // reduction result
};

// Kernel argument shared by every stage.
struct BCtl {
    const BSlot *slots;
    const int *tab;  // nact, stride, act[stride], first[K_COUNT][stride], groups
    int *flags, *done, *cycles;
    double *last;
    int grp = -1;  // solve group whose slots a launch covers (-1: every slot)
};

// Tiles of the plane-wise stages flatten each xy-plane over a workgroup's
// threads, so short rows (small grids) keep every lane busy: a tile is
// `threads` consecutive points of one plane (x fastest) for one z-block.
__host__ __device__ inline int64_t plane_tiles(int64_t plane, int threads)
{
    return (plane + threads - 1) / threads;
}

// K_PAIR items of a slot: 256-entry tiles of an NXH x NY split plane, for
// each group of body::kMergeZ planes.  The fused close writes one residual
// partial per item.
__host__ __device__ inline int64_t pair_items(const Geom &G)
{
    return plane_tiles((int64_t)((G.NX + 1) / 2) * G.NY, BX * BY) *
           ((G.NZ + body::kMergeZ - 1) / body::kMergeZ);
}

inline int64_t kind_items(int kind, const Geom &G, const DField &P, int plan, int precision)
{
    const int64_t nxh = (G.NX + 1) / 2;
    const int shape = plan_gather(plan), lds = plan_lds(plan);
    const bool staged = lds >= 0 && lds_slot(G);
    switch (kind) {
    case K_VBC: {
        int64_t cu, cv, cw;
        body::velocity_bc_counts(G, cu, cv, cw);
        return (cu + cv + cw + CHUNK - 1) / CHUNK;
    }
    case K_SWEEP: {
        if (staged) return 0;
        const int r = kShapeRows[shape], z = kShapeZ[shape];
        return plane_tiles((int64_t)(G.NX + 1) * (G.NY + 1), BX * r) * ((G.NZ + z) / z);
    }
    case K_CELL: return plane_tiles((int64_t)G.NX * G.NY, BX * BY) * ((G.NZ + CZ - 1) / CZ);
    case K_GRAD: return plane_tiles((int64_t)G.NX * G.NY, BX * BY) * (G.NZ + 1);
    case K_PASS: {
        const int pz = precision == AF_PRECISION_MIXED ? pass_planes<float>()
                                                       : pass_planes<double>();
        return plane_tiles(nxh * G.NY, BX * BY) * ((G.NZ + pz - 1) / pz);
    }
    case K_PBC:
        return (body::pressure_shell_count(P.nx, P.ny, P.nz, G.g) + CHUNK - 1) / CHUNK;
    case K_COL: return plane_tiles((int64_t)G.NX * G.NY, BX * BY);
    case K_LDS: return staged ? lds_items(G, lds) : 0;
    case K_PAIR: return pair_items(G);
    default: return 1;
    }
}

// q = j / n, r = j % n for 0 <= j < 2^24, n >= 1: a float estimate of the
// quotient is off by at most one, fixed by one correction step.
__device__ __forceinline__ void divmod_small(int j, int n, int &q, int &r)
{
    q = (int)((float)j * (1.0f / (float)n));
    r = j - q * n;
    if (r < 0) {
        q--;
        r += n;
    } else if (r >= n) {
        q++;
        r -= n;
    }
}

// Point (x, y) of tile `t` of a plane-wise stage: `row` points per row,
// `plane` points per plane, `threads` per tile; returns the z-block.
__device__ __forceinline__ int plane_point(int t, int row, int plane, int threads, int &x, int &y)
{
    const int nchunk = (plane + threads - 1) / threads;
    int zb, chunk;
    divmod_small(t, nchunk, zb, chunk);
    const int j = chunk * threads + (int)(threadIdx.y * blockDim.x + threadIdx.x);
    divmod_small(j, row, y, x);
    return zb;
}

// Persistent loop over the work-list of stage KIND: every active slot's
// items, or those of solve group c.grp's slots.  XCD x (= blockIdx.x % 8)
// owns the x-th eighth of the items, in order; its workgroups stride
// through that range.  f(slot, tile) runs with a workgroup-uniform slot and
// tile, so bodies may synchronise the workgroup.  gridDim.x % 8 == 0.
template <int KIND, typename F>
__device__ __forceinline__ void for_items(const BCtl &c, F f)
{
    const int *tab = c.tab;
    const int nact = tab[0], stride = tab[1];
    const int *act = tab + 2;
    const int *first = tab + 2 + stride * (1 + KIND);
    int a_lo = 0, a_hi = nact;
    if (c.grp >= 0) {
        // Groups: count, then count + 1 boundaries in act[] order.
        const int *gr = tab + 2 + stride * (1 + K_COUNT);
        if (c.grp >= gr[0]) return;
        a_lo = gr[1 + c.grp];
        a_hi = gr[2 + c.grp];
    }
    const int base = first[a_lo];
    const int total = first[a_hi] - base;
    const int lanes = (int)(gridDim.x / XCDS);
    const int xcd = (int)(blockIdx.x % XCDS);
    const int lo = base + (int)(((int64_t)total * xcd) / XCDS);
    const int hi = base + (int)(((int64_t)total * (xcd + 1)) / XCDS);
    int i = lo + (int)(blockIdx.x / XCDS);
    if (i >= hi) return;
    int a = a_lo, r = a_hi - 1;  // largest a with first[a] <= i
    while (a < r) {
        const int m = (a + r + 1) >> 1;
        if (first[m] <= i) a = m; else r = m - 1;
    }
    for (; i < hi; i += lanes) {
        while (first[a + 1] <= i) a++;
        f(act[a], i - first[a]);
    }
}

__device__ __forceinline__ const DField &velf(const BSlot &e, int sw, int c)
{
    return e.f[(sw ? AF_FIELD_TMP0 : AF_FIELD_U) + c];
}
__device__ __forceinline__ const DField &tmpf(const BSlot &e, int sw, int c)
{
    return e.f[(sw ? AF_FIELD_U : AF_FIELD_TMP0) + c];
}

// applyVelocityBC, stencil.go:246-315.
__global__ __launch_bounds__(256) void kb_vbc(BCtl c, int sw)
{
    for_items<K_VBC>(c, [&](int s, int t) {
        const BSlot &e = c.slots[s];
        int64_t cu, cv, cw;
        body::velocity_bc_counts(e.G, cu, cv, cw);
        body::velocity_bc_cell(velf(e, sw, 0), velf(e, sw, 1), velf(e, sw, 2), e.G,
                               (int64_t)t * CHUNK + threadIdx.x, cu, cv, cw);
    });
}

// advectDiffuseAxis, stencil.go:21-119.
template <int AX, int BYT, int ZR>
__global__ __launch_bounds__(BX *BYT) void kb_sweep(BCtl c, int sw)
{
    for_items<K_SWEEP>(c, [&](int s, int t) {
        const BSlot &e = c.slots[s];
        SweepArgs a;
        a.U = velf(e, sw, 0); a.V = velf(e, sw, 1); a.W = velf(e, sw, 2);
        a.T0 = tmpf(e, sw, 0); a.T1 = tmpf(e, sw, 1); a.T2 = tmpf(e, sw, 2);
        a.G = e.G;
        a.dt = e.dt_half;
        a.invH = e.invH;
        a.lam = e.lam_half;
        a.flag = c.flags + s;
        a.bc = 1;  // batched sweeps always apply the velocity BC to their output
        const int row = e.G.NX + 1;
        int x, y;
        const int zb = plane_point(t, row, row * (e.G.NY + 1), BX * BYT, x, y);
        const int g = e.G.g;
        body::sweep_column<AX, ZR>(a, g + x, g + y, g + zb * ZR);
    });
}

// advectDiffuseAxis, stencil.go:21-119, with the sources staged in LDS
// (af_sweep_lds.h) for the slots that take it: one workgroup per item of
// the K_LDS work-list, which numbers a slot's tiles x fastest, then y, then
// z-chunk, followed by its wall-column blocks.  XCD x (= blockIdx.x % 8)
// takes the x-th eighth of the list in order; workgroups past the list
// return.  The grid must hold at least 8 * ceil(items / 8) workgroups.
template <int AX, int LBY, int ZC>
__global__ __launch_bounds__(BX *LBY) void kb_sweep_lds(BCtl c, int sw)
{
    using Bx = slds::Box<AX, LBY, 2>;
    __shared__ double lds_mem[Bx::TOTAL];
    const int *tab = c.tab;
    const int nact = tab[0], stride = tab[1];
    const int *act = tab + 2;
    const int *first = tab + 2 + stride * (1 + K_LDS);
    const int total = first[nact];
    const int per = (total + XCDS - 1) / XCDS;
    const int k = (int)(blockIdx.x / XCDS);
    const int i = (int)(blockIdx.x % XCDS) * per + k;
    if (k >= per || i >= total) return;
    int a = 0, r = nact - 1;  // largest a with first[a] <= i
    while (a < r) {
        const int m = (a + r + 1) >> 1;
        if (first[m] <= i) a = m; else r = m - 1;
    }
    const int s = act[a], t = i - first[a];
    const BSlot &e = c.slots[s];
    SweepArgs sa;
    sa.U = velf(e, sw, 0); sa.V = velf(e, sw, 1); sa.W = velf(e, sw, 2);
    sa.T0 = tmpf(e, sw, 0); sa.T1 = tmpf(e, sw, 1); sa.T2 = tmpf(e, sw, 2);
    sa.G = e.G;
    sa.dt = e.dt_half;
    sa.invH = e.invH;
    sa.lam = e.lam_half;
    sa.flag = c.flags + s;
    sa.bc = 1;
    const int3 nb = slds::tiles<LBY, ZC>(e.G);
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if (t >= ntile) {
        slds::wall_face<AX, true>(sa, (t - ntile) * (BX * LBY) + threadIdx.y * BX + threadIdx.x,
                                  bad);
    } else {
        const int rest = t / nb.x;
        // Descriptors by run-time index come from the slot table in global
        // memory (a local copy indexed at run time would live in scratch).
        auto src_of = [&](int q) -> const DField & { return velf(e, sw, q); };
        int zs, zend;
        slds::zchunk(e.G, nb.z, rest / nb.y, zs, zend);
        slds::tile<AX, LBY, 2, true>(sa, src_of, (slds::lds_t *)lds_mem, t - rest * nb.x,
                                     rest % nb.y, zs, zend, bad);
    }
    if (bad) *body::gp(sa.flag) = 1;
}

// The same work-list with per-field staged boxes (af_sweep_stage.h).  Six
// waves per SIMD are asked for where the registers allow it (the 4-row Y
// sweep stages two pairs of its velocity field per thread).
template <int AX, int LBY, int ZC>
__global__ __launch_bounds__(BX *LBY) __attribute__((amdgpu_waves_per_eu(
    AX == 1 && LBY == 4 ? 4 : 6))) void kb_sweep_stg(
    BCtl c, int sw)
{
    using L = stg::Layout<AX, LBY>;
    __shared__ __attribute__((aligned(16))) double lds_mem[L::TOTAL];
    const int *tab = c.tab;
    const int nact = tab[0], stride = tab[1];
    const int *act = tab + 2;
    const int *first = tab + 2 + stride * (1 + K_LDS);
    const int total = first[nact];
    const int per = (total + XCDS - 1) / XCDS;
    const int k = (int)(blockIdx.x / XCDS);
    const int i = (int)(blockIdx.x % XCDS) * per + k;
    if (k >= per || i >= total) return;
    int a = 0, r = nact - 1;  // largest a with first[a] <= i
    while (a < r) {
        const int m = (a + r + 1) >> 1;
        if (first[m] <= i) a = m; else r = m - 1;
    }
    const int s = act[a], t = i - first[a];
    const BSlot &e = c.slots[s];
    SweepArgs sa;
    sa.U = velf(e, sw, 0); sa.V = velf(e, sw, 1); sa.W = velf(e, sw, 2);
    sa.T0 = tmpf(e, sw, 0); sa.T1 = tmpf(e, sw, 1); sa.T2 = tmpf(e, sw, 2);
    sa.G = e.G;
    sa.dt = e.dt_half;
    sa.invH = e.invH;
    sa.lam = e.lam_half;
    sa.flag = c.flags + s;
    sa.bc = 1;
    const int3 nb = slds::tiles<LBY, ZC>(e.G);
    const int ntile = nb.x * nb.y * nb.z;
    bool bad = false;
    if (t >= ntile) {
        slds::wall_face<AX, true>(sa, (t - ntile) * (BX * LBY) + threadIdx.y * BX + threadIdx.x,
                                  bad);
    } else {
        const int rest = t / nb.x;
        auto src_of = [&](int q) -> const DField & { return velf(e, sw, q); };
        int zs, zend;
        slds::zchunk(e.G, nb.z, rest / nb.y, zs, zend);
        stg::tile<AX, LBY, true>(sa, src_of, (slds::lds_t *)lds_mem, t - rest * nb.x,
                                 rest % nb.y, zs, zend, bad);
    }
    if (bad) *body::gp(sa.flag) = 1;
}

// computeDivergence, project.go:42-62, also writing the solve's split
// right-hand side: the FP64 divergence itself, or the mixed solve's FP32
// -(H^2 div).  Same values as a separate split pass over Div.  It also
// writes the solve's first red pass (the first cycle's, for the mixed
// solve): from zero, every neighbour and the mirrored own value are +0, so
// a red cell's value depends only on its right-hand side.  DIV false skips
// the store to Div (a Div nothing reads before the next projection's
// divergence rewrites it).
template <bool MIXED, bool DIV>
__global__ __launch_bounds__(BX *BY) void kb_div(BCtl c, int sw)
{
    for_items<K_CELL>(c, [&](int s, int t) {
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        int xi, yi;
        const int zb = plane_point(t, G.NX, G.NX * G.NY, BX * BY, xi, yi);
        // Plane by plane (measured on the 96 engines: issuing the CZ planes'
        // loads together, body::divergence_column, was 3% slower).
        for (int r = 0; r < CZ; r++) {
            const int zi = zb * CZ + r;
            if (zi >= G.NZ) break;
            double v;
            // (false: x or y outside the interior, the same on every plane)
            if (!body::divergence_cell(velf(e, sw, 0), velf(e, sw, 1), velf(e, sw, 2),
                                       e.f[AF_FIELD_DIV], G, e.div_scale, xi + G.g, yi + G.g,
                                       zi + G.g, v, DIV))
                break;
            const int i = (zi * G.NY + yi) * ((G.NX + 1) / 2) + (xi >> 1);
            const bool black = body::parity(xi, yi, zi, G.g);
            // The solve's first red pass from zero (s_red): relax(+0 + ... +
            // +0, rhs), the colour pass's own MODE 1 arithmetic.
            if (MIXED) {
                const float rf = (float)(-(e.h2 * v));
                body::gp(static_cast<float *>(black ? e.d_black : e.d_red))[i] = rf;
                if (!black) {
                    const float z = 0.0f;
                    body::gp(static_cast<float *>(e.s_red))[i] =
                        body::relax<float>(z + z + z + z + z + z, rf, 0.0);
                }
            } else {
                body::gp(static_cast<double *>(black ? e.d_black : e.d_red))[i] = v;
                if (!black) {
                    const double z = 0.0;
                    body::gp(static_cast<double *>(e.s_red))[i] =
                        body::relax<double>(z + z + z + z + z + z, v, e.h2);
                }
            }
        }
    });
}

// One red-black colour pass on the split layout (gaussSeidelSweep,
// project.go:103-130).  gated: skip slots whose mixed solve is done.
template <typename T, int MODE>
__global__ __launch_bounds__(BX *BY) void kb_pass(BCtl c, int color, int gated)
{
    for_items<K_PASS>(c, [&](int s, int t) {
        if (gated && c.done[s]) return;
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        SplitArgs<T> a;
        a.red = static_cast<T *>(e.s_red);
        a.black = static_cast<T *>(e.s_black);
        a.rhs_red = static_cast<const T *>(e.d_red);
        a.rhs_black = static_cast<const T *>(e.d_black);
        a.NX = G.NX; a.NY = G.NY; a.NZ = G.NZ; a.NXH = (G.NX + 1) / 2; a.g = G.g;
        a.h2 = e.h2;
        int k, yi;
        const int zb = plane_point(t, a.NXH, a.NXH * G.NY, BX * BY, k, yi);
        // FP64: two-plane chunks, software-pipelined (87 -> 86 us per pass on the
        // 96 engines; with four-plane chunks 94 us, 142 VGPRs).
        if constexpr (sizeof(T) == 8)
            body::split_pass_column_pipelined<T, MODE, pass_planes<T>(), 2>(a, color, k, yi,
                                                                           zb * pass_planes<T>());
        else
            body::split_pass_column<T, MODE, pass_planes<T>()>(a, color, k, yi,
                                                          zb * pass_planes<T>());
    });
}

// Split -> P interior (ADD: P += split).
template <typename TS, bool ADD>
__global__ __launch_bounds__(BX *BY) void kb_from_split(BCtl c, int gated)
{
    for_items<K_PAIR>(c, [&](int s, int t) {
        if (gated && c.done[s]) return;
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        const DField &P = e.f[AF_FIELD_P];
        const int NXH = (G.NX + 1) / 2;
        int k, yi;
        const int zb = plane_point(t, NXH, NXH * G.NY, BX * BY, k, yi);
        body::from_split_pairs<TS, ADD, body::kMergeZ>(
            P.d, P.sy, P.sz, G.g, G.NX, G.NY, G.NZ, NXH, static_cast<const TS *>(e.s_red),
            static_cast<const TS *>(e.s_black), k, yi, zb * body::kMergeZ);
    });
}

// The mixed solve's first-cycle close (body::merge_residual_pair): P takes
// the FP32 correction and the FP64 residual of that P goes to the split
// right-hand side, with one max|r| partial per item.  The same bits as
// kb_from_split<float, false> followed by kb_resid.
// SR: also store the residual as the next cycle's FP32 right-hand side (the
// start of a defect-correction cycle); else P = e and the max only.
template <bool SR>
__global__ __launch_bounds__(BX *BY) void kb_merge_resid(BCtl c, int gated)
{
    __shared__ double sh[BX * BY];
    for_items<K_PAIR>(c, [&](int s, int t) {
        if (gated && c.done[s]) return;
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        const int NXH = (G.NX + 1) / 2;
        int k, yi;
        const int zb = plane_point(t, NXH, NXH * G.NY, BX * BY, k, yi);
        const double m = body::merge_residual_pair<true, SR>(
            e.f[AF_FIELD_P], e.f[AF_FIELD_DIV].d, static_cast<const float *>(e.s_red),
            static_cast<const float *>(e.s_black), static_cast<float *>(e.d_red),
            static_cast<float *>(e.d_black), G, e.h2, k, yi, zb * body::kMergeZ);
        const double b = body::block_max256(m, sh);
        if (threadIdx.x == 0 && threadIdx.y == 0) body::gp(e.partials)[t] = b;
    });
}

// Final applyPressureBC of the solve, project.go:98 (outer layers zeroed for
// Ghost > 1, as after the solve's zeroing).
__global__ __launch_bounds__(256) void kb_pbc(BCtl c)
{
    for_items<K_PBC>(c, [&](int s, int t) {
        const BSlot &e = c.slots[s];
        const DField &P = e.f[AF_FIELD_P];
        const int64_t k = (int64_t)t * CHUNK + threadIdx.x;
        if (k >= body::pressure_shell_count(P.nx, P.ny, P.nz, e.G.g)) return;
        body::pressure_bc_cell(P, e.G, k, e.G.g > 1);
    });
}

// subtractGradient, project.go:172-205, with the velocity BC that follows it
// in project applied to the updated faces.
__global__ __launch_bounds__(BX *BY) void kb_grad(BCtl c, int sw)
{
    for_items<K_GRAD>(c, [&](int s, int t) {
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        int xi, yi;
        const int zi = plane_point(t, G.NX, G.NX * G.NY, BX * BY, xi, yi);
        body::gradient_cell<true>(velf(e, sw, 0), velf(e, sw, 1), velf(e, sw, 2),
                                  e.f[AF_FIELD_P], G, e.grad_scale, G.g + xi, G.g + yi, G.g + zi);
    });
}

// Mixed solve: FP64 residual with FP32 split output and per-column partials.
__global__ __launch_bounds__(BX *BY) void kb_resid(BCtl c, int gated)
{
    __shared__ double sh[BX * BY];
    for_items<K_COL>(c, [&](int s, int t) {
        if (gated && c.done[s]) return;
        const BSlot &e = c.slots[s];
        const Geom &G = e.G;
        int xi, yi;
        plane_point(t, G.NX, G.NX * G.NY, BX * BY, xi, yi);
        const double m = body::residual_column(e.f[AF_FIELD_P], e.f[AF_FIELD_DIV].d,
                                               static_cast<float *>(e.d_red),
                                               static_cast<float *>(e.d_black), G, e.h2, xi, yi);
        const double b = body::block_max256(m, sh);
        if (threadIdx.x == 0 && threadIdx.y == 0) body::gp(e.partials)[t] = b;
    });
}

// Mixed solve: fold a slot's partials (fixed order), record max|div| and
// raise the slot's done flag once its target is met.
// pairs: the partials are kb_merge_resid's (one per K_PAIR item), else
// kb_resid's (one per K_COL column tile).
__global__ __launch_bounds__(256) void kb_res_final(BCtl c, int gated, int pairs)
{
    __shared__ double sh[256];
    for_items<K_ENG>(c, [&](int s, int) {
        if (gated && c.done[s]) return;
        const BSlot &e = c.slots[s];
        const int np = pairs ? (int)pair_items(e.G)
                             : (e.G.NX * e.G.NY + BX * BY - 1) / (BX * BY);
        double v = 0;
        const auto part = body::gp((const double *)e.partials);
        for (int k = threadIdx.x; k < np; k += 256) v = part[k] > v ? part[k] : v;
        const double m = body::block_max256(v, sh);
        if (threadIdx.x == 0) {
            const double d = e.res_to_div * m;
            c.last[s] = d;
            c.cycles[s] = c.cycles[s] + 1;
            if (d <= e.target) c.done[s] = 1;
        }
    });
}

__global__ __launch_bounds__(256) void kb_mixed_reset(BCtl c, int checked)
{
    for_items<K_ENG>(c, [&](int s, int) {
        if (threadIdx.x == 0) {
            c.done[s] = 0;
            c.cycles[s] = checked ? 0 : 1;
        }
    });
}

__global__ __launch_bounds__(256) void kb_fill_int(int *p, int n, int v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ host side

int64_t round256(int64_t n) { return (n + 255) & ~(int64_t)255; }

struct SlotLayout {
    af_field f[AF_FIELD_COUNT];
    int64_t split_off[4];
    int64_t partials_off, out_off, total;
};

bool slot_layout(const af_config &c, int precision, SlotLayout *L)
{
    const int64_t arena = mac_layout(L->f, c.NX, c.NY, c.NZ, c.Ghost);
    int64_t off = round256(arena);
    const int64_t elem = precision == AF_PRECISION_MIXED ? 4 : 8;
    const int64_t n = c.NZ * c.NY * ((c.NX + 1) / 2) * elem;
    for (int k = 0; k < 4; k++) {
        L->split_off[k] = off;
        off += round256(n);
    }
    Geom G{(int)c.NX, (int)c.NY, (int)c.NZ, (int)c.Ghost};
    L->partials_off = off;
    const int64_t np = std::max<int64_t>(reduce_partials_needed(G), pair_items(G));
    off += round256(np * 8);
    L->out_off = off;
    off += 256;
    L->total = off;
    for (int i = 0; i < AF_FIELD_COUNT; i++)
        if (L->f[i].Count >= (int64_t)1 << 31) return false;
    return true;
}

// First-fit pool over one device allocation.
struct Pool {
    char *base = nullptr;
    int64_t cap = 0, used = 0;
    std::map<int64_t, int64_t> free_;  // offset -> length

    bool alloc(int64_t n, int64_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const int64_t rest = it->second - n;
            const int64_t at = it->first + n;
            free_.erase(it);
            if (rest > 0) free_[at] = rest;
            used += n;
            return true;
        }
        return false;
    }
    void release(int64_t off, int64_t n)
    {
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t largest() const
    {
        int64_t m = 0;
        for (const auto &kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }
};

struct HSlot {
    bool used = false;
    af_config cfg{};
    SlotLayout L{};
    int64_t off = 0;
    int64_t desc_step = -1;  // serving: batched step before which the slot's
                             // descriptor upload was enqueued
};

constexpr int kTabRing = 3;  // pinned work-list staging buffers

}  // namespace
}  // namespace af

using namespace af;

struct af_batch {
    af_batch_config cfg{};
    int device = 0;
    bool closed = true;
    hipStream_t stream = nullptr;
    Pool pool;
    std::vector<HSlot> slots;
    BSlot *d_slots = nullptr, *h_slots = nullptr;  // h_slots pinned
    int *d_tab = nullptr, *h_tab = nullptr;        // h_tab pinned, kTabRing tables
    int64_t tab_ints = 0;
    int *d_ctl = nullptr;    // flags | done | cycles, MaxSlots each
    double *d_last = nullptr;
    int *h_ints = nullptr;   // pinned, MaxSlots
    double *h_dbl = nullptr; // pinned, MaxSlots
    bool tab_dirty = true;
    int tab_shape = -1;
    hipGraphExec_t graph = nullptr;  // step graph for the current table's group count
    std::map<int, hipGraphExec_t> graphs;  // step graphs by solve-group count
    int ngroups = 1;                       // solve groups of the current table
    int shape = 4;  // sweep plan (make_plan): gather shape + LDS-staged variant
    int nwg_sweep = 2048, nwg_pass = 2048, nwg_other = 2048;
    int lds_grid = 8;      // workgroups of the LDS-staged sweep launches in the graph
    int lds_need = 8;      // workgroups the current table's K_LDS list needs
    bool tuned = false;
    int nact = 0;
    int64_t cells = 0;
    // Serving policy 2: slots the next batched steps advance (empty: every
    // resident slot).  Resident slots outside it keep their state.
    std::vector<char> step_mask;
    ICScratch ic;  // initial-condition generator scratch (lazy, outside the pool)
};

namespace {

int check_open(const af_batch *b)
{
    if (!b) return fail(AF_ERR_ARG, "aeroflux: null batch");
    if (b->closed) return fail(AF_ERR_CLOSED, "aeroflux: batch is closed");
    return AF_OK;
}

int check_slot(const af_batch *b, int32_t slot)
{
    if (int rc = check_open(b)) return rc;
    if (slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot].used)
        return fail(AF_ERR_ARG, "aeroflux: no engine in batch slot %d", (int)slot);
    return AF_OK;
}

BCtl ctl(const af_batch *b, int grp = -1)
{
    BCtl c;
    c.grp = grp;
    c.slots = b->d_slots;
    c.tab = b->d_tab;
    const int m = b->cfg.MaxSlots;
    c.flags = b->d_ctl;
    c.done = b->d_ctl + m;
    c.cycles = b->d_ctl + 2 * m;
    c.last = b->d_last;
    return c;
}

DField slot_field(const af_batch *b, const HSlot &h, int fid)
{
    const af_field &f = h.L.f[fid];
    DField d;
    d.d = reinterpret_cast<double *>(b->pool.base + h.off + f.Off);
    d.sy = (int)f.SY; d.sz = (int)f.SZ;
    d.nx = (int)f.NX; d.ny = (int)f.NY; d.nz = (int)f.NZ;
    return d;
}

Geom slot_geom(const HSlot &h)
{
    return Geom{(int)h.cfg.NX, (int)h.cfg.NY, (int)h.cfg.NZ, (int)h.cfg.Ghost};
}

// Rebuilds staging table `ring` for the current slots and sweep plan.
int build_table(af_batch *b, int plan, int ring = 0)
{
    const int stride = b->cfg.MaxSlots + 1;
    int *t = b->h_tab + (int64_t)ring * b->tab_ints;
    int *act = t + 2;
    int n = 0;
    const bool masked = !b->step_mask.empty();
    for (int s = 0; s < (int)b->slots.size(); s++)
        if (b->slots[s].used &&
            (!masked || ((size_t)s < b->step_mask.size() && b->step_mask[(size_t)s])))
            act[n++] = s;
    t[0] = n;
    t[1] = stride;
    for (int k = 0; k < K_COUNT; k++) {
        int *first = t + 2 + (int64_t)stride * (1 + k);
        int64_t acc = 0;
        first[0] = 0;
        for (int a = 0; a < n; a++) {
            const HSlot &h = b->slots[act[a]];
            acc += kind_items(k, slot_geom(h), slot_field(b, h, AF_FIELD_P), plan,
                              b->cfg.Precision);
            if (acc > INT_MAX - 8 * 1024)
                return fail(AF_ERR_CONFIG, "aeroflux: batch work-list exceeds 2^31 tiles");
            first[a + 1] = (int)acc;
        }
    }
    b->nact = n;
    // Solve groups: contiguous runs of act[] of about equal split-array
    // bytes, as many as the group byte budget needs (one when it is -1); a
    // group also closes before a slot that would take it past the budget.
    {
        int *gr = t + 2 + (int64_t)stride * (1 + K_COUNT);
        const int64_t elem = b->cfg.Precision == AF_PRECISION_MIXED ? 4 : 8;
        std::vector<int64_t> acc((size_t)n + 1, 0);
        for (int a = 0; a < n; a++) {
            const af_config &c = b->slots[act[a]].cfg;
            acc[(size_t)a + 1] = acc[(size_t)a] + 4 * c.NZ * c.NY * ((c.NX + 1) / 2) * elem;
        }
        const int64_t cap = b->cfg.SolveGroupBytes;
        int64_t want = cap < 0 || n == 0 ? 1 : (acc[(size_t)n] + cap - 1) / cap;
        want = want < 1 ? 1 : (want > n && n > 0 ? n : want);
        int g = 0, start = 0;
        gr[1] = 0;
        for (int a = 1; a < n; a++) {
            // A boundary before act[a] when a passes the next equal share, or
            // when act[a] would take the open group past the budget.
            const bool share = g + 1 < want && acc[(size_t)a] * want >= acc[(size_t)n] * (g + 1);
            const bool over = cap > 0 && acc[(size_t)a + 1] - acc[(size_t)start] > cap;
            if (share || over) {
                gr[2 + g++] = a;
                start = a;
            }
        }
        gr[2 + g] = n;
        gr[0] = g + 1;
        b->ngroups = g + 1;
    }
    const int lds_total = t[2 + (int64_t)stride * (1 + K_LDS) + n];
    b->lds_need = 8 * ((lds_total + 7) / 8);
    if (b->lds_need < 8) b->lds_need = 8;
    return AF_OK;
}

void drop_graph(af_batch *b);

// The LDS-staged sweep runs one workgroup per work item, so its grid is
// fixed in the step graph: when a table needs more, the graph is dropped
// (the caller has synchronised the stream) and recaptured with a grid 25%
// larger than needed, so a growing resident set recaptures rarely.
void fit_lds_grid(af_batch *b)
{
    if (b->lds_need <= b->lds_grid) return;
    const int64_t want = 8 * (((int64_t)b->lds_need + b->lds_need / 4 + 7) / 8);
    b->lds_grid = (int)(want < (int64_t)(INT_MAX & ~7) ? want : (int64_t)(INT_MAX & ~7));
    drop_graph(b);
}

// Uploads the table through the pinned staging copy, after any earlier
// copy from it has run.
int upload_table(af_batch *b, int plan)
{
    if (!b->tab_dirty && b->tab_shape == plan) return AF_OK;
    HIP_TRY(hipStreamSynchronize(b->stream), "upload batch work-list");
    if (int rc = build_table(b, plan)) return rc;
    fit_lds_grid(b);
    HIP_TRY(hipMemcpyAsync(b->d_tab, b->h_tab, (size_t)b->tab_ints * sizeof(int),
                           hipMemcpyHostToDevice, b->stream),
            "upload batch work-list");
    b->tab_dirty = false;
    b->tab_shape = plan;
    return AF_OK;
}

int sync(af_batch *b, const char *what)
{
    HIP_TRY(hipStreamSynchronize(b->stream), what);
    return AF_OK;
}

// ---- stage launchers ----

template <int BYT, int ZR>
hipError_t sweep_shape_launch(const af_batch *b, int nwg, int axis, int sw)
{
    const BCtl c = ctl(b);
    const dim3 grid(nwg), blk(BX, BYT);
    switch (axis) {
    case 0: hipLaunchKernelGGL((kb_sweep<0, BYT, ZR>), grid, blk, 0, b->stream, c, sw); break;
    case 1: hipLaunchKernelGGL((kb_sweep<1, BYT, ZR>), grid, blk, 0, b->stream, c, sw); break;
    default: hipLaunchKernelGGL((kb_sweep<2, BYT, ZR>), grid, blk, 0, b->stream, c, sw); break;
    }
    return hipGetLastError();
}

template <int LBY, int ZC, bool OWN>
hipError_t sweep_lds_launch(const af_batch *b, int axis, int sw)
{
    const BCtl c = ctl(b);
    const dim3 grid(b->lds_grid), blk(BX, LBY);
    if (OWN) {
        switch (axis) {
        case 0: hipLaunchKernelGGL((kb_sweep_stg<0, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
        case 1: hipLaunchKernelGGL((kb_sweep_stg<1, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
        default: hipLaunchKernelGGL((kb_sweep_stg<2, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
        }
        return hipGetLastError();
    }
    switch (axis) {
    case 0: hipLaunchKernelGGL((kb_sweep_lds<0, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
    case 1: hipLaunchKernelGGL((kb_sweep_lds<1, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
    default: hipLaunchKernelGGL((kb_sweep_lds<2, LBY, ZC>), grid, blk, 0, b->stream, c, sw); break;
    }
    return hipGetLastError();
}

// One sweep of every slot under `plan`: the gather over the K_SWEEP list,
// then (if the plan has a variant) the LDS-staged sweep over the K_LDS list.
hipError_t enq_sweep(const af_batch *b, int plan, int nwg, int axis, int sw)
{
    hipError_t err;
    switch (plan_gather(plan)) {
    case 0: err = sweep_shape_launch<1, 1>(b, nwg, axis, sw); break;
    case 1: err = sweep_shape_launch<1, 2>(b, nwg, axis, sw); break;
    case 2: err = sweep_shape_launch<2, 1>(b, nwg, axis, sw); break;
    case 3: err = sweep_shape_launch<2, 2>(b, nwg, axis, sw); break;
    case 4: err = sweep_shape_launch<4, 1>(b, nwg, axis, sw); break;
    case 5: err = sweep_shape_launch<4, 2>(b, nwg, axis, sw); break;
    case 6: err = sweep_shape_launch<8, 1>(b, nwg, axis, sw); break;
    default: err = sweep_shape_launch<16, 1>(b, nwg, axis, sw); break;
    }
    if (err != hipSuccess) return err;
    switch (plan_lds(plan)) {
    case -1: return hipSuccess;
    case 0: return sweep_lds_launch<4, 16, false>(b, axis, sw);
    case 1: return sweep_lds_launch<4, 32, false>(b, axis, sw);
    case 2: return sweep_lds_launch<8, 16, false>(b, axis, sw);
    case 3: return sweep_lds_launch<8, 32, false>(b, axis, sw);
    case 4: return sweep_lds_launch<8, 16, true>(b, axis, sw);
    case 5: return sweep_lds_launch<8, 32, true>(b, axis, sw);
    case 6: return sweep_lds_launch<4, 16, true>(b, axis, sw);
    default: return sweep_lds_launch<4, 32, true>(b, axis, sw);
    }
}

template <typename T>
hipError_t enq_pass(const af_batch *b, int nwg, int color, int mode, int gated, int grp = -1)
{
    const BCtl c = ctl(b, grp);
    const dim3 grid(nwg), blk(BX, BY);
    switch (mode) {
    case 0: hipLaunchKernelGGL((kb_pass<T, 0>), grid, blk, 0, b->stream, c, color, gated); break;
    case 1: hipLaunchKernelGGL((kb_pass<T, 1>), grid, blk, 0, b->stream, c, color, gated); break;
    default: hipLaunchKernelGGL((kb_pass<T, 2>), grid, blk, 0, b->stream, c, color, gated); break;
    }
    return hipGetLastError();
}

// Shell and per-slot stages have few items: a smaller persistent grid.
constexpr int kSmallGrid = 1024;

hipError_t enq_vbc(const af_batch *b, int sw)
{
    hipLaunchKernelGGL(kb_vbc, dim3(kSmallGrid), dim3(256), 0, b->stream, ctl(b), sw);
    return hipGetLastError();
}

hipError_t enq_div(const af_batch *b, int sw, bool store_div = true)
{
    const dim3 grid(b->nwg_other), blk(BX, BY);
    const bool mixed = b->cfg.Precision == AF_PRECISION_MIXED;
    if (mixed && store_div)
        hipLaunchKernelGGL((kb_div<true, true>), grid, blk, 0, b->stream, ctl(b), sw);
    else if (mixed)
        hipLaunchKernelGGL((kb_div<true, false>), grid, blk, 0, b->stream, ctl(b), sw);
    else if (store_div)
        hipLaunchKernelGGL((kb_div<false, true>), grid, blk, 0, b->stream, ctl(b), sw);
    else
        hipLaunchKernelGGL((kb_div<false, false>), grid, blk, 0, b->stream, ctl(b), sw);
    return hipGetLastError();
}

hipError_t enq_grad(const af_batch *b, int sw)
{
    hipLaunchKernelGGL(kb_grad, dim3(b->nwg_other), dim3(BX, BY), 0, b->stream, ctl(b), sw);
    return hipGetLastError();
}

// solvePressureRedBlack, project.go:73-100, on every slot: FP64 colour-split
// passes (the first from P == 0), or the mixed-precision defect correction.
// Slots' solves are independent, so the solve runs group by group (the
// table's solve groups, `ngroups` of them): all of one group's passes, then
// the next group's, and a group's split arrays stay in the last-level cache
// from one pass to the next instead of being streamed from HBM for every pass.
//
// checked: the mixed solve iterates to each slot's MixedTarget, as the
// single engine's checked solve does (af_mixed.hip): the first cycle's
// close checks the FP64 residual of p = e; a slot that misses its target
// continues its FP32 iteration by MixedFixSweeps pairs and checks again,
// then runs FP64 defect-correction cycles.  Each slot takes the engine's
// operations in its order, so its values are the engine's.
hipError_t enq_solve(const af_batch *b, int ngroups, bool checked)
{
    const BCtl c = ctl(b);
    const dim3 gp(b->nwg_other), blk(BX, BY);
    hipError_t err = hipSuccess;
    auto mode = [](int p) { return p == 0 ? 1 : (p == 1 ? 2 : 0); };
    const bool mixed = b->cfg.Precision == AF_PRECISION_MIXED;
    checked = checked && mixed;
    if (mixed)
        hipLaunchKernelGGL(kb_mixed_reset, dim3(kSmallGrid), dim3(256), 0, b->stream, c,
                           checked ? 1 : 0);
    for (int g = 0; g < ngroups && err == hipSuccess; g++) {
        const BCtl cg = ctl(b, ngroups > 1 ? g : -1);
        const int grp = cg.grp;
        // The divergence already wrote the first red pass (kb_div): the
        // passes start at the first black one.
        if (!mixed) {
            const int passes = (int)(2 * b->cfg.GaussSeidelIters);
            for (int p = 1; p < passes && err == hipSuccess; p++)
                err = enq_pass<double>(b, b->nwg_pass, p & 1, mode(p), 0, grp);
            if (err != hipSuccess) return err;
            hipLaunchKernelGGL((kb_from_split<double, false>), gp, blk, 0, b->stream, cg, 0);
            continue;
        }
        for (int p = 1; p < 2 * b->cfg.MixedSweeps && err == hipSuccess; p++)
            err = enq_pass<float>(b, b->nwg_pass, p & 1, mode(p), 0, grp);
        if (err != hipSuccess) return err;
        if (!checked) {
            hipLaunchKernelGGL((kb_from_split<float, false>), gp, blk, 0, b->stream, cg, 0);
            continue;
        }
        hipLaunchKernelGGL((kb_merge_resid<false>), gp, blk, 0, b->stream, cg, 0);
        hipLaunchKernelGGL(kb_res_final, dim3(kSmallGrid), dim3(256), 0, b->stream, cg, 0, 1);
    }
    if (err != hipSuccess) return err;
    // The later cycles (a few passes, only for slots whose target is not met
    // yet) run over every slot at once: each slot still takes the same
    // operations in the same order.  First the FP32 iteration continued.
    if (checked && b->cfg.MixedMaxCycles > 1) {
        for (int p = 0; p < 2 * b->cfg.MixedFixSweeps && err == hipSuccess; p++)
            err = enq_pass<float>(b, b->nwg_pass, p & 1, 0, 1);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL((kb_merge_resid<false>), gp, blk, 0, b->stream, c, 1);
        hipLaunchKernelGGL(kb_res_final, dim3(kSmallGrid), dim3(256), 0, b->stream, c, 1, 1);
    }
    // Then defect correction: the residual of P = e as the right-hand side.
    if (checked && b->cfg.MixedMaxCycles > 2)
        hipLaunchKernelGGL((kb_merge_resid<true>), gp, blk, 0, b->stream, c, 1);
    for (int cyc = 2; checked && cyc < b->cfg.MixedMaxCycles; cyc++) {
        for (int p = 0; p < 2 * b->cfg.MixedFixSweeps && err == hipSuccess; p++)
            err = enq_pass<float>(b, b->nwg_pass, p & 1, mode(p), 1);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL((kb_from_split<float, true>), gp, blk, 0, b->stream, c, 1);
        hipLaunchKernelGGL(kb_resid, gp, blk, 0, b->stream, c, 1);
        hipLaunchKernelGGL(kb_res_final, dim3(kSmallGrid), dim3(256), 0, b->stream, c, 1, 0);
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    hipLaunchKernelGGL(kb_pbc, dim3(kSmallGrid), dim3(256), 0, b->stream, c);
    return hipGetLastError();
}

// project, project.go:12-34.  The velocity already holds the BC the
// projection starts with (the sweep before it applied it), and the gradient
// subtraction applies the closing one.  div_dead (the step's first
// projection): nothing reads Div before the second projection's divergence
// rewrites it, so Div is not stored, and a checked mixed solve runs its one
// cycle unchecked there (the second projection, whose divergence the step
// leaves, is checked), as the single engine's does.  (The single engine's first projection also
// leaves the solution in the split arrays for its gradient; the batched
// gradient reading the split arrays, one position per work item, measured
// 839 us (FP64) and 1008 us (FP32) against 585 us for kb_grad on the 96
// engines, more than the merge and pressure BC it saves.)
hipError_t enq_project(const af_batch *b, int sw, int ngroups, bool div_dead = false)
{
    hipError_t err = enq_div(b, sw, !div_dead);
    if (err == hipSuccess) err = enq_solve(b, ngroups, b->cfg.MixedChecked != 0 && !div_dead);
    if (err == hipSuccess) err = enq_grad(b, sw);  // applies the closing velocity BC
    return err;
}

// Step, aeroflux.go:348-409, for every slot.  sw tracks the commitVelocity
// swaps (0: velocity in U/V/W); the six swaps return it to 0.
hipError_t enq_step(const af_batch *b, int ngroups)
{
    static const int fwd[3] = {AF_AXIS_X, AF_AXIS_Y, AF_AXIS_Z};
    static const int rev[3] = {AF_AXIS_Z, AF_AXIS_Y, AF_AXIS_X};
    // Batched sweeps and gradient subtractions apply the velocity BC to
    // their own output, so only the step's first BC runs as a launch.
    hipError_t err = enq_vbc(b, 0);
    int sw = 0;
    for (int k = 0; k < 3 && err == hipSuccess; k++) {
        err = enq_sweep(b, b->shape, b->nwg_sweep, fwd[k], sw);
        sw ^= 1;
    }
    if (err == hipSuccess) err = enq_project(b, sw, ngroups, true);
    for (int k = 0; k < 3 && err == hipSuccess; k++) {
        err = enq_sweep(b, b->shape, b->nwg_sweep, rev[k], sw);
        sw ^= 1;
    }
    if (err == hipSuccess) err = enq_project(b, sw, ngroups);
    return err;
}

void drop_graph(af_batch *b)
{
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    b->graph = nullptr;
}

// The step graph for the current table's solve-group count.  Graphs read
// the work-list table at run time, so one graph per group count serves
// every table with that count; they are kept until the plan or grids change.
int get_graph(af_batch *b)
{
    auto it = b->graphs.find(b->ngroups);
    if (it != b->graphs.end()) {
        b->graph = it->second;
        return AF_OK;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamSynchronize(b->stream), "batch step capture");
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
            "begin batch step capture");
    hipError_t err = enq_step(b, b->ngroups);
    hipError_t err2 = hipStreamEndCapture(b->stream, &g);
    if (err != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(err, "capture batch step");
    }
    if (err2 != hipSuccess) return hip_fail(err2, "end batch step capture");
    hipGraphExec_t x = nullptr;
    hipError_t err3 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err3 != hipSuccess) return hip_fail(err3, "instantiate batch step graph");
    b->graphs[b->ngroups] = x;
    b->graph = x;
    return AF_OK;
}

// Times `reps` launches of `one` (after one warm-up) with events.
template <typename F>
int time_launches(af_batch *b, int reps, F one, double *ms_out)
{
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "create event");
    HIP_TRY(hipEventCreate(&e1), "create event");
    hipError_t err = one();
    if (err == hipSuccess) err = hipEventRecord(e0, b->stream);
    for (int r = 0; r < reps && err == hipSuccess; r++) err = one();
    if (err == hipSuccess) err = hipEventRecord(e1, b->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return hip_fail(err, "time batched stage");
    *ms_out = (double)ms / reps;
    return AF_OK;
}

hipError_t enq_timed_pass(const af_batch *b, int nwg, int grp = -1)
{
    if (b->cfg.Precision == AF_PRECISION_MIXED) return enq_pass<float>(b, nwg, 0, 0, 0, grp);
    return enq_pass<double>(b, nwg, 0, 0, 0, grp);
}

// One colour pass over every slot as the step runs it, group by group with
// each group's passes back to back: the sum over groups of one pass's
// average time.
int time_grouped_pass(af_batch *b, int nwg, int reps, double *ms_out)
{
    if (b->ngroups <= 1)
        return time_launches(b, reps, [&] { return enq_timed_pass(b, nwg); }, ms_out);
    double sum = 0;
    for (int g = 0; g < b->ngroups; g++) {
        double ms = 0;
        if (int rc = time_launches(b, reps, [&] { return enq_timed_pass(b, nwg, g); }, &ms))
            return rc;
        sum += ms;
    }
    *ms_out = sum;
    return AF_OK;
}

// Autotuner: the gather shape (timed on X, Y and Z sweeps with every slot
// gathering), then the LDS-staged variant for the slots that can take it
// (or none), then persistent workgroup counts for the gather sweeps and the
// colour passes, each timed on batched launches over the resident slots.  A
// sweep writes only the face cells of the Tmp fields, and the next step's
// first sweep rewrites every one of them before anything reads them; colour
// passes touch only solve buffers that every solve rewrites before reading.
// Results therefore do not depend on tuning.
int tune(af_batch *b)
{
    if (b->nact == 0) return AF_OK;
    static const int kWg[] = {1024, 2048, 4096, 8192};
    const int reps = 3;
    const int wg0 = b->cfg.Workgroups > 0 ? b->cfg.Workgroups : 2048;
    auto time_plan = [&](int plan, int nwg, double *ms) -> int {
        if (int rc = upload_table(b, plan)) return rc;
        *ms = 0;
        for (int ax = 0; ax < 3; ax++) {
            double t;
            if (int rc = time_launches(b, reps, [&] { return enq_sweep(b, plan, nwg, ax, 0); },
                                       &t))
                return rc;
            *ms += t;
        }
        return AF_OK;
    };
    int gather = b->cfg.SweepShape > 0 ? b->cfg.SweepShape - 1 : plan_gather(b->shape);
    double best = 1e30;
    if (b->cfg.SweepShape <= 0) {
        for (int s = 0; s < kShapes; s++) {
            double ms;
            if (int rc = time_plan(make_plan(s, -1), wg0, &ms)) return rc;
            if (ms < best) {
                best = ms;
                gather = s;
            }
        }
    }
    int lds = b->cfg.SweepStaged > 0 ? b->cfg.SweepStaged - 1 : -1;
    if (b->cfg.SweepStaged == 0) {
        if (best >= 1e30)
            if (int rc = time_plan(make_plan(gather, -1), wg0, &best)) return rc;
        for (int v = 0; v < kLdsVariants; v++) {
            double ms;
            if (int rc = time_plan(make_plan(gather, v), wg0, &ms)) return rc;
            if (ms < best) {
                best = ms;
                lds = v;
            }
        }
    }
    b->shape = make_plan(gather, lds);
    if (int rc = upload_table(b, b->shape)) return rc;
    b->nwg_sweep = b->nwg_pass = b->nwg_other = wg0;
    if (b->cfg.Workgroups <= 0) {
        double bs = 1e30, bp = 1e30;
        for (int w : kWg) {
            double ms;
            if (int rc = time_launches(b, reps, [&] { return enq_sweep(b, b->shape, w, 0, 0); },
                                       &ms))
                return rc;
            if (ms < bs) {
                bs = ms;
                b->nwg_sweep = w;
            }
            if (int rc = time_grouped_pass(b, w, reps, &ms)) return rc;
            if (ms < bp) {
                bp = ms;
                b->nwg_pass = w;
            }
        }
        b->nwg_other = b->nwg_pass;
    }
    drop_graph(b);
    b->tuned = true;
    return sync(b, "tune batch");
}

void release(af_batch *b)
{
    if (b->closed) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    b->ic.release();
    if (b->pool.base) (void)hipFree(b->pool.base);
    if (b->d_slots) (void)hipFree(b->d_slots);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_ctl) (void)hipFree(b->d_ctl);
    if (b->d_last) (void)hipFree(b->d_last);
    if (b->h_slots) (void)hipHostFree(b->h_slots);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_ints) (void)hipHostFree(b->h_ints);
    if (b->h_dbl) (void)hipHostFree(b->h_dbl);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->pool = Pool();
    b->d_slots = b->h_slots = nullptr;
    b->d_tab = b->h_tab = nullptr;
    b->d_ctl = nullptr;
    b->d_last = nullptr;
    b->h_ints = nullptr;
    b->h_dbl = nullptr;
    b->stream = nullptr;
    b->slots.clear();
    b->nact = 0;
    b->cells = 0;
    b->closed = true;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

int64_t af_batch_slot_bytes(const af_config *cfg, int32_t precision)
{
    if (!cfg) return 0;
    af_config c = *cfg;
    if (af_config_normalize(&c) != AF_OK) return 0;
    SlotLayout L;
    if (!slot_layout(c, precision, &L)) return 0;
    return L.total;
}

int af_batch_new(const af_batch_config *cfg_in, af_batch **out)
{
    if (!cfg_in || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *out = nullptr;
    af_batch_config c = *cfg_in;
    if (c.Precision != AF_PRECISION_FP64 && c.Precision != AF_PRECISION_MIXED)
        return fail(AF_ERR_CONFIG, "aeroflux: unknown precision mode %d", (int)c.Precision);
    if (c.MaxSlots == 0) c.MaxSlots = 1024;
    if (c.MemoryBudget == 0) c.MemoryBudget = (int64_t)1 << 30;
    if (c.GaussSeidelIters == 0) c.GaussSeidelIters = 24;
    if (c.MixedMaxCycles == 0) c.MixedMaxCycles = 4;
    if (c.MixedSweeps == 0) c.MixedSweeps = (int32_t)c.GaussSeidelIters;
    if (c.MixedFixSweeps == 0) c.MixedFixSweeps = 2;
    if (c.MaxSlots < 1 || c.MaxSlots > (1 << 20) || c.MemoryBudget < 0 ||
        c.GaussSeidelIters < 1 || c.MixedMaxCycles < 1 || c.MixedSweeps < 1 ||
        c.MixedFixSweeps < 1)
        return fail(AF_ERR_CONFIG, "aeroflux: bad batch capacity, budget or solve schedule");
    if (c.SweepShape < 0 || c.SweepShape > kShapes)
        return fail(AF_ERR_CONFIG, "aeroflux: sweep shape %d out of range 1..%d",
                    (int)c.SweepShape, kShapes);
    if (c.Workgroups < 0 || c.Workgroups % XCDS != 0)
        return fail(AF_ERR_CONFIG, "aeroflux: persistent workgroups must be a multiple of 8");
    if (c.SweepStaged < -1 || c.SweepStaged > kLdsVariants)
        return fail(AF_ERR_CONFIG, "aeroflux: staged sweep variant %d out of range -1..%d",
                    (int)c.SweepStaged, kLdsVariants);
    if (c.PreemptRatio == 0) c.PreemptRatio = 4.0;
    if (c.SnapshotBudget == 0) c.SnapshotBudget = (int64_t)32 << 30;
    if (c.StepCellBudget == 0) c.StepCellBudget = kDefaultStepCells;
    if (c.ServePolicy < 0 || c.ServePolicy > 3 || !(c.PreemptRatio >= 1.0) ||
        std::isinf(c.PreemptRatio) || c.SnapshotBudget < 0 || c.StepCellBudget < 1)
        return fail(AF_ERR_CONFIG,
                    "aeroflux: serve policy must be 0, 1, 2 or 3, PreemptRatio >= 1, "
                    "SnapshotBudget >= 0, StepCellBudget >= 1");
    if (c.SolveGroupBytes == 0) c.SolveGroupBytes = kDefaultSolveGroupBytes;
    if (c.SolveGroupBytes < -1)
        return fail(AF_ERR_CONFIG, "aeroflux: SolveGroupBytes must be -1, 0 or positive");
    if (c.StepMaxWait < 0)
        return fail(AF_ERR_CONFIG, "aeroflux: StepMaxWait must be 0 (no bound) or positive");

    af_batch *b = new (std::nothrow) af_batch();
    if (!b) return fail(AF_ERR_OOM, "aeroflux: out of host memory");
    b->cfg = c;
    b->closed = false;
    auto bail = [&](hipError_t er, const char *what, int code) {
        int rc = code == AF_ERR_OOM ? fail(AF_ERR_OOM, "aeroflux: %s: %s", what,
                                           hipGetErrorString(er))
                                    : hip_fail(er, what);
        release(b);
        delete b;
        return rc;
    };
    hipError_t err = hipGetDevice(&b->device);
    if (err != hipSuccess) return bail(err, "no HIP device", AF_ERR_HIP);
    const int m = c.MaxSlots;
    b->slots.resize(m);
    b->tab_ints = 2 + (int64_t)(m + 1) * (1 + K_COUNT) + (m + 2);
    b->shape = make_plan(c.SweepShape > 0 ? c.SweepShape - 1 : 4,
                         c.SweepStaged > 0 ? c.SweepStaged - 1 : -1);
    if (c.Workgroups > 0) b->nwg_sweep = b->nwg_pass = b->nwg_other = c.Workgroups;
    if ((err = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(err, "create stream", AF_ERR_HIP);
    if (c.MemoryBudget > 0) {
        if ((err = hipMalloc(&b->pool.base, (size_t)c.MemoryBudget)) != hipSuccess)
            return bail(err, "allocate batch pool", AF_ERR_OOM);
        b->pool.cap = c.MemoryBudget;
        b->pool.free_[0] = c.MemoryBudget;
    }
    if ((err = hipMalloc(&b->d_slots, sizeof(BSlot) * (size_t)m)) != hipSuccess ||
        (err = hipMalloc(&b->d_tab, sizeof(int) * (size_t)b->tab_ints)) != hipSuccess ||
        (err = hipMalloc(&b->d_ctl, sizeof(int) * 3 * (size_t)m)) != hipSuccess ||
        (err = hipMalloc(&b->d_last, sizeof(double) * (size_t)m)) != hipSuccess)
        return bail(err, "allocate batch tables", AF_ERR_OOM);
    if ((err = hipHostMalloc(&b->h_slots, sizeof(BSlot) * (size_t)m, hipHostMallocDefault)) !=
            hipSuccess ||
        (err = hipHostMalloc(&b->h_tab, sizeof(int) * (size_t)b->tab_ints * kTabRing,
                             hipHostMallocDefault)) != hipSuccess ||
        (err = hipHostMalloc(&b->h_ints, sizeof(int) * (size_t)m, hipHostMallocDefault)) !=
            hipSuccess ||
        (err = hipHostMalloc(&b->h_dbl, sizeof(double) * (size_t)m, hipHostMallocDefault)) !=
            hipSuccess)
        return bail(err, "allocate pinned batch tables", AF_ERR_OOM);
    const int nctl = 3 * m;
    hipLaunchKernelGGL(kb_fill_int, dim3((nctl + 255) / 256), dim3(256), 0, b->stream, b->d_ctl,
                       nctl, 0);
    if ((err = hipGetLastError()) != hipSuccess) return bail(err, "clear batch flags", AF_ERR_HIP);
    if ((err = launch_zero(b->d_last, m, b->stream)) != hipSuccess)
        return bail(err, "clear batch stats", AF_ERR_HIP);
    if (int rc = build_table(b, b->shape)) {
        release(b);
        delete b;
        return rc;
    }
    b->tab_dirty = true;
    if ((err = hipStreamSynchronize(b->stream)) != hipSuccess)
        return bail(err, "initialise batch", AF_ERR_HIP);
    *out = b;
    return AF_OK;
}

int af_batch_close(af_batch *b)
{
    if (!b) return fail(AF_ERR_ARG, "aeroflux: null batch");
    release(b);
    return AF_OK;
}

void af_batch_free(af_batch *b)
{
    if (!b) return;
    release(b);
    delete b;
}

int af_batch_get_info(const af_batch *b, af_batch_info *out)
{
    if (!b || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    af_batch_info i{};
    for (const HSlot &h : b->slots) i.Slots += h.used ? 1 : 0;
    i.SweepShape = plan_gather(b->shape);
    i.SweepStaged = plan_lds(b->shape);
    if (i.SweepStaged >= 0)
        for (const HSlot &h : b->slots) i.StagedSlots += h.used && lds_slot(slot_geom(h)) ? 1 : 0;
    i.SweepWorkgroups = b->nwg_sweep;
    i.PassWorkgroups = b->nwg_pass;
    i.OtherWorkgroups = b->nwg_other;
    i.Tuned = b->tuned ? 1 : 0;
    i.Cells = b->cells;
    i.BytesInUse = b->pool.used;
    i.BytesLargestFree = b->pool.largest();
    i.MemoryBudget = b->cfg.MemoryBudget;
    i.SolveGroups = b->ngroups;
    *out = i;
    return AF_OK;
}

// Checks a normalised engine config against the batch's precision and
// solve schedule.
static int check_cfg(const af_batch *b, const af_config &c)
{
    // UseSparseSolver is accepted and the red-black solve runs, as for one
    // engine (New never reads the flag; project.go:19 falls back off Darwin,
    // sparse_fallback.go:20-24).
    const af_batch_config &bc = b->cfg;
    if (c.Precision != bc.Precision)
        return fail(AF_ERR_CONFIG, "aeroflux: engine precision %d differs from the batch's %d",
                    (int)c.Precision, (int)bc.Precision);
    if (c.GaussSeidelIters != bc.GaussSeidelIters)
        return fail(AF_ERR_CONFIG,
                    "aeroflux: engine runs %ld sweep pairs, the batch's schedule %ld",
                    (long)c.GaussSeidelIters, (long)bc.GaussSeidelIters);
    if (c.Precision == AF_PRECISION_MIXED) {
        if (c.MixedSweeps != bc.MixedSweeps || c.MixedFixSweeps != bc.MixedFixSweeps ||
            c.MixedMaxCycles != bc.MixedMaxCycles)
            return fail(AF_ERR_CONFIG,
                        "aeroflux: engine mixed-precision schedule differs from the batch's");
        if (bc.MixedChecked ? !(c.MixedTarget > 0) : c.MixedTarget != 0)
            return fail(AF_ERR_CONFIG,
                        bc.MixedChecked
                            ? "aeroflux: a checked mixed-precision batch needs MixedTarget > 0"
                            : "aeroflux: a one-cycle mixed-precision batch needs MixedTarget == 0");
    }
    return AF_OK;
}

static void free_slot(af_batch *b, int slot)
{
    HSlot &h = b->slots[slot];
    b->pool.release(h.off, h.L.total);
    b->cells -= h.cfg.NX * h.cfg.NY * h.cfg.NZ;
    h.used = false;
    b->tab_dirty = true;
}

// Takes a free slot id and pool block for a checked config and enqueues its
// zeroing (MAC.Zero, grid.go:158-163), descriptor upload and flag resets on
// the batch stream.  AF_ERR_OOM, with nothing changed, if either is missing.
// The caller must know that no earlier upload from the slot's pinned
// descriptor is still pending.
static int pick_slot(const af_batch *b)
{
    for (int s = 0; s < (int)b->slots.size(); s++)
        if (!b->slots[s].used) return s;
    return -1;
}

static int place_slot(af_batch *b, const af_config &c, const SlotLayout &L, int *slot_out)
{
    const int slot = pick_slot(b);
    if (slot < 0) return fail(AF_ERR_OOM, "aeroflux: batch has no free slot");
    int64_t off;
    if (!b->pool.alloc(L.total, &off))
        return fail(AF_ERR_OOM,
                    "aeroflux: batch pool cannot place %ld bytes (%ld of %ld in use, largest "
                    "free block %ld)",
                    (long)L.total, (long)b->pool.used, (long)b->pool.cap,
                    (long)b->pool.largest());
    HSlot &h = b->slots[slot];
    h.used = true;
    h.cfg = c;
    h.L = L;
    h.off = off;
    char *base = b->pool.base + off;
    BSlot &d = b->h_slots[slot];
    for (int i = 0; i < AF_FIELD_COUNT; i++) d.f[i] = slot_field(b, h, i);
    d.G = slot_geom(h);
    // The same expressions, in the same order, as the single engine.
    d.dt_half = c.DT * 0.5;
    d.invH = 1.0 / c.H;
    d.lam_half = c.Viscosity * d.dt_half * d.invH * d.invH;
    d.h2 = c.H * c.H;
    d.div_scale = c.Density / (c.DT * c.H);
    d.grad_scale = c.DT / (c.Density * c.H);
    d.res_to_div = c.DT / (c.Density * d.h2);
    d.target = c.MixedTarget;
    d.s_red = base + L.split_off[0];
    d.s_black = base + L.split_off[1];
    d.d_red = base + L.split_off[2];
    d.d_black = base + L.split_off[3];
    d.partials = reinterpret_cast<double *>(base + L.partials_off);
    d.out = reinterpret_cast<double *>(base + L.out_off);
    b->cells += c.NX * c.NY * c.NZ;
    b->tab_dirty = true;
    hipError_t err = launch_zero(reinterpret_cast<double *>(base), L.total / 8, b->stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(b->d_slots + slot, b->h_slots + slot, sizeof(BSlot),
                             hipMemcpyHostToDevice, b->stream);
    if (err == hipSuccess) {
        int *ctl3[3] = {b->d_ctl + slot, b->d_ctl + b->cfg.MaxSlots + slot,
                        b->d_ctl + 2 * b->cfg.MaxSlots + slot};
        for (int *p : ctl3)
            hipLaunchKernelGGL(kb_fill_int, dim3(1), dim3(256), 0, b->stream, p, 1, 0);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = launch_zero(b->d_last + slot, 1, b->stream);
    if (err != hipSuccess) {
        free_slot(b, slot);
        return hip_fail(err, "add engine to batch");
    }
    *slot_out = slot;
    return AF_OK;
}

int af_batch_add(af_batch *b, const af_config *cfg_in, int32_t *slot_out)
{
    if (int rc = check_open(b)) return rc;
    if (!cfg_in || !slot_out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    af_config c = *cfg_in;
    if (int rc = af_config_normalize(&c)) return rc;
    if (int rc = check_cfg(b, c)) return rc;
    SlotLayout L;
    if (!slot_layout(c, c.Precision, &L))
        return fail(AF_ERR_CONFIG, "aeroflux: a %ldx%ldx%ld grid exceeds 2^31 elements per field",
                    (long)c.NX, (long)c.NY, (long)c.NZ);
    HIP_TRY(hipSetDevice(b->device), "set device");
    // Every earlier upload from the pinned descriptors has run once the
    // stream is idle.
    HIP_TRY(hipStreamSynchronize(b->stream), "add engine to batch");
    int slot;
    if (int rc = place_slot(b, c, L, &slot)) return rc;
    hipError_t err = hipStreamSynchronize(b->stream);
    if (err != hipSuccess) {
        free_slot(b, slot);
        return hip_fail(err, "add engine to batch");
    }
    *slot_out = slot;
    return AF_OK;
}

int af_batch_remove(af_batch *b, int32_t slot)
{
    if (int rc = check_slot(b, slot)) return rc;
    HIP_TRY(hipSetDevice(b->device), "set device");
    HIP_TRY(hipStreamSynchronize(b->stream), "remove engine from batch");
    free_slot(b, slot);
    return AF_OK;
}

int af_batch_tune(af_batch *b)
{
    if (int rc = check_open(b)) return rc;
    HIP_TRY(hipSetDevice(b->device), "set device");
    if (int rc = upload_table(b, b->shape)) return rc;
    return tune(b);
}

int af_batch_step_n(af_batch *b, int64_t n, int32_t *faulted, int32_t cap, int32_t *nfaulted)
{
    if (int rc = check_open(b)) return rc;
    if (n < 0) return fail(AF_ERR_ARG, "aeroflux: negative step count");
    HIP_TRY(hipSetDevice(b->device), "set device");
    if (nfaulted) *nfaulted = 0;
    if (int rc = upload_table(b, b->shape)) return rc;
    if (!b->tuned && b->nact > 0) {
        if (int rc = tune(b)) return rc;
        if (int rc = upload_table(b, b->shape)) return rc;
    }
    if (int rc = get_graph(b)) return rc;
    const int m = b->cfg.MaxSlots;
    hipLaunchKernelGGL(kb_fill_int, dim3((m + 255) / 256), dim3(256), 0, b->stream, b->d_ctl, m,
                       0);
    HIP_TRY(hipGetLastError(), "clear fault flags");
    for (int64_t k = 0; k < n; k++) HIP_TRY(hipGraphLaunch(b->graph, b->stream), "launch batch step");
    HIP_TRY(hipMemcpyAsync(b->h_ints, b->d_ctl, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost,
                           b->stream),
            "read fault flags");
    if (int rc = sync(b, "batch step")) return rc;
    int nf = 0;
    for (int s = 0; s < m; s++) {
        if (!b->slots[s].used || !b->h_ints[s]) continue;
        if (faulted && nf < cap) faulted[nf] = s;
        nf++;
    }
    if (nfaulted) *nfaulted = nf;
    if (nf)
        return fail(AF_ERR_FAULT,
                    "aeroflux: device fault contained: non-finite departure coordinate in an "
                    "advection sweep (%d batch slots)",
                    nf);
    return AF_OK;
}

int af_batch_field_desc(const af_batch *b, int32_t slot, int id, af_field *out)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (!out || id < 0 || id >= AF_FIELD_COUNT)
        return fail(AF_ERR_ARG, "aeroflux: bad field descriptor arguments");
    *out = b->slots[slot].L.f[id];
    return AF_OK;
}

static int batch_copy(af_batch *b, int32_t slot, int id, void *host, int64_t count, bool write)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (id < 0 || id >= AF_FIELD_COUNT || !host)
        return fail(AF_ERR_ARG, "aeroflux: bad field copy arguments");
    const HSlot &h = b->slots[slot];
    const af_field &f = h.L.f[id];
    if (count != f.Count)
        return fail(AF_ERR_ARG, "aeroflux: field %d has %ld elements, got %ld", id, (long)f.Count,
                    (long)count);
    HIP_TRY(hipSetDevice(b->device), "set device");
    char *dev = b->pool.base + h.off + f.Off;
    if (write)
        HIP_TRY(hipMemcpyAsync(dev, host, (size_t)count * 8, hipMemcpyHostToDevice, b->stream),
                "write field");
    else
        HIP_TRY(hipMemcpyAsync(host, dev, (size_t)count * 8, hipMemcpyDeviceToHost, b->stream),
                "read field");
    return sync(b, "field copy");
}

int af_batch_field_read(af_batch *b, int32_t slot, int id, double *host, int64_t count)
{
    return batch_copy(b, slot, id, host, count, false);
}

int af_batch_field_write(af_batch *b, int32_t slot, int id, const double *host, int64_t count)
{
    return batch_copy(b, slot, id, const_cast<double *>(host), count, true);
}

int af_batch_snapshot_bytes(const af_batch *b, int32_t slot, int64_t *bytes)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (!bytes) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *bytes = snapshot_bytes(b->slots[slot].L.f);
    return AF_OK;
}

// The slot's reduction output word serves as the checksum accumulator; the
// stream is drained first, so no step is using it.
static int batch_snapshot_io(af_batch *b, int32_t slot, void *buf, int64_t bytes, bool save)
{
    if (int rc = check_slot(b, slot)) return rc;
    HIP_TRY(hipSetDevice(b->device), "set device");
    HIP_TRY(hipStreamSynchronize(b->stream), "batch snapshot");
    const HSlot &h = b->slots[slot];
    unsigned char *base = reinterpret_cast<unsigned char *>(b->pool.base + h.off);
    auto *acc = reinterpret_cast<unsigned long long *>(base + h.L.out_off);
    if (save) return snapshot_save(h.L.f, base, h.cfg, buf, bytes, acc, b->stream);
    return snapshot_load(h.L.f, base, h.cfg, buf, bytes, acc, b->stream);
}

int af_batch_snapshot(af_batch *b, int32_t slot, void *dst, int64_t bytes)
{
    return batch_snapshot_io(b, slot, dst, bytes, true);
}

int af_batch_restore(af_batch *b, int32_t slot, const void *src, int64_t bytes)
{
    return batch_snapshot_io(b, slot, const_cast<void *>(src), bytes, false);
}

int af_batch_apply_velocity_bc(af_batch *b, int32_t slot)
{
    if (int rc = check_slot(b, slot)) return rc;
    HIP_TRY(hipSetDevice(b->device), "set device");
    const HSlot &h = b->slots[slot];
    HIP_TRY(launch_velocity_bc(slot_field(b, h, AF_FIELD_U), slot_field(b, h, AF_FIELD_V),
                               slot_field(b, h, AF_FIELD_W), slot_geom(h), b->stream),
            "velocity bc");
    return sync(b, "velocity bc");
}

int af_batch_init_band_limited(af_batch *b, int32_t slot, uint64_t seed, int32_t kmax,
                               double slope, double rms)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (kmax < 1 || kmax > 64 || !(rms > 0) || std::isinf(rms) || std::isnan(slope))
        return fail(AF_ERR_ARG, "aeroflux: band-limited field needs 1 <= kmax <= 64, rms > 0");
    HIP_TRY(hipSetDevice(b->device), "set device");
    const HSlot &h = b->slots[slot];
    const DField U = slot_field(b, h, AF_FIELD_U), V = slot_field(b, h, AF_FIELD_V),
                 W = slot_field(b, h, AF_FIELD_W);
    HIP_TRY(launch_zero(U.d, h.L.f[AF_FIELD_U].Count, b->stream), "zero velocity");
    HIP_TRY(launch_zero(V.d, h.L.f[AF_FIELD_V].Count, b->stream), "zero velocity");
    HIP_TRY(launch_zero(W.d, h.L.f[AF_FIELD_W].Count, b->stream), "zero velocity");
    HIP_TRY(launch_band_limited(U, V, W, slot_geom(h), h.cfg.H, seed, kmax, slope, rms, b->ic,
                                b->stream),
            "band-limited initial condition");
    HIP_TRY(launch_velocity_bc(U, V, W, slot_geom(h), b->stream), "velocity bc");
    return sync(b, "band-limited initial condition");
}

static int slot_reduce(af_batch *b, int32_t slot, bool ke, double *out)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (!out) return fail(AF_ERR_ARG, "aeroflux: null output");
    HIP_TRY(hipSetDevice(b->device), "set device");
    const HSlot &h = b->slots[slot];
    const BSlot &d = b->h_slots[slot];
    const DField U = slot_field(b, h, AF_FIELD_U), V = slot_field(b, h, AF_FIELD_V),
                 W = slot_field(b, h, AF_FIELD_W);
    if (ke)
        HIP_TRY(launch_kinetic_energy(U, V, W, d.G, d.partials, d.out, b->stream),
                "kinetic energy");
    else
        HIP_TRY(launch_max_divergence(U, V, W, d.G, d.invH, d.partials, d.out, b->stream),
                "max divergence");
    HIP_TRY(hipMemcpyAsync(b->h_dbl, d.out, sizeof(double), hipMemcpyDeviceToHost, b->stream),
            "copy reduction");
    if (int rc = sync(b, "reduction")) return rc;
    double v = b->h_dbl[0];
    if (ke) {
        const double vol = h.cfg.H * h.cfg.H * h.cfg.H;
        v = 0.5 * v * vol;  // project.go:257
    }
    *out = v;
    return AF_OK;
}

int af_batch_max_divergence(af_batch *b, int32_t slot, double *out)
{
    return slot_reduce(b, slot, false, out);
}

int af_batch_kinetic_energy(af_batch *b, int32_t slot, double *out)
{
    return slot_reduce(b, slot, true, out);
}

int af_batch_mixed_stats(af_batch *b, int32_t slot, double *last_div, int32_t *cycles)
{
    if (int rc = check_slot(b, slot)) return rc;
    if (b->cfg.Precision != AF_PRECISION_MIXED)
        return fail(AF_ERR_ARG, "aeroflux: batch is not in mixed precision");
    HIP_TRY(hipSetDevice(b->device), "set device");
    HIP_TRY(hipMemcpyAsync(b->h_dbl, b->d_last + slot, sizeof(double), hipMemcpyDeviceToHost,
                           b->stream),
            "read mixed stats");
    HIP_TRY(hipMemcpyAsync(b->h_ints, b->d_ctl + 2 * b->cfg.MaxSlots + slot, sizeof(int),
                           hipMemcpyDeviceToHost, b->stream),
            "read mixed stats");
    if (int rc = sync(b, "mixed stats")) return rc;
    if (last_div) *last_div = b->h_dbl[0];
    if (cycles) *cycles = b->h_ints[0];
    return AF_OK;
}

int af_batch_time_stage(af_batch *b, int stage, int reps, double *avg_ms)
{
    if (int rc = check_open(b)) return rc;
    if (reps < 1 || !avg_ms || stage < 0 || stage > 9)
        return fail(AF_ERR_ARG, "aeroflux: bad timing arguments");
    const bool mixed = b->cfg.Precision == AF_PRECISION_MIXED;
    if (stage == 9 && !mixed)
        return fail(AF_ERR_ARG, "aeroflux: the residual stage needs a mixed-precision batch");
    HIP_TRY(hipSetDevice(b->device), "set device");
    if (int rc = upload_table(b, b->shape)) return rc;
    if (stage == 0) return time_grouped_pass(b, b->nwg_pass, reps, avg_ms);
    const BCtl c = ctl(b);
    const dim3 gp(b->nwg_other), blk(BX, BY);
    return time_launches(
        b, reps,
        [&]() -> hipError_t {
            switch (stage) {
            case 1: return enq_sweep(b, b->shape, b->nwg_sweep, 0, 0);
            case 2: return enq_div(b, 0);
            case 3: return enq_grad(b, 0);
            case 4: return enq_vbc(b, 0);
            case 5: return enq_sweep(b, b->shape, b->nwg_sweep, 1, 0);
            case 6: return enq_sweep(b, b->shape, b->nwg_sweep, 2, 0);
            case 7:
                hipLaunchKernelGGL(kb_pbc, dim3(kSmallGrid), dim3(256), 0, b->stream, c);
                return hipGetLastError();
            case 8:
                if (mixed)
                    hipLaunchKernelGGL((kb_from_split<float, false>), gp, blk, 0, b->stream, c, 0);
                else
                    hipLaunchKernelGGL((kb_from_split<double, false>), gp, blk, 0, b->stream, c, 0);
                return hipGetLastError();
            default:
                hipLaunchKernelGGL(kb_resid, gp, blk, 0, b->stream, c, 0);
                return hipGetLastError();
            }
        },
        avg_ms);
}

// ------------------------------------------------------------- serving

// Resident interior cells at which af_serve runs the autotuner.
constexpr int64_t kServeTuneCells = (int64_t)128 << 20;

int af_serve(af_batch *b, const af_serve_job *jobs, int64_t n, int32_t kmax, double slope,
             double rms, int32_t realtime, af_serve_result *out, af_serve_stats *stats,
             af_serve_progress progress, double progress_interval, void *user)
{
    if (int rc = check_open(b)) return rc;
    if ((!jobs || !out) && n > 0) return fail(AF_ERR_ARG, "aeroflux: null argument");
    if (n < 0 || kmax < 1 || kmax > 64 || !(rms > 0))
        return fail(AF_ERR_ARG, "aeroflux: bad serving arguments");
    for (const HSlot &h : b->slots)
        if (h.used) return fail(AF_ERR_ARG, "aeroflux: af_serve needs an empty batch");
    for (int64_t i = 1; i < n; i++)
        if (jobs[i].Arrival < jobs[i - 1].Arrival)
            return fail(AF_ERR_ARG, "aeroflux: jobs must be sorted by arrival time");
    HIP_TRY(hipSetDevice(b->device), "set device");

    // Job configs, as New would normalise them, and their pool footprints.
    std::vector<af_config> cfgs((size_t)n);
    std::vector<SlotLayout> lays((size_t)n);
    std::vector<int32_t> status((size_t)n, AF_OK);
    int64_t ic_bytes = 0;
    for (int64_t i = 0; i < n; i++) {
        const af_serve_job &j = jobs[i];
        af_config c{};
        c.NX = j.NX; c.NY = j.NY; c.NZ = j.NZ;
        c.H = j.H; c.Viscosity = j.Viscosity; c.DT = j.DT;
        c.GaussSeidelIters = b->cfg.GaussSeidelIters;
        c.Precision = b->cfg.Precision;
        c.MixedMaxCycles = b->cfg.MixedMaxCycles;
        c.MixedSweeps = b->cfg.MixedSweeps;
        c.MixedFixSweeps = b->cfg.MixedFixSweeps;
        if (b->cfg.MixedChecked) {
            if (!(j.MixedTarget > 0))
                return fail(AF_ERR_ARG,
                            "aeroflux: job %ld needs MixedTarget > 0 in a checked mixed batch",
                            (long)i);
            c.MixedTarget = j.MixedTarget;
        }
        if (j.Steps < 1) return fail(AF_ERR_ARG, "aeroflux: job %ld has no steps", (long)i);
        if (int rc = af_config_normalize(&c)) return rc;
        if (!slot_layout(c, c.Precision, &lays[i]))
            return fail(AF_ERR_CONFIG, "aeroflux: job %ld grid exceeds 2^31 elements per field",
                        (long)i);
        cfgs[i] = c;
        if (lays[i].total > b->pool.cap) status[i] = AF_ERR_OOM;
        const Geom G{(int)c.NX, (int)c.NY, (int)c.NZ, (int)c.Ghost};
        ic_bytes = std::max(ic_bytes, ic_scratch_bytes(G, kmax));
    }
    // Pinned per-job results: max|div|, kinetic-energy sum, fault flag.
    double *res = nullptr;
    int *flags = nullptr;
    HIP_TRY(hipStreamSynchronize(b->stream), "serve");
    HIP_TRY(b->ic.reserve(ic_bytes), "reserve initial-condition scratch");
    HIP_TRY(hipHostMalloc(&res, sizeof(double) * 2 * (size_t)std::max<int64_t>(n, 1),
                          hipHostMallocDefault),
            "allocate pinned results");
    if (hipHostMalloc(&flags, sizeof(int) * (size_t)std::max<int64_t>(n, 1),
                      hipHostMallocDefault) != hipSuccess) {
        (void)hipHostFree(res);
        return fail(AF_ERR_OOM, "aeroflux: allocate pinned results");
    }
    constexpr int DEPTH = 2;  // batched steps in flight behind the host
    constexpr int R = DEPTH + 1;
    static_assert(R <= kTabRing, "work-list staging ring too small");
    hipEvent_t ev[R] = {};
    struct Res {
        int64_t job;
        int slot;
        int64_t left;
        int64_t idle;  // consecutive batched steps that passed it over
    };
    std::vector<Res> resident;
    std::deque<std::pair<int64_t, std::vector<int64_t>>> pending;
    // Waiting jobs, ordered by key: arrival index (first come first served)
    // or remaining work, cells x steps left, ties by arrival (policy 1).
    // Policies 1 and 3 admit the least remaining work first; 1 preempts;
    // 2 and 3 step the least remaining work first under the cell budget.
    const bool by_work = b->cfg.ServePolicy == 1 || b->cfg.ServePolicy == 3;
    const bool preempting = b->cfg.ServePolicy == 1;
    const bool by_steps = b->cfg.ServePolicy == 2 || b->cfg.ServePolicy == 3;
    std::vector<std::pair<double, int64_t>> order;  // policy 2: (work left, resident index)
    std::set<std::pair<double, int64_t>> queue;
    auto work_of = [&](int64_t j, int64_t left) {
        return (double)(cfgs[j].NX * cfgs[j].NY * cfgs[j].NZ) * (double)left;
    };
    auto enqueue = [&](int64_t j, int64_t left) {
        queue.insert({by_work ? work_of(j, left) : (double)j, j});
    };
    // Preempted jobs: host snapshot of the slot's arena and its fault flag.
    struct Held {
        std::unique_ptr<unsigned char[]> buf;
        int64_t bytes = 0;
        int64_t left = 0;
        int flag = 0;
        bool once = false;  // preempted already
    };
    std::vector<Held> held(by_work ? (size_t)n : 0);
    int64_t held_bytes = 0;
    auto left_now = [&](int64_t j) {
        return by_work && held[(size_t)j].buf ? held[(size_t)j].left : jobs[j].Steps;
    };
    int64_t next = 0, step = 0, completed = -1, jobs_done = 0;
    double last_report = 0;
    af_serve_stats st{};
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    int rc = AF_OK;
    auto finish_oldest = [&]() -> hipError_t {
        auto &p = pending.front();
        hipError_t e = hipEventSynchronize(ev[p.first % R]);
        if (e != hipSuccess) return e;
        const double t = now();
        for (int64_t j : p.second) {
            out[j].Done = t;
            st.Wall = t;
        }
        jobs_done += (int64_t)p.second.size();
        completed = p.first;
        pending.pop_front();
        return hipSuccess;
    };
    hipError_t err = hipSuccess;
    for (int k = 0; k < R && err == hipSuccess; k++)
        err = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    for (int64_t i = 0; i < n; i++) {
        out[i] = af_serve_result{};
        out[i].Slot = -1;
        out[i].Status = status[i];
        out[i].FirstStep = -1;
        out[i].DoneStep = -1;
    }
    // Policy 1: snapshot the resident job with the most remaining work to
    // host memory, free its slot and requeue it, if it has more than
    // PreemptRatio times the work of queued job j and the snapshot fits the
    // host budget.  Every step in flight completes first (the snapshot is
    // stream-ordered and synchronises), so the snapshot is the state after
    // the job's enqueued steps and `left` steps remain.
    auto preempt_for = [&](int64_t j, bool *done) -> int {
        *done = false;
        if (!preempting || resident.empty()) return AF_OK;
        // A job is preempted at most once: after it resumes it runs to the
        // end, which bounds the snapshot traffic a steady stream of short
        // arrivals can cause.
        size_t v = resident.size();
        for (size_t r = 0; r < resident.size(); r++)
            if (!held[(size_t)resident[r].job].once &&
                (v == resident.size() || work_of(resident[r].job, resident[r].left) >
                                             work_of(resident[v].job, resident[v].left)))
                v = r;
        if (v == resident.size()) return AF_OK;
        const Res x = resident[v];
        // Strictly more: within one admission pass (no step in between) two
        // jobs can then never evict each other in turn.
        if (!(work_of(x.job, x.left) > b->cfg.PreemptRatio * work_of(j, left_now(j))))
            return AF_OK;
        HSlot &h = b->slots[x.slot];
        const int64_t nb = snapshot_bytes(h.L.f);
        if (held_bytes + nb > b->cfg.SnapshotBudget) return AF_OK;
        std::unique_ptr<unsigned char[]> buf(new (std::nothrow) unsigned char[(size_t)nb]);
        if (!buf) return AF_OK;
        unsigned char *base = reinterpret_cast<unsigned char *>(b->pool.base + h.off);
        auto *acc = reinterpret_cast<unsigned long long *>(base + h.L.out_off);
        if (int e = snapshot_save(h.L.f, base, h.cfg, buf.get(), nb, acc, b->stream)) return e;
        int flag = 0;
        HIP_TRY(hipMemcpy(&flag, b->d_ctl + x.slot, sizeof(int), hipMemcpyDeviceToHost),
                "preempt");
        while (!pending.empty()) HIP_TRY(finish_oldest(), "preempt");
        free_slot(b, x.slot);
        Held &k = held[(size_t)x.job];
        k.buf = std::move(buf);
        k.bytes = nb;
        k.left = x.left;
        k.flag = flag;
        k.once = true;
        held_bytes += nb;
        resident[v] = resident.back();
        resident.pop_back();
        enqueue(x.job, x.left);
        st.Preemptions++;
        *done = true;
        return AF_OK;
    };
    while (err == hipSuccess && rc == AF_OK) {
        const double t = now();
        while (next < n && (!realtime || jobs[next].Arrival <= t)) {
            enqueue(next, jobs[next].Steps);
            next++;
        }
        // Admission in queue order while the pool can place the head job
        // (policy 1 may preempt to make room).
        while (!queue.empty() && err == hipSuccess && rc == AF_OK) {
            const int64_t j = queue.begin()->second;
            if (status[j] == AF_ERR_OOM) {
                out[j].Admitted = out[j].Done = t;
                queue.erase(queue.begin());
                continue;
            }
            const int cand = pick_slot(b);
            if (cand >= 0) {
                // The slot's pinned descriptor is rewritten below; an upload
                // of it for an earlier job must have run (it has once that
                // job's first step is complete).
                while (b->slots[cand].desc_step > completed && !pending.empty() &&
                       (err = finish_oldest()) == hipSuccess) {
                }
                if (err != hipSuccess) break;
            }
            int slot;
            if (cand < 0 || place_slot(b, cfgs[j], lays[j], &slot) != AF_OK) {
                bool made_room = false;
                rc = preempt_for(j, &made_room);
                if (made_room) continue;
                break;
            }
            HSlot &h = b->slots[slot];
            h.desc_step = step;
            const DField U = slot_field(b, h, AF_FIELD_U), V = slot_field(b, h, AF_FIELD_V),
                         W = slot_field(b, h, AF_FIELD_W);
            const Geom G = slot_geom(h);
            int64_t left = jobs[j].Steps;
            if (by_work && held[(size_t)j].buf) {
                // Resume a preempted job from its snapshot.
                Held &k = held[(size_t)j];
                unsigned char *base = reinterpret_cast<unsigned char *>(b->pool.base + h.off);
                auto *acc = reinterpret_cast<unsigned long long *>(base + h.L.out_off);
                rc = snapshot_load(h.L.f, base, h.cfg, k.buf.get(), k.bytes, acc, b->stream);
                if (rc != AF_OK) break;
                if (k.flag) {
                    err = hipMemcpyAsync(b->d_ctl + slot, &k.flag, sizeof(int),
                                         hipMemcpyHostToDevice, b->stream);
                    if (err == hipSuccess) err = hipStreamSynchronize(b->stream);
                }
                left = k.left;
                held_bytes -= k.bytes;
                k.buf.reset();
            } else {
                if (err == hipSuccess)
                    err = launch_band_limited(U, V, W, G, h.cfg.H, jobs[j].Seed, kmax, slope,
                                              rms, b->ic, b->stream);
                if (err == hipSuccess) err = launch_velocity_bc(U, V, W, G, b->stream);
                out[j].Admitted = t;
            }
            out[j].Slot = slot;
            resident.push_back({j, slot, left, 0});
            queue.erase(queue.begin());
            st.MaxResident = std::max<int32_t>(st.MaxResident, (int32_t)resident.size());
            st.PeakBytes = std::max(st.PeakBytes, b->pool.used);
        }
        if (err != hipSuccess || rc != AF_OK) break;
        if (resident.empty() && !queue.empty()) {
            rc = fail(AF_ERR_OOM, "aeroflux: an empty batch cannot place job %ld",
                      (long)queue.begin()->second);
            break;
        }
        auto retire = [&](const std::vector<int64_t> &done_jobs) -> hipError_t {
            for (int64_t j : done_jobs) {
                const int slot = out[j].Slot;
                const HSlot &h = b->slots[slot];
                const BSlot &d = b->h_slots[slot];
                const DField U = slot_field(b, h, AF_FIELD_U), V = slot_field(b, h, AF_FIELD_V),
                             W = slot_field(b, h, AF_FIELD_W);
                hipError_t e = launch_max_divergence(U, V, W, d.G, d.invH, d.partials, d.out,
                                                     b->stream);
                if (e == hipSuccess)
                    e = launch_kinetic_energy(U, V, W, d.G, d.partials, d.out + 1, b->stream);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(res + 2 * j, d.out, 2 * sizeof(double),
                                       hipMemcpyDeviceToHost, b->stream);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(flags + j, b->d_ctl + slot, sizeof(int),
                                       hipMemcpyDeviceToHost, b->stream);
                if (e != hipSuccess) return e;
                st.CellUpdates += cfgs[j].NX * cfgs[j].NY * cfgs[j].NZ * jobs[j].Steps;
                free_slot(b, slot);
            }
            return hipSuccess;
        };
        if (resident.empty()) {
            while (!pending.empty() && err == hipSuccess) err = finish_oldest();
            if (err != hipSuccess) break;
            if (next >= n && queue.empty()) break;
            if (queue.empty()) {
                const double wait = jobs[next].Arrival - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            continue;
        }
        // Policies 2 and 3: this step advances the resident jobs with the
        // least remaining work (cells x steps left, ties by arrival), as many
        // as fit the cell budget, at least one; a job that does not fit is
        // passed over for smaller ones behind it.  With StepMaxWait > 0, a
        // job passed over by that many consecutive steps goes first (longest
        // wait first), which bounds how long a large job can be held back.
        // Every job still takes exactly its own steps, so results do not
        // depend on the choice.
        if (by_steps) {
            order.clear();
            const int64_t max_wait = b->cfg.StepMaxWait;
            for (size_t r = 0; r < resident.size(); r++) {
                const Res &x = resident[r];
                const bool starved = max_wait > 0 && x.idle >= max_wait;
                order.push_back({starved ? -(double)x.idle : work_of(x.job, x.left), (int64_t)r});
            }
            std::sort(order.begin(), order.end(), [&](const auto &p, const auto &q) {
                return p.first != q.first ? p.first < q.first
                                          : resident[(size_t)p.second].job <
                                                resident[(size_t)q.second].job;
            });
            std::vector<char> mask(b->slots.size(), 0);
            int64_t cells = 0;
            for (const auto &o : order) {
                const Res &x = resident[(size_t)o.second];
                const int64_t c = cfgs[x.job].NX * cfgs[x.job].NY * cfgs[x.job].NZ;
                if (cells > 0 && cells + c > b->cfg.StepCellBudget) continue;
                mask[(size_t)x.slot] = 1;
                cells += c;
            }
            if (mask != b->step_mask) {
                b->step_mask.swap(mask);
                b->tab_dirty = true;
            }
        }
        // Tune once the resident set is representative: a working set that
        // fills the GPU, or everything the stream will admit.  Until then the
        // default shape and grids run.
        if (!b->tuned && (b->cells >= kServeTuneCells || (next >= n && queue.empty()))) {
            if ((err = hipStreamSynchronize(b->stream)) != hipSuccess) break;
            while (!pending.empty() && err == hipSuccess) err = finish_oldest();
            if (err != hipSuccess) break;
            if ((rc = upload_table(b, b->shape)) != AF_OK) break;
            if ((rc = tune(b)) != AF_OK) break;
        }
        if (b->tab_dirty || b->tab_shape != b->shape) {
            const int ring = (int)(step % R);
            if ((rc = build_table(b, b->shape, ring)) != AF_OK) break;
            if (b->lds_need > b->lds_grid) {
                // The staged sweep's grid no longer covers the work-list: the
                // graph is rebuilt once the steps in flight have run.
                if ((err = hipStreamSynchronize(b->stream)) != hipSuccess) break;
                fit_lds_grid(b);
            }
            if ((err = hipMemcpyAsync(b->d_tab, b->h_tab + (int64_t)ring * b->tab_ints,
                                      (size_t)b->tab_ints * sizeof(int), hipMemcpyHostToDevice,
                                      b->stream)) != hipSuccess)
                break;
            b->tab_dirty = false;
            b->tab_shape = b->shape;
        }
        if ((rc = get_graph(b)) != AF_OK) break;
        if ((err = hipGraphLaunch(b->graph, b->stream)) != hipSuccess) break;
        if ((err = hipEventRecord(ev[step % R], b->stream)) != hipSuccess) break;
        std::vector<int64_t> done_now;
        for (size_t r = 0; r < resident.size();) {
            Res &x = resident[r];
            if (by_steps && !b->step_mask[(size_t)x.slot]) {
                x.idle++;  // not advanced by this step
                r++;
                continue;
            }
            x.idle = 0;
            if (out[x.job].FirstStep < 0) out[x.job].FirstStep = step;
            if (--x.left == 0) {
                out[x.job].DoneStep = step;
                done_now.push_back(x.job);
                std::swap(resident[r], resident.back());
                resident.pop_back();
            } else {
                r++;
            }
        }
        if ((err = retire(done_now)) != hipSuccess) break;
        pending.emplace_back(step, std::move(done_now));
        step++;
        while (pending.size() > (size_t)DEPTH && err == hipSuccess) err = finish_oldest();
        if (progress && err == hipSuccess && now() - last_report >= progress_interval) {
            last_report = now();
            af_serve_stats so_far = st;
            so_far.Wall = last_report;
            so_far.Steps = step;
            progress(&so_far, jobs_done, user);
        }
    }
    while (!pending.empty() && err == hipSuccess) err = finish_oldest();
    if (err == hipSuccess) err = hipStreamSynchronize(b->stream);
    for (int k = 0; k < R; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    if (!b->step_mask.empty()) {
        b->step_mask.clear();  // later batched steps advance every slot again
        b->tab_dirty = true;
    }
    if (err == hipSuccess && rc == AF_OK) {
        for (int64_t j = 0; j < n; j++) {
            if (status[j] == AF_ERR_OOM) continue;
            out[j].MaxDivergence = res[2 * j];
            const double h = cfgs[j].H;
            out[j].KineticEnergy = 0.5 * res[2 * j + 1] * (h * h * h);  // project.go:257
            if (flags[j]) out[j].Status = AF_ERR_FAULT;
        }
        st.Steps = step;
        if (stats) *stats = st;
    }
    (void)hipHostFree(res);
    (void)hipHostFree(flags);
    if (err != hipSuccess) return hip_fail(err, "serve");
    return rc;
}

}  // extern "C"
