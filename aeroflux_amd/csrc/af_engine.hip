// af_engine.hip -- the engine behind the C ABI (include/aeroflux.h).
//
// One engine = one zeroed HBM arena laid out exactly like the reference's
// MAC arena (grid.go:79-134), one HIP stream, and one hipGraph per Step
// (aeroflux.go:348-409) that replays the whole Strang palindrome.  The
// reference's per-phase goroutine dispatch + join becomes stream order; its
// serial boundary-condition fills become gather kernels (or are fused into
// the red-black passes); its guard-page fault containment becomes a device
// flag raised by a NaN departure coordinate and returned by Step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unistd.h>

#include "../../include/aeroflux.h"
#include "af_host.h"
#include "af_ic.h"
#include "af_internal.h"
#include "af_mixed.h"
#include "af_direct.h"
#include "af_snapshot.h"
#include "af_rbgs_split.h"
#include "af_rbgs_tb.h"

#include <map>
#include <mutex>
#include <tuple>

namespace af {

namespace {
thread_local std::string g_err;
constexpr int kCacheLine = 128;  // layout.go:13

af_field field_layout(int64_t *off, int64_t nx, int64_t ny, int64_t nz)
{
    af_field f;
    f.Off = *off;
    f.NX = nx; f.NY = ny; f.NZ = nz;
    f.SY = nx; f.SZ = nx * ny;
    f.Count = nx * ny * nz;
    int64_t sz = f.Count * 8;
    if (int64_t r = sz % kCacheLine) sz += kCacheLine - r;
    *off += sz;
    return f;
}
}  // namespace

const char *last_error() { return g_err.c_str(); }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t err, const char *what)
{
    return fail(AF_ERR_HIP, "aeroflux: %s: %s", what, hipGetErrorString(err));
}

int64_t mac_layout(af_field *f, int64_t nx, int64_t ny, int64_t nz, int64_t g)
{
    const int64_t g2 = 2 * g, cx = nx + g2, cy = ny + g2, cz = nz + g2;
    int64_t off = 0;
    f[AF_FIELD_P] = field_layout(&off, cx, cy, cz);
    f[AF_FIELD_U] = field_layout(&off, cx + 1, cy, cz);
    f[AF_FIELD_V] = field_layout(&off, cx, cy + 1, cz);
    f[AF_FIELD_W] = field_layout(&off, cx, cy, cz + 1);
    f[AF_FIELD_DIV] = field_layout(&off, cx, cy, cz);
    f[AF_FIELD_TMP0] = field_layout(&off, cx + 1, cy, cz);
    f[AF_FIELD_TMP1] = field_layout(&off, cx, cy + 1, cz);
    f[AF_FIELD_TMP2] = field_layout(&off, cx, cy, cz + 1);
    return off;
}

}  // namespace af

using namespace af;

namespace {

// Current descriptors of the six velocity-sized fields, as the step walks
// through its commitVelocity swaps.
struct Desc {
    af_field f[AF_FIELD_COUNT];
};

void commit(Desc &d)  // commitVelocity, stencil.go:230-235
{
    std::swap(d.f[AF_FIELD_U], d.f[AF_FIELD_TMP0]);
    std::swap(d.f[AF_FIELD_V], d.f[AF_FIELD_TMP1]);
    std::swap(d.f[AF_FIELD_W], d.f[AF_FIELD_TMP2]);
}

// How the FP64 red-black solve is scheduled.  k == 0: one pass per launch on
// the colour-split layout (the default); k < 0: one pass per launch on the
// arena layout; k > 0: k passes per launch, temporally blocked, arena layout.
struct Tile {
    int k = 0;
    int rows = 2;   // rows per thread (blocked)
    int cz = 16;    // z-planes per workgroup (blocked)
};

}  // namespace

struct af_engine {
    af_config cfg;
    int device = 0;
    bool closed = false;
    hipStream_t stream = nullptr;
    unsigned char *arena = nullptr;
    int64_t bytes = 0;
    Desc cur;          // descriptors after any stage-level swaps
    int64_t orig_off[AF_FIELD_COUNT];
    Geom G;
    double invH, h2, div_scale, grad_scale;
    int *flag = nullptr;        // device fault flag
    int *flag_host = nullptr;   // pinned mirror
    double *partials = nullptr;
    double *red_out = nullptr;
    double *red_host = nullptr; // pinned
    double *pbuf = nullptr;     // FP64 ping-pong partner of P for the blocked solve
    // Colour-split FP64 solve buffers: P and Div, red and black, interior only.
    double *ps_red = nullptr, *ps_black = nullptr, *ds_red = nullptr, *ds_black = nullptr;
    Tile tile64;                // red-black schedule (see Tile)
    int sweep_shape = 4;        // advection workgroup shape (tuned at New)
    int sweep_slab = 1 << 30;   // advection y-slab rows of the block order (tuned at New)
    int sweep_chunks = 0;       // LDS-staged sweep z-chunks per tile column (0: the shape's)
    int pass_slab = 1 << 30;    // colour-pass y-slab (workgroup rows) of the block order (tuned)
    hipGraphExec_t graphs[8] = {};  // keyed by the swap state of U/V/W
    MixedState mixed;
    ICScratch ic;               // initial-condition generator scratch (lazy)
    DirectState direct;         // direct pressure solve, once enabled
};

namespace {

DField dfield(const af_engine *e, const af_field &f)
{
    DField d;
    d.d = reinterpret_cast<double *>(e->arena + f.Off);
    d.sy = (int)f.SY;
    d.sz = (int)f.SZ;
    d.nx = (int)f.NX;
    d.ny = (int)f.NY;
    d.nz = (int)f.NZ;
    return d;
}

int swap_state(const af_engine *e, const Desc &d)
{
    int s = 0;
    if (d.f[AF_FIELD_U].Off != e->orig_off[AF_FIELD_U]) s |= 1;
    if (d.f[AF_FIELD_V].Off != e->orig_off[AF_FIELD_V]) s |= 2;
    if (d.f[AF_FIELD_W].Off != e->orig_off[AF_FIELD_W]) s |= 4;
    return s;
}

int check_open(const af_engine *e)
{
    if (!e) return fail(AF_ERR_ARG, "aeroflux: null engine");
    if (e->closed) return fail(AF_ERR_CLOSED, "aeroflux: engine is closed");
    return AF_OK;
}

// ---- enqueue helpers (used both eagerly and under stream capture) -------

hipError_t enq_velocity_bc(af_engine *e, const Desc &d)
{
    return launch_velocity_bc(dfield(e, d.f[AF_FIELD_U]), dfield(e, d.f[AF_FIELD_V]),
                              dfield(e, d.f[AF_FIELD_W]), e->G, e->stream);
}

// With bc the sweep also applies applyVelocityBC to its output
// (body::store_face_bc), which the step runs after every sweep.
hipError_t enq_sweep(af_engine *e, const Desc &d, int axis, double dt, bool bc = false)
{
    SweepArgs a;
    a.U = dfield(e, d.f[AF_FIELD_U]);
    a.V = dfield(e, d.f[AF_FIELD_V]);
    a.W = dfield(e, d.f[AF_FIELD_W]);
    a.T0 = dfield(e, d.f[AF_FIELD_TMP0]);
    a.T1 = dfield(e, d.f[AF_FIELD_TMP1]);
    a.T2 = dfield(e, d.f[AF_FIELD_TMP2]);
    a.G = e->G;
    a.dt = dt;
    // advectComponent, stencil.go:47,73: invH = 1.0/h; lam = nu*dt*invH*invH.
    a.invH = e->invH;
    a.lam = e->cfg.Viscosity * dt * e->invH * e->invH;
    a.flag = e->flag;
    a.shape = e->sweep_shape;
    a.slab = e->sweep_slab;
    a.zchunks = e->sweep_chunks;
    a.bc = bc ? 1 : 0;
    return launch_sweep(axis, a, e->stream);
}

// Whether the projection's divergence also writes the solve's split
// right-hand side (the colour-split FP64 solve and the mixed solve read it).
bool fused_rhs(const af_engine *e)
{
    if (e->direct.ready()) return false;
    return e->cfg.Precision == AF_PRECISION_MIXED || e->tile64.k == 0;
}

// computeDivergence; with `fuse` also the solve's split right-hand side
// and the solve's first (red) pass, whose values depend only on that
// right-hand side (k_divergence<3>, or <2> for the mixed solve's FP32 pass).
// div_dead (with fuse): Div itself is not stored, because nothing reads it
// before the next projection's divergence overwrites it.
hipError_t enq_divergence(af_engine *e, const Desc &d, bool fuse = false, bool div_dead = false)
{
    int split = 0;
    void *red = nullptr, *black = nullptr, *first = nullptr;
    if (fuse && e->cfg.Precision == AF_PRECISION_MIXED) {
        split = 2;
        red = e->mixed.r_red;
        black = e->mixed.r_black;
        first = e->mixed.e_red;
    } else if (fuse && e->tile64.k == 0) {
        split = 3;
        red = e->ds_red;
        black = e->ds_black;
        first = e->ps_red;
    }
    return launch_divergence(dfield(e, d.f[AF_FIELD_U]), dfield(e, d.f[AF_FIELD_V]),
                             dfield(e, d.f[AF_FIELD_W]), dfield(e, d.f[AF_FIELD_DIV]), e->G,
                             e->div_scale, e->stream, split, red, black, e->h2, first,
                             !(div_dead && split >= 2));
}

// subtractGradient; with `velocity_bc` also the velocity BC that follows it
// in project.
hipError_t enq_gradient(af_engine *e, const Desc &d, bool velocity_bc = false)
{
    return launch_gradient(dfield(e, d.f[AF_FIELD_U]), dfield(e, d.f[AF_FIELD_V]),
                           dfield(e, d.f[AF_FIELD_W]), dfield(e, d.f[AF_FIELD_P]), e->G,
                           e->grad_scale, e->stream, velocity_bc);
}

// solvePressureRedBlack (project.go:73-100), FP64, production form: the
// zeroing is folded into the first red pass and every intermediate
// applyPressureBC into the passes themselves; the final fill is materialised.
TbArgs<double> tb_args(const af_engine *e, const DField &P, const DField &D)
{
    TbArgs<double> a;
    a.src = P.d;
    a.dst = P.d;
    a.rhs = D.d;
    a.NX = e->G.NX; a.NY = e->G.NY; a.NZ = e->G.NZ; a.g = e->G.g;
    a.sy = P.sy; a.sz = P.sz;
    a.h2 = e->h2;
    a.color0 = AF_COLOR_RED;
    a.cz = e->tile64.cz;
    return a;
}

// Temporally blocked form: 2*GaussSeidelIters colour passes in launches of
// tile64.k passes, ping-ponging between P and pbuf so that the last launch
// writes P; the first launch starts from P == 0 without reading it (the
// solve's zeroing).  Bit-identical to the single-pass form.
hipError_t enq_solve_fp64_blocked(af_engine *e, const DField &P, const DField &D)
{
    const int passes = (int)(2 * e->cfg.GaussSeidelIters);
    const int K = e->tile64.k;
    const int launches = (passes + K - 1) / K;
    TbArgs<double> a = tb_args(e, P, D);
    // Launch i (0-based) writes P when (launches - 1 - i) is even.
    double *bufs[2] = {P.d, e->pbuf};
    int done = 0;
    for (int i = 0; i < launches; i++) {
        const int k = (passes - done) < K ? (passes - done) : K;
        double *dst = bufs[(launches - 1 - i) & 1];
        a.dst = dst;
        hipError_t err = launch_rbgs_tb<double>(a, k, e->tile64.rows, i == 0, e->stream);
        if (err != hipSuccess) return err;
        a.src = dst;
        done += k;
    }
    return launch_pressure_bc(P, e->G, e->G.g > 1, e->stream);
}

SplitArgs<double> split_args(const af_engine *e)
{
    SplitArgs<double> a;
    a.red = e->ps_red;
    a.black = e->ps_black;
    a.rhs_red = e->ds_red;
    a.rhs_black = e->ds_black;
    a.NX = e->G.NX; a.NY = e->G.NY; a.NZ = e->G.NZ; a.NXH = (e->G.NX + 1) / 2; a.g = e->G.g;
    a.h2 = e->h2;
    a.slab = e->pass_slab;
    return a;
}

// Colour-split form: Div is split once, the 2*GaussSeidelIters passes run on
// the compacted colour arrays (the first from P == 0, which is the solve's
// zeroing), and the result is merged back into P's interior before the final
// ghost fill.  Bit-identical to the reference's sequence.  rhs_ready: the
// projection's divergence (k_divergence<3>) already wrote the split
// right-hand side and the first red pass, so the passes start at the first
// black one.
hipError_t enq_solve_fp64_split(af_engine *e, const DField &P, const DField &D, bool rhs_ready,
                                bool merge = true)
{
    hipError_t err = rhs_ready ? hipSuccess
                               : launch_to_split<double, double>(D.d, D.sy, D.sz, e->G, e->ds_red,
                                                                 e->ds_black, e->stream);
    const SplitArgs<double> a = split_args(e);
    const int passes = (int)(2 * e->cfg.GaussSeidelIters);
    for (int p = rhs_ready ? 1 : 0; p < passes && err == hipSuccess; p++)
        err = launch_split_pass<double>(a, p & 1, p == 0 ? 1 : (p == 1 ? 2 : 0), e->stream);
    if (!merge) return err;  // the caller reads the solution from the split arrays
    if (err == hipSuccess)
        err = launch_from_split<double>(P.d, P.sy, P.sz, e->G, e->ps_red, e->ps_black, false,
                                        nullptr, e->stream);
    if (err == hipSuccess) err = launch_pressure_bc(P, e->G, e->G.g > 1, e->stream);
    return err;
}

hipError_t enq_solve_fp64(af_engine *e, const Desc &d, bool rhs_ready)
{
    const DField P = dfield(e, d.f[AF_FIELD_P]);
    const DField D = dfield(e, d.f[AF_FIELD_DIV]);
    if (e->tile64.k > 0) return enq_solve_fp64_blocked(e, P, D);
    if (e->tile64.k == 0) return enq_solve_fp64_split(e, P, D, rhs_ready);
    hipError_t err = launch_rbgs_first(P, D, e->G, e->h2, e->stream);
    if (err != hipSuccess) return err;
    err = launch_rbgs(P, D, e->G, e->h2, AF_COLOR_BLACK, true, e->stream);
    for (int64_t it = 1; it < e->cfg.GaussSeidelIters && err == hipSuccess; it++) {
        err = launch_rbgs(P, D, e->G, e->h2, AF_COLOR_RED, true, e->stream);
        if (err == hipSuccess) err = launch_rbgs(P, D, e->G, e->h2, AF_COLOR_BLACK, true, e->stream);
    }
    if (err == hipSuccess) err = launch_pressure_bc(P, e->G, e->G.g > 1, e->stream);
    return err;
}

// rhs_ready: the divergence of this projection already wrote the split
// right-hand side (enq_divergence with fuse), and on the FP64 colour-split
// path also the solve's first red pass.
// check: a mixed-precision solve iterates to MixedTarget (when set); else
// it runs one cycle.
hipError_t enq_solve(af_engine *e, const Desc &d, bool rhs_ready = false, bool merge = true,
                     bool check = true)
{
    if (e->direct.ready()) {
        int bs = 0;
        return enqueue_direct_solve(e->direct, dfield(e, d.f[AF_FIELD_P]),
                                    dfield(e, d.f[AF_FIELD_DIV]), e->G, e->stream, &bs);
    }
    if (e->cfg.Precision == AF_PRECISION_MIXED) {
        MixedCtx c;
        c.P = dfield(e, d.f[AF_FIELD_P]);
        c.D = dfield(e, d.f[AF_FIELD_DIV]);
        c.G = e->G;
        c.h2 = e->h2;
        c.sweeps = e->cfg.MixedSweeps;
        c.fix_sweeps = e->cfg.MixedFixSweeps;
        c.cycles = e->cfg.MixedMaxCycles;
        c.target = check ? e->cfg.MixedTarget : 0;
        c.res_to_div = e->cfg.DT / (e->cfg.Density * e->h2);
        c.rhs_ready = rhs_ready;
        c.merge = merge;
        c.slab = e->pass_slab;
        return enqueue_mixed_solve(e->mixed, c, e->stream);
    }
    return enq_solve_fp64(e, d, rhs_ready);
}

// project, project.go:12-34.  The gradient subtraction applies the closing
// velocity BC itself (body::gradient_pair), which is the same values as the
// separate launch.  bc_applied: the velocity already holds the opening BC
// (the sweep before it applied the BC to its output).
//
// On the FP64 colour-split path and on the mixed path the solution is not
// merged into P by the solve: the gradient reads it from the split arrays
// (launch_gradient_split; the FP64 solution, or the FP32 correction, which
// is the whole of a mixed solution unless defect-correction cycles ran: the
// same values) and writes P's interior as it goes, and P's ghost fill
// follows.  A mixed solve with a max|div| target (MixedTarget) is checked
// on the projection whose divergence the step leaves; when it needed
// defect-correction cycles the solution is in P instead, and the gradient
// reading P runs in place of the split one (both launched, each gated on
// the solve's device flag).  p_dead: nothing reads P or Div after this
// projection's gradient before the next projection overwrites them (the
// step's first projection), so Div is not stored, P not written at all, and
// the mixed solve runs its one cycle unchecked (the second projection
// removes what divergence it leaves, and is checked).
hipError_t enq_project(af_engine *e, const Desc &d, bool bc_applied = false, bool p_dead = false)
{
    hipError_t err = bc_applied ? hipSuccess : enq_velocity_bc(e, d);
    const bool mixed = e->cfg.Precision == AF_PRECISION_MIXED;
    const bool split_path = !e->direct.ready() && (mixed || e->tile64.k == 0);
    const bool checked = mixed && e->cfg.MixedTarget > 0 && !p_dead;
    const DField P = dfield(e, d.f[AF_FIELD_P]);
    if (e->direct.method == DIRECT_FFT) {
        // The direct solve's first pass computes the divergence itself (the
        // same values, stored to Div), which saves re-reading Div.
        if (err == hipSuccess)
            err = enqueue_direct_project_solve(
                e->direct, P, dfield(e, d.f[AF_FIELD_DIV]), dfield(e, d.f[AF_FIELD_U]),
                dfield(e, d.f[AF_FIELD_V]), dfield(e, d.f[AF_FIELD_W]), e->div_scale, e->G,
                e->stream);
    } else {
        const bool fuse = fused_rhs(e);
        if (err == hipSuccess) err = enq_divergence(e, d, fuse, split_path && p_dead);
        if (err == hipSuccess && split_path && !mixed)
            err = enq_solve_fp64_split(e, P, dfield(e, d.f[AF_FIELD_DIV]), fuse, false);
        else if (err == hipSuccess)
            err = enq_solve(e, d, fuse, !split_path, !p_dead);
    }
    if (err == hipSuccess && split_path) {
        const DField U = dfield(e, d.f[AF_FIELD_U]), V = dfield(e, d.f[AF_FIELD_V]),
                     W = dfield(e, d.f[AF_FIELD_W]);
        const DField *Pout = p_dead ? nullptr : &P;
        const int *merged = checked ? e->mixed.merged : nullptr;
        err = mixed ? launch_gradient_split<float>(U, V, W, e->mixed.e_red, e->mixed.e_black, e->G,
                                                   e->grad_scale, e->stream, Pout, merged, 0)
                    : launch_gradient_split<double>(U, V, W, e->ps_red, e->ps_black, e->G,
                                                    e->grad_scale, e->stream, Pout);
        if (err == hipSuccess && checked)
            err = launch_gradient(U, V, W, P, e->G, e->grad_scale, e->stream, true, merged, 1);
        // The solve's final applyPressureBC (project.go:98), on the P the
        // gradient wrote.
        if (err == hipSuccess && !p_dead)
            err = launch_pressure_bc(P, e->G, e->G.g > 1, e->stream, merged, 0);
        return err;
    }
    if (err == hipSuccess) err = enq_gradient(e, d, true);
    return err;
}

// Step, aeroflux.go:348-409.  d is advanced through the six swaps, which
// return it to its starting state.
hipError_t enq_step(af_engine *e, Desc &d)
{
    static const int fwd[3] = {AF_AXIS_X, AF_AXIS_Y, AF_AXIS_Z};
    static const int rev[3] = {AF_AXIS_Z, AF_AXIS_Y, AF_AXIS_X};
    const double half = e->cfg.DT * 0.5;
    hipError_t err = hipSuccess;
    // Every sweep is followed by applyVelocityBC: the next sweep's, or the
    // one that opens project (project.go:13).  Each sweep applies it to its
    // own output (body::store_face_bc), so only the step's first BC is
    // launched.  The BC writes only cells it never reads, so applying it a
    // second time would leave the same values.  The projection ends with
    // the BC (applied by the gradient subtraction) and nothing touches the
    // velocity before the reverse sweeps' first BC, which is skipped too.
    // Whole planes of V and W, ghosts included, are then written in one
    // pass, which also keeps the sweep's row writes to whole cache lines.
    err = enq_velocity_bc(e, d);
    for (int k = 0; k < 3 && err == hipSuccess; k++) {
        err = enq_sweep(e, d, fwd[k], half, true);
        commit(d);
    }
    // P after the first projection is overwritten by the second solve.
    if (err == hipSuccess) err = enq_project(e, d, true, true);
    for (int k = 0; k < 3 && err == hipSuccess; k++) {
        err = enq_sweep(e, d, rev[k], half, true);
        commit(d);
    }
    if (err == hipSuccess) err = enq_project(e, d, true);
    return err;
}

int get_graph(af_engine *e, hipGraphExec_t *out)
{
    const int s = swap_state(e, e->cur);
    if (!e->graphs[s]) {
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal),
                "begin step capture");
        Desc d = e->cur;
        hipError_t err = enq_step(e, d);
        hipError_t err2 = hipStreamEndCapture(e->stream, &g);
        if (err != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(err, "capture step");
        }
        if (err2 != hipSuccess) return hip_fail(err2, "end step capture");
        hipGraphExec_t x = nullptr;
        hipError_t err3 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err3 != hipSuccess) return hip_fail(err3, "instantiate step graph");
        e->graphs[s] = x;
    }
    *out = e->graphs[s];
    return AF_OK;
}

int sync_stream(af_engine *e, const char *what)
{
    HIP_TRY(hipStreamSynchronize(e->stream), what);
    return AF_OK;
}

int run_reduction(af_engine *e, double *out)
{
    HIP_TRY(hipMemcpyAsync(e->red_host, e->red_out, sizeof(double), hipMemcpyDeviceToHost,
                           e->stream),
            "copy reduction");
    if (int rc = sync_stream(e, "reduction")) return rc;
    *out = *e->red_host;
    return AF_OK;
}

void release(af_engine *e)
{
    if (e->closed) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &g : e->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    release_mixed(e->mixed);
    release_direct(e->direct);
    e->ic.release();
    if (e->arena) (void)hipFree(e->arena);
    if (e->pbuf) (void)hipFree(e->pbuf);
    e->pbuf = nullptr;
    double **split[4] = {&e->ps_red, &e->ps_black, &e->ds_red, &e->ds_black};
    for (double **b : split) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (e->flag) (void)hipFree(e->flag);
    if (e->partials) (void)hipFree(e->partials);
    if (e->red_out) (void)hipFree(e->red_out);
    if (e->flag_host) (void)hipHostFree(e->flag_host);
    if (e->red_host) (void)hipHostFree(e->red_host);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    e->arena = nullptr;
    e->flag = nullptr;
    e->partials = e->red_out = e->red_host = nullptr;
    e->flag_host = nullptr;
    e->stream = nullptr;
    e->closed = true;
}

// ---- schedule selection ---------------------------------------------------

// Tuned blocked tiles are cached per (grid, ghost, passes per launch) for the
// process, so engines of the same shape (batches, served jobs) tune once.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, Tile> g_tune_cache;

bool tile_fits(const Tile &t)
{
    return rbgs_tb_supported(t.k, t.rows) && 16 * t.rows - 2 * t.k > 0 && t.cz >= 1;
}

// Times a few launches of every (rows, z-chunk) candidate for k passes per
// launch on the engine's own buffers (src is never written) and keeps the
// fastest.
int tune_blocked(af_engine *e, int k, Tile *out)
{
    const DField P = dfield(e, e->cur.f[AF_FIELD_P]);
    const DField D = dfield(e, e->cur.f[AF_FIELD_DIV]);
    const int rowss[] = {1, 2};
    const int czs[] = {8, 16, 32, 64};
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), "create event");
    HIP_TRY(hipEventCreate(&b), "create event");
    float best = 1e30f;
    Tile bt;
    bt.k = k;
    for (int rows : rowss)
        for (int cz : czs) {
            Tile t;
            t.k = k; t.rows = rows; t.cz = cz;
            if (!tile_fits(t)) continue;
            TbArgs<double> ta = tb_args(e, P, D);
            ta.dst = e->pbuf;
            ta.cz = cz;
            hipError_t err = launch_rbgs_tb<double>(ta, k, rows, false, e->stream);
            const int reps = 3;
            if (err == hipSuccess) err = hipEventRecord(a, e->stream);
            for (int r = 0; r < reps && err == hipSuccess; r++)
                err = launch_rbgs_tb<double>(ta, k, rows, false, e->stream);
            if (err == hipSuccess) err = hipEventRecord(b, e->stream);
            if (err == hipSuccess) err = hipEventSynchronize(b);
            float ms = 0;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
            if (err != hipSuccess) {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                return hip_fail(err, "tune red-black tile");
            }
            if (ms < best) {
                best = ms;
                bt = t;
            }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *out = bt;
    return AF_OK;
}

// Advection workgroup shape, y-slab height of the block order and, for the
// LDS-staged shapes, z-chunks per tile column (the shape's own, or the count
// that best fills whole rounds of resident workgroups): every choice gives
// the same bits (faces are independent), so the fastest on this grid is
// kept.  Timed on the engine's own (zeroed) arena, X, Y and Z sweeps twice
// each, as the step launches them (with the velocity BC); the sweeps write
// only the Tmp fields, whose BC ghosts then hold -0, so the Tmp fields are
// zeroed again afterwards.  Cached per grid for the process.
std::map<std::tuple<int, int, int, int>, std::tuple<int, int, int>> g_sweep_cache;

constexpr int kPlaneOrder = 1 << 30;

int choose_sweep_shape(af_engine *e)
{
    const af_config &c = e->cfg;
    if (c.SweepShape > sweep_shape_count() || c.SweepShape < 0)
        return fail(AF_ERR_CONFIG, "aeroflux: sweep shape %d out of range 1..%d",
                    (int)c.SweepShape, sweep_shape_count());
    if (c.SweepSlab < -1)
        return fail(AF_ERR_CONFIG, "aeroflux: sweep slab %d must be -1, 0 or positive",
                    (int)c.SweepSlab);
    if (c.SweepShape > 0) e->sweep_shape = c.SweepShape - 1;
    if (c.SweepSlab != 0) e->sweep_slab = c.SweepSlab < 0 ? kPlaneOrder : (int)c.SweepSlab;
    if (c.SweepShape > 0 && c.SweepSlab != 0) return AF_OK;
    const Geom &G = e->G;
    const auto key = std::make_tuple(G.NX, G.NY, G.NZ, G.g);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_sweep_cache.find(key);
        if (it != g_sweep_cache.end() && c.SweepShape == 0 && c.SweepSlab == 0) {
            e->sweep_shape = std::get<0>(it->second);
            e->sweep_slab = std::get<1>(it->second);
            e->sweep_chunks = std::get<2>(it->second);
            return AF_OK;
        }
    }
    static const int kSlabs[] = {kPlaneOrder, 8, 16, 32, 64};
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), "create event");
    HIP_TRY(hipEventCreate(&b), "create event");
    float best = 1e30f;
    int bs = e->sweep_shape, bl = e->sweep_slab, bc = 0;
    hipError_t err = hipSuccess;
    for (int s = 0; s < sweep_shape_count() && err == hipSuccess; s++) {
        if (c.SweepShape > 0 && s != c.SweepShape - 1) continue;
        // Staged shapes: their own chunk count, and the round-filling one.
        int chunks[2] = {0, 0};
        if (s >= 8) {
            const int v = s - 8, m = sweep_lds_chunks(G, v);
            if (m != sweep_lds_default_chunks(G, v)) chunks[1] = m;
        }
        for (int sl : kSlabs) {
            if (c.SweepSlab != 0 && sl != e->sweep_slab) continue;
            if (s >= 8 && c.SweepSlab == 0 && sl != kPlaneOrder) continue;  // no slab order
            for (int ci = 0; ci < 2 && err == hipSuccess; ci++) {
                if (ci == 1 && chunks[1] == 0) continue;
                e->sweep_shape = s;
                e->sweep_slab = sl;
                e->sweep_chunks = chunks[ci];
                err = enq_sweep(e, e->cur, AF_AXIS_X, e->cfg.DT * 0.5, true);
                if (err == hipSuccess) err = hipEventRecord(a, e->stream);
                for (int r = 0; r < 2 && err == hipSuccess; r++)
                    for (int ax = 0; ax < 3 && err == hipSuccess; ax++)
                        err = enq_sweep(e, e->cur, ax, e->cfg.DT * 0.5, true);
                if (err == hipSuccess) err = hipEventRecord(b, e->stream);
                if (err == hipSuccess) err = hipEventSynchronize(b);
                float ms = 0;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
                if (err == hipSuccess && ms < best) {
                    best = ms;
                    bs = s;
                    bl = sl;
                    bc = chunks[ci];
                }
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (int f = AF_FIELD_TMP0; f <= AF_FIELD_TMP2 && err == hipSuccess; f++) {
        const DField t = dfield(e, e->cur.f[f]);
        err = hipMemsetAsync(t.d, 0, (size_t)t.sz * t.nz * sizeof(double), e->stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return hip_fail(err, "tune advection workgroup shape");
    e->sweep_shape = bs;
    e->sweep_slab = bl;
    e->sweep_chunks = bc;
    if (c.SweepShape == 0 && c.SweepSlab == 0) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_sweep_cache[key] = std::make_tuple(bs, bl, bc);
    }
    return AF_OK;
}

// The colour passes' workgroup order: plane order, or y-slabs of 2-32
// workgroup rows walked through z.  A pass reads the planes above and below
// each cell, and once three planes of the other colour no longer fit an
// XCD's L2 the plane order fetches them again; a slab keeps them.  Timed on
// the engine's own passes (FP64 split, or the mixed solve's FP32 ones), which
// touch only solve buffers every solve rewrites before reading; cached per
// grid and precision.
std::map<std::tuple<int, int, int, int, int>, int> g_pass_cache;

int choose_pass_slab(af_engine *e)
{
    const bool mixed = e->cfg.Precision == AF_PRECISION_MIXED;
    if (e->cfg.PassSlab < -1)
        return fail(AF_ERR_CONFIG, "aeroflux: pass slab %d must be -1, 0 or positive",
                    (int)e->cfg.PassSlab);
    if (e->cfg.PassSlab != 0) {
        e->pass_slab = e->cfg.PassSlab < 0 ? kPlaneOrder : (int)e->cfg.PassSlab;
        return AF_OK;
    }
    if (!mixed && e->tile64.k != 0) return AF_OK;
    const Geom &G = e->G;
    const auto key = std::make_tuple(G.NX, G.NY, G.NZ, G.g, (int)mixed);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_pass_cache.find(key);
        if (it != g_pass_cache.end()) {
            e->pass_slab = it->second;
            return AF_OK;
        }
    }
    static const int kSlabs[] = {kPlaneOrder, 2, 4, 8, 16, 32};
    SplitArgs<double> a64 = split_args(e);
    SplitArgs<float> a32;
    if (mixed) {
        a32.red = e->mixed.e_red;
        a32.black = e->mixed.e_black;
        a32.rhs_red = e->mixed.r_red;
        a32.rhs_black = e->mixed.r_black;
        a32.NX = G.NX; a32.NY = G.NY; a32.NZ = G.NZ; a32.NXH = (G.NX + 1) / 2; a32.g = G.g;
        a32.h2 = 0;
    }
    auto pass = [&](int sl) -> hipError_t {
        if (mixed) {
            a32.slab = sl;
            return launch_split_pass<float>(a32, AF_COLOR_RED, 0, e->stream);
        }
        a64.slab = sl;
        return launch_split_pass<double>(a64, AF_COLOR_RED, 0, e->stream);
    };
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0), "create event");
    HIP_TRY(hipEventCreate(&t1), "create event");
    float best = 1e30f;
    int bl = kPlaneOrder;
    hipError_t err = hipSuccess;
    for (int sl : kSlabs) {
        if (err == hipSuccess) err = pass(sl);
        if (err == hipSuccess) err = hipEventRecord(t0, e->stream);
        for (int r = 0; r < 6 && err == hipSuccess; r++) err = pass(sl);
        if (err == hipSuccess) err = hipEventRecord(t1, e->stream);
        if (err == hipSuccess) err = hipEventSynchronize(t1);
        float ms = 0;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
        if (err == hipSuccess && ms < best) {
            best = ms;
            bl = sl;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (err != hipSuccess) return hip_fail(err, "tune colour-pass order");
    e->pass_slab = bl;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_pass_cache[key] = bl;
    return AF_OK;
}

// TileK = 0: one colour pass per launch, on the colour-split layout or on the
// arena layout, whichever solves this grid faster.  The split layout moves
// half the bytes per pass but adds a gather and a scatter per solve and
// launches more workgroups; at 256^3 it is 1.7x faster, at 64^3 and below the
// arena layout wins (fewer, lighter launches).  Timed on the engine's own
// zeroed arena (a solve from zeros writes +0 everywhere, so nothing changes);
// cached per grid, ghost width and sweep count for the process.
std::map<std::tuple<int, int, int, int, int64_t>, int> g_layout_cache;

int choose_one_pass_layout(af_engine *e)
{
    const Geom &G = e->G;
    Tile &t = e->tile64;
    // The mixed solve has its own FP32 split buffers and never runs the FP64
    // schedule (whose split buffers it does not allocate).
    if (e->cfg.Precision == AF_PRECISION_MIXED) {
        t.k = 0;
        return AF_OK;
    }
    const auto key = std::make_tuple(G.NX, G.NY, G.NZ, G.g, (int64_t)e->cfg.GaussSeidelIters);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_layout_cache.find(key);
        if (it != g_layout_cache.end()) {
            t.k = it->second;
            return AF_OK;
        }
    }
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0), "create event");
    HIP_TRY(hipEventCreate(&t1), "create event");
    float best = 1e30f;
    int pick = 0;
    hipError_t err = hipSuccess;
    for (int k : {0, -1}) {
        t.k = k;
        if (err == hipSuccess) err = enq_solve_fp64(e, e->cur, false);
        if (err == hipSuccess) err = hipEventRecord(t0, e->stream);
        for (int r = 0; r < 2 && err == hipSuccess; r++) err = enq_solve_fp64(e, e->cur, false);
        if (err == hipSuccess) err = hipEventRecord(t1, e->stream);
        if (err == hipSuccess) err = hipEventSynchronize(t1);
        float ms = 0;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
        if (err == hipSuccess && ms < best) {
            best = ms;
            pick = k;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (err != hipSuccess) return hip_fail(err, "tune solve layout");
    t.k = pick;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_layout_cache[key] = pick;
    return AF_OK;
}

int choose_schedule(af_engine *e)
{
    const af_config &c = e->cfg;
    const Geom &G = e->G;
    Tile &t = e->tile64;
    if (c.TileK < -2)
        return fail(AF_ERR_CONFIG, "aeroflux: TileK %d must be 0, -1, -2 or 2, 4, 6",
                    (int)c.TileK);
    if (c.TileK < 0) {
        t.k = c.TileK == -1 ? -1 : 0;
        return AF_OK;
    }
    if (c.TileK == 0) return choose_one_pass_layout(e);
    t.k = c.TileK;
    t.rows = c.TileRows > 0 ? c.TileRows : 2;
    t.cz = c.TileZ > 0 ? c.TileZ : 16;
    if (!tile_fits(t))
        return fail(AF_ERR_CONFIG,
                    "aeroflux: unsupported red-black tile k=%d rows=%d z=%d (k in {2,4,6}, rows "
                    "in {1,2}, 16*rows > 2*k)",
                    t.k, t.rows, t.cz);
    if (c.TileRows > 0 && c.TileZ > 0) return AF_OK;
    const auto key = std::make_tuple(G.NX, G.NY, G.NZ, G.g, t.k);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            t = it->second;
            return AF_OK;
        }
    }
    if (int rc = tune_blocked(e, t.k, &t)) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = t;
    return AF_OK;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

const char *af_last_error(void) { return af::last_error(); }

int af_config_normalize(af_config *c)
{
    if (!c) return fail(AF_ERR_ARG, "aeroflux: null config");
    // Config.Normalize, aeroflux.go:47-113, check for check.
    if (c->NX <= 0 || c->NY <= 0 || c->NZ <= 0)
        return fail(AF_ERR_CONFIG, "aeroflux: interior extents must be positive, got %ldx%ldx%ld",
                    (long)c->NX, (long)c->NY, (long)c->NZ);
    if (c->Ghost == 0) c->Ghost = 1;
    if (c->Ghost < 1)
        return fail(AF_ERR_CONFIG, "aeroflux: ghost thickness must be >= 1, got %ld", (long)c->Ghost);
    if (c->H == 0) {
        int64_t m = c->NX > c->NY ? c->NX : c->NY;
        m = m > c->NZ ? m : c->NZ;
        c->H = 1.0 / (double)m;
    }
    if (c->H <= 0 || std::isnan(c->H) || std::isinf(c->H))
        return fail(AF_ERR_CONFIG, "aeroflux: cell size must be finite and positive, got %g", c->H);
    if (c->Density == 0) c->Density = 1.0;
    if (c->Density <= 0 || std::isnan(c->Density) || std::isinf(c->Density))
        return fail(AF_ERR_CONFIG, "aeroflux: density must be finite and positive, got %g",
                    c->Density);
    if (c->Viscosity < 0 || std::isnan(c->Viscosity) || std::isinf(c->Viscosity))
        return fail(AF_ERR_CONFIG, "aeroflux: viscosity must be finite and non-negative, got %g",
                    c->Viscosity);
    if (c->DT == 0) {
        if (c->Viscosity > 0)
            c->DT = 0.5 * c->H * c->H / (6 * c->Viscosity);
        else
            c->DT = 0.25 * c->H;
    }
    if (c->DT <= 0 || std::isnan(c->DT) || std::isinf(c->DT))
        return fail(AF_ERR_CONFIG, "aeroflux: timestep must be finite and positive, got %g", c->DT);
    if (c->Workers == 0) c->Workers = (int64_t)sysconf(_SC_NPROCESSORS_ONLN);
    if (c->Workers < 1)
        return fail(AF_ERR_CONFIG, "aeroflux: worker count must be >= 1, got %ld", (long)c->Workers);
    if (c->Workers > c->NZ) c->Workers = c->NZ;
    if (c->GaussSeidelIters == 0) c->GaussSeidelIters = 24;
    if (c->GaussSeidelIters < 1)
        return fail(AF_ERR_CONFIG, "aeroflux: gauss-seidel iterations must be >= 1, got %ld",
                    (long)c->GaussSeidelIters);
    if (c->Viscosity > 0) {
        const double limit = c->H * c->H / (6 * c->Viscosity);
        if (c->DT > limit)
            return fail(AF_ERR_CONFIG,
                        "aeroflux: dt %g exceeds explicit diffusion stability limit %g (h=%g, "
                        "nu=%g); reduce dt or h",
                        c->DT, limit, c->H, c->Viscosity);
    }
    // Extensions (no reference counterpart).
    if (c->Precision != AF_PRECISION_FP64 && c->Precision != AF_PRECISION_MIXED)
        return fail(AF_ERR_CONFIG, "aeroflux: unknown precision mode %d", (int)c->Precision);
    if (c->MixedMaxCycles == 0) c->MixedMaxCycles = 4;
    if (c->MixedSweeps == 0) c->MixedSweeps = (int32_t)c->GaussSeidelIters;
    if (c->MixedFixSweeps == 0) c->MixedFixSweeps = 2;
    if (c->MixedMaxCycles < 1 || c->MixedSweeps < 1 || c->MixedFixSweeps < 1)
        return fail(AF_ERR_CONFIG, "aeroflux: mixed-precision cycles and sweeps must be >= 1");
    if (c->MixedTarget < 0 || std::isnan(c->MixedTarget))
        return fail(AF_ERR_CONFIG, "aeroflux: mixed-precision target must be >= 0");
    return AF_OK;
}

int64_t af_mac_bytes(int64_t nx, int64_t ny, int64_t nz, int64_t ghost)
{
    if (nx < 1 || ny < 1 || nz < 1 || ghost < 1) return 0;
    af_field f[AF_FIELD_COUNT];
    return mac_layout(f, nx, ny, nz, ghost);
}

int af_new(const af_config *cfg_in, af_engine **out)
{
    if (!cfg_in || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *out = nullptr;
    af_config cfg = *cfg_in;
    if (int rc = af_config_normalize(&cfg)) return rc;
    // UseSparseSolver is accepted and changes nothing: New never reads it,
    // and project() takes the direct solve only once EnableSparseSolver has
    // attached a factorisation, which the non-Darwin build never does
    // (aeroflux.go:176-213, project.go:19, sparse_fallback.go:20-24).  The
    // red-black solve runs, exactly as with the flag clear.

    af_engine *e = new (std::nothrow) af_engine();
    if (!e) return fail(AF_ERR_OOM, "aeroflux: out of host memory");
    e->cfg = cfg;
    e->bytes = mac_layout(e->cur.f, cfg.NX, cfg.NY, cfg.NZ, cfg.Ghost);
    for (int i = 0; i < AF_FIELD_COUNT; i++) {
        e->orig_off[i] = e->cur.f[i].Off;
        if (e->cur.f[i].Count >= (int64_t)1 << 31) {
            delete e;
            return fail(AF_ERR_CONFIG,
                        "aeroflux: a %ldx%ldx%ld grid exceeds 2^31 elements per field",
                        (long)cfg.NX, (long)cfg.NY, (long)cfg.NZ);
        }
    }
    e->G.NX = (int)cfg.NX;
    e->G.NY = (int)cfg.NY;
    e->G.NZ = (int)cfg.NZ;
    e->G.g = (int)cfg.Ghost;
    e->invH = 1.0 / cfg.H;                          // stencil.go:47, project.go:216
    e->h2 = cfg.H * cfg.H;                          // project.go:107
    e->div_scale = cfg.Density / (cfg.DT * cfg.H);  // project.go:49
    e->grad_scale = cfg.DT / (cfg.Density * cfg.H); // project.go:179

    hipError_t err = hipGetDevice(&e->device);
    if (err != hipSuccess) {
        delete e;
        return hip_fail(err, "no HIP device");
    }
    e->closed = false;
    auto bail = [&](hipError_t er, const char *what, int code) {
        int rc = code == AF_ERR_OOM
                     ? fail(AF_ERR_OOM, "aeroflux: %s (%ld bytes): %s", what, (long)e->bytes,
                            hipGetErrorString(er))
                     : hip_fail(er, what);
        release(e);
        delete e;
        return rc;
    };
    if ((err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(err, "create stream", AF_ERR_HIP);
    if ((err = hipMalloc(&e->arena, (size_t)e->bytes)) != hipSuccess)
        return bail(err, "allocate arena", AF_ERR_OOM);
    if ((err = hipMemsetAsync(e->arena, 0, (size_t)e->bytes, e->stream)) != hipSuccess)
        return bail(err, "zero arena", AF_ERR_HIP);  // MAC.Zero, grid.go:158-163
    if ((err = hipMalloc(&e->flag, sizeof(int))) != hipSuccess)
        return bail(err, "allocate flag", AF_ERR_OOM);
    if ((err = hipMemsetAsync(e->flag, 0, sizeof(int), e->stream)) != hipSuccess)
        return bail(err, "zero flag", AF_ERR_HIP);
    const int np = reduce_partials_needed(e->G);
    if ((err = hipMalloc(&e->partials, sizeof(double) * (size_t)np)) != hipSuccess)
        return bail(err, "allocate partials", AF_ERR_OOM);
    if ((err = hipMalloc(&e->red_out, sizeof(double) * 4)) != hipSuccess)
        return bail(err, "allocate reduction", AF_ERR_OOM);
    if ((err = hipHostMalloc(&e->flag_host, sizeof(int), hipHostMallocDefault)) != hipSuccess)
        return bail(err, "allocate pinned flag", AF_ERR_OOM);
    if ((err = hipHostMalloc(&e->red_host, sizeof(double) * 4, hipHostMallocDefault)) != hipSuccess)
        return bail(err, "allocate pinned result", AF_ERR_OOM);
    if (cfg.Precision == AF_PRECISION_MIXED) {
        if ((err = init_mixed(e->mixed, e->G, e->stream)) != hipSuccess)
            return bail(err, "allocate mixed-precision buffers", AF_ERR_OOM);
    } else if (cfg.TileK == 0 || cfg.TileK == -2) {
        const size_t n = (size_t)split_count(e->G) * sizeof(double);
        double **split[4] = {&e->ps_red, &e->ps_black, &e->ds_red, &e->ds_black};
        for (double **b : split) {
            if ((err = hipMalloc(b, n)) != hipSuccess)
                return bail(err, "allocate colour-split buffers", AF_ERR_OOM);
            if ((err = hipMemsetAsync(*b, 0, n, e->stream)) != hipSuccess)
                return bail(err, "zero colour-split buffers", AF_ERR_HIP);
        }
    } else if (cfg.TileK > 0) {
        if ((err = hipMalloc(&e->pbuf, (size_t)e->cur.f[AF_FIELD_P].Count * 8)) != hipSuccess)
            return bail(err, "allocate pressure ping-pong buffer", AF_ERR_OOM);
        if ((err = hipMemsetAsync(e->pbuf, 0, (size_t)e->cur.f[AF_FIELD_P].Count * 8,
                                  e->stream)) != hipSuccess)
            return bail(err, "zero pressure ping-pong buffer", AF_ERR_HIP);
        if ((err = rbgs_tb_init()) != hipSuccess)
            return bail(err, "configure red-black kernels", AF_ERR_HIP);
    }
    if ((err = hipStreamSynchronize(e->stream)) != hipSuccess)
        return bail(err, "initialise engine", AF_ERR_HIP);
    int rc = choose_schedule(e);
    if (rc == AF_OK) rc = choose_sweep_shape(e);
    if (rc == AF_OK) rc = choose_pass_slab(e);
    if (rc != AF_OK) {
        const std::string msg = af::last_error();
        release(e);
        delete e;
        return fail(rc, "%s", msg.c_str());
    }
    // Tuning launches wrote only pbuf / the FP32 buffers; restore their zeros.
    if (e->pbuf &&
        (err = hipMemsetAsync(e->pbuf, 0, (size_t)e->cur.f[AF_FIELD_P].Count * 8, e->stream)) !=
            hipSuccess)
        return bail(err, "zero pressure ping-pong buffer", AF_ERR_HIP);
    if ((err = hipStreamSynchronize(e->stream)) != hipSuccess)
        return bail(err, "initialise engine", AF_ERR_HIP);
    *out = e;
    return AF_OK;
}

int af_close(af_engine *e)
{
    if (!e) return fail(AF_ERR_ARG, "aeroflux: null engine");
    release(e);
    return AF_OK;
}

void af_free(af_engine *e)
{
    if (!e) return;
    release(e);
    delete e;
}

int af_get_config(const af_engine *e, af_config *out)
{
    if (!e || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *out = e->cfg;
    return AF_OK;
}

int af_step(af_engine *e) { return af_step_n(e, 1, nullptr); }

int af_step_n(af_engine *e, int64_t n, int64_t *first_fault_step)
{
    if (int rc = check_open(e)) return rc;
    if (n < 0) return fail(AF_ERR_ARG, "aeroflux: negative step count");
    HIP_TRY(hipSetDevice(e->device), "set device");
    hipGraphExec_t g;
    if (int rc = get_graph(e, &g)) return rc;
    // The flag is cleared once; a fault in step k stays raised.  To name the
    // first faulting step, the per-step flag is read only when one is asked
    // for and n > 1.
    const bool per_step = first_fault_step && n > 1;
    HIP_TRY(launch_set_int(e->flag, 0, e->stream), "clear fault flag");
    for (int64_t k = 0; k < n; k++) {
        HIP_TRY(hipGraphLaunch(g, e->stream), "launch step graph");
        if (per_step) {
            HIP_TRY(hipMemcpyAsync(e->flag_host, e->flag, sizeof(int), hipMemcpyDeviceToHost,
                                   e->stream),
                    "read fault flag");
            if (int rc = sync_stream(e, "step")) return rc;
            if (*e->flag_host) {
                *first_fault_step = k;
                return fail(AF_ERR_FAULT,
                            "aeroflux: device fault contained: non-finite departure coordinate "
                            "in an advection sweep (step %ld)",
                            (long)k);
            }
        }
    }
    HIP_TRY(hipMemcpyAsync(e->flag_host, e->flag, sizeof(int), hipMemcpyDeviceToHost, e->stream),
            "read fault flag");
    if (int rc = sync_stream(e, "step")) return rc;
    if (*e->flag_host) {
        if (first_fault_step) *first_fault_step = 0;
        return fail(AF_ERR_FAULT,
                    "aeroflux: device fault contained: non-finite departure coordinate in an "
                    "advection sweep");
    }
    return AF_OK;
}

int af_max_divergence(af_engine *e, double *out)
{
    if (int rc = check_open(e)) return rc;
    if (!out) return fail(AF_ERR_ARG, "aeroflux: null output");
    HIP_TRY(hipSetDevice(e->device), "set device");
    const Desc &d = e->cur;
    HIP_TRY(launch_max_divergence(dfield(e, d.f[AF_FIELD_U]), dfield(e, d.f[AF_FIELD_V]),
                                  dfield(e, d.f[AF_FIELD_W]), e->G, e->invH, e->partials,
                                  e->red_out, e->stream),
            "max divergence");
    return run_reduction(e, out);
}

int af_kinetic_energy(af_engine *e, double *out)
{
    if (int rc = check_open(e)) return rc;
    if (!out) return fail(AF_ERR_ARG, "aeroflux: null output");
    HIP_TRY(hipSetDevice(e->device), "set device");
    const Desc &d = e->cur;
    HIP_TRY(launch_kinetic_energy(dfield(e, d.f[AF_FIELD_U]), dfield(e, d.f[AF_FIELD_V]),
                                  dfield(e, d.f[AF_FIELD_W]), e->G, e->partials, e->red_out,
                                  e->stream),
            "kinetic energy");
    double sum;
    if (int rc = run_reduction(e, &sum)) return rc;
    const double vol = e->cfg.H * e->cfg.H * e->cfg.H;
    *out = 0.5 * sum * vol;  // project.go:257
    return AF_OK;
}

int af_max_velocity(af_engine *e, double *out)
{
    if (int rc = check_open(e)) return rc;
    if (!out) return fail(AF_ERR_ARG, "aeroflux: null output");
    HIP_TRY(hipSetDevice(e->device), "set device");
    const Desc &d = e->cur;
    const DField U = dfield(e, d.f[AF_FIELD_U]), V = dfield(e, d.f[AF_FIELD_V]),
                 W = dfield(e, d.f[AF_FIELD_W]);
    HIP_TRY(launch_max_abs3(U.d, d.f[AF_FIELD_U].Count, V.d, d.f[AF_FIELD_V].Count, W.d,
                            d.f[AF_FIELD_W].Count, e->partials, e->red_out, e->stream),
            "max velocity");
    return run_reduction(e, out);
}

int af_cfl(af_engine *e, double *out)
{
    double mv;
    if (int rc = af_max_velocity(e, &mv)) return rc;
    *out = mv * e->cfg.DT / e->cfg.H;  // stencil.go:337
    return AF_OK;
}

int af_enable_sparse_solver(af_engine *e)
{
    (void)e;
    return fail(AF_ERR_UNSUPPORTED,
                "aeroflux: sparse direct solver requires darwin with cgo (Accelerate); the "
                "Red-Black iterative solver is used instead");
}

int af_enable_direct_solver(af_engine *e)
{
    return af_enable_direct_solver_method(e, DIRECT_AUTO);
}

int af_enable_direct_solver_method(af_engine *e, int32_t method)
{
    if (int rc = check_open(e)) return rc;
    if (method < DIRECT_AUTO || method > DIRECT_DENSE)
        return fail(AF_ERR_ARG, "aeroflux: direct solver method %d is not 0, 1 or 2", (int)method);
    if (method == DIRECT_FFT && !direct_fft_supported(e->G))
        return fail(AF_ERR_UNSUPPORTED,
                    "aeroflux: the FFT direct solve needs every extent a power of two in 8..512 "
                    "(grid %dx%dx%d)",
                    e->G.NX, e->G.NY, e->G.NZ);
    const int want = method == DIRECT_AUTO
                         ? (direct_fft_supported(e->G) ? DIRECT_FFT : DIRECT_DENSE)
                         : method;
    if (e->direct.method == want) return AF_OK;
    HIP_TRY(hipSetDevice(e->device), "set device");
    HIP_TRY(hipStreamSynchronize(e->stream), "enable direct solver");
    release_direct(e->direct);
    int bs = 0;
    const hipError_t err = init_direct(e->direct, e->G, e->cfg.H, want, e->stream, &bs);
    if (err != hipSuccess) {
        release_direct(e->direct);
        if (bs != 0) return fail(AF_ERR_HIP, "aeroflux: direct solver: BLAS status %d", bs);
        return hip_fail(err, "direct solver");
    }
    // The step graphs were captured with the red-black solve.
    for (auto &g : e->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return AF_OK;
}

int af_direct_solver_enabled(const af_engine *e, int32_t *on)
{
    if (!e || !on) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *on = e->direct.method;
    return AF_OK;
}

int af_field_desc(const af_engine *e, int id, af_field *out)
{
    if (!e || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    if (id < 0 || id >= AF_FIELD_COUNT) return fail(AF_ERR_ARG, "aeroflux: bad field id %d", id);
    *out = e->cur.f[id];
    return AF_OK;
}

int af_field_read(af_engine *e, int id, double *host, int64_t count)
{
    if (int rc = check_open(e)) return rc;
    if (id < 0 || id >= AF_FIELD_COUNT || !host)
        return fail(AF_ERR_ARG, "aeroflux: bad field read arguments");
    const af_field &f = e->cur.f[id];
    if (count != f.Count)
        return fail(AF_ERR_ARG, "aeroflux: field %d has %ld elements, got %ld", id, (long)f.Count,
                    (long)count);
    HIP_TRY(hipSetDevice(e->device), "set device");
    HIP_TRY(hipMemcpyAsync(host, e->arena + f.Off, (size_t)count * 8, hipMemcpyDeviceToHost,
                           e->stream),
            "read field");
    return sync_stream(e, "read field");
}

int af_field_write(af_engine *e, int id, const double *host, int64_t count)
{
    if (int rc = check_open(e)) return rc;
    if (id < 0 || id >= AF_FIELD_COUNT || !host)
        return fail(AF_ERR_ARG, "aeroflux: bad field write arguments");
    const af_field &f = e->cur.f[id];
    if (count != f.Count)
        return fail(AF_ERR_ARG, "aeroflux: field %d has %ld elements, got %ld", id, (long)f.Count,
                    (long)count);
    HIP_TRY(hipSetDevice(e->device), "set device");
    HIP_TRY(hipMemcpyAsync(e->arena + f.Off, host, (size_t)count * 8, hipMemcpyHostToDevice,
                           e->stream),
            "write field");
    return sync_stream(e, "write field");
}

int af_field_device_ptr(const af_engine *e, int id, uint64_t *out)
{
    if (!e || !out) return fail(AF_ERR_ARG, "aeroflux: null argument");
    if (e->closed) return fail(AF_ERR_CLOSED, "aeroflux: engine is closed");
    if (id < 0 || id >= AF_FIELD_COUNT) return fail(AF_ERR_ARG, "aeroflux: bad field id %d", id);
    *out = (uint64_t)(uintptr_t)(e->arena + e->cur.f[id].Off);
    return AF_OK;
}

int64_t af_arena_bytes(const af_engine *e) { return e ? e->bytes : 0; }

int af_snapshot_bytes(const af_engine *e, int64_t *bytes)
{
    if (int rc = check_open(e)) return rc;
    if (!bytes) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *bytes = snapshot_bytes(e->cur.f);
    return AF_OK;
}

int af_snapshot(af_engine *e, void *dst, int64_t bytes)
{
    if (int rc = check_open(e)) return rc;
    HIP_TRY(hipSetDevice(e->device), "set device");
    return snapshot_save(e->cur.f, e->arena, e->cfg, dst, bytes,
                         reinterpret_cast<unsigned long long *>(e->red_out), e->stream);
}

int af_restore(af_engine *e, const void *src, int64_t bytes)
{
    if (int rc = check_open(e)) return rc;
    HIP_TRY(hipSetDevice(e->device), "set device");
    return snapshot_load(e->cur.f, e->arena, e->cfg, src, bytes,
                         reinterpret_cast<unsigned long long *>(e->red_out), e->stream);
}

// ---- stage entry points ------------------------------------------------

#define STAGE_BEGIN(e)                                  \
    if (int rc = check_open(e)) return rc;              \
    HIP_TRY(hipSetDevice((e)->device), "set device")

int af_apply_velocity_bc(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(enq_velocity_bc(e, e->cur), "velocity bc");
    return sync_stream(e, "velocity bc");
}

int af_sweep_axis(af_engine *e, int axis, double dt)
{
    STAGE_BEGIN(e);
    if (axis < 0 || axis > 2) return fail(AF_ERR_ARG, "aeroflux: bad axis %d", axis);
    HIP_TRY(launch_set_int(e->flag, 0, e->stream), "clear fault flag");
    HIP_TRY(enq_sweep(e, e->cur, axis, dt), "advection sweep");
    HIP_TRY(hipMemcpyAsync(e->flag_host, e->flag, sizeof(int), hipMemcpyDeviceToHost, e->stream),
            "read fault flag");
    if (int rc = sync_stream(e, "advection sweep")) return rc;
    if (*e->flag_host)
        return fail(AF_ERR_FAULT,
                    "aeroflux: device fault contained: non-finite departure coordinate in an "
                    "advection sweep");
    return AF_OK;
}

int af_commit_velocity(af_engine *e)
{
    if (int rc = check_open(e)) return rc;
    commit(e->cur);
    return AF_OK;
}

int af_compute_divergence(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(enq_divergence(e, e->cur), "divergence");
    return sync_stream(e, "divergence");
}

int af_gs_sweep(af_engine *e, int color)
{
    STAGE_BEGIN(e);
    if (color != 0 && color != 1) return fail(AF_ERR_ARG, "aeroflux: bad colour %d", color);
    HIP_TRY(launch_rbgs(dfield(e, e->cur.f[AF_FIELD_P]), dfield(e, e->cur.f[AF_FIELD_DIV]), e->G,
                        e->h2, color, false, e->stream),
            "gauss-seidel sweep");
    return sync_stream(e, "gauss-seidel sweep");
}

int af_apply_pressure_bc(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(launch_pressure_bc(dfield(e, e->cur.f[AF_FIELD_P]), e->G, false, e->stream),
            "pressure bc");
    return sync_stream(e, "pressure bc");
}

int af_solve_pressure(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(enq_solve(e, e->cur), "pressure solve");
    return sync_stream(e, "pressure solve");
}

int af_subtract_gradient(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(enq_gradient(e, e->cur), "gradient");
    return sync_stream(e, "gradient");
}

int af_project(af_engine *e)
{
    STAGE_BEGIN(e);
    HIP_TRY(enq_project(e, e->cur), "project");
    return sync_stream(e, "project");
}

int af_init_band_limited(af_engine *e, uint64_t seed, int32_t kmax, double slope, double rms)
{
    STAGE_BEGIN(e);
    if (kmax < 1 || kmax > 64 || !(rms > 0) || std::isinf(rms) || std::isnan(slope))
        return fail(AF_ERR_ARG, "aeroflux: band-limited field needs 1 <= kmax <= 64, rms > 0");
    const Desc &d = e->cur;
    const DField U = dfield(e, d.f[AF_FIELD_U]), V = dfield(e, d.f[AF_FIELD_V]),
                 W = dfield(e, d.f[AF_FIELD_W]);
    HIP_TRY(launch_zero(U.d, d.f[AF_FIELD_U].Count, e->stream), "zero velocity");
    HIP_TRY(launch_zero(V.d, d.f[AF_FIELD_V].Count, e->stream), "zero velocity");
    HIP_TRY(launch_zero(W.d, d.f[AF_FIELD_W].Count, e->stream), "zero velocity");
    HIP_TRY(launch_band_limited(U, V, W, e->G, e->cfg.H, seed, kmax, slope, rms, e->ic,
                                e->stream),
            "band-limited initial condition");
    HIP_TRY(enq_velocity_bc(e, d), "velocity bc");
    return sync_stream(e, "band-limited initial condition");
}

int af_solver_tile(const af_engine *e, int32_t *k, int32_t *rows, int32_t *zchunk)
{
    if (!e) return fail(AF_ERR_ARG, "aeroflux: null engine");
    if (k) *k = e->tile64.k;
    if (rows) *rows = e->tile64.rows;
    if (zchunk) *zchunk = e->tile64.cz;
    return AF_OK;
}

int af_sweep_shape(const af_engine *e, int32_t *shape, int32_t *rows, int32_t *zper)
{
    if (!e) return fail(AF_ERR_ARG, "aeroflux: null engine");
    // Gathering shapes: rows x z per thread; LDS-staged: rows x the largest
    // z-chunk (the NZ+1 face planes split into chunks of sizes that differ
    // by at most one).
    static const int kRows[] = {1, 1, 2, 2, 4, 4, 8, 16, 4, 4, 8, 8, 8, 8};
    static const int kZ[] = {1, 2, 1, 2, 1, 2, 1, 1, 16, 32, 16, 32, 16, 32};
    static_assert(sizeof(kRows) == sizeof(kZ), "one entry per shape");
    if (e->sweep_shape < 0 || e->sweep_shape >= (int)(sizeof(kRows) / sizeof(kRows[0])) ||
        e->sweep_shape >= sweep_shape_count())
        return fail(AF_ERR_ARG, "aeroflux: sweep shape %d has no table entry", e->sweep_shape);
    if (shape) *shape = e->sweep_shape;
    if (rows) *rows = kRows[e->sweep_shape];
    if (zper) {
        *zper = kZ[e->sweep_shape];
        if (e->sweep_shape >= 8) {
            const int n = e->sweep_chunks > 0 ? e->sweep_chunks
                                              : sweep_lds_default_chunks(e->G, e->sweep_shape - 8);
            *zper = (e->G.NZ + 1 + n - 1) / n;
        }
    }
    return AF_OK;
}

int af_sweep_slab(const af_engine *e, int32_t *rows)
{
    if (!e || !rows) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *rows = e->sweep_slab >= (1 << 30) ? -1 : e->sweep_slab;
    return AF_OK;
}

int af_pass_slab(const af_engine *e, int32_t *rows)
{
    if (!e || !rows) return fail(AF_ERR_ARG, "aeroflux: null argument");
    *rows = e->pass_slab >= (1 << 30) ? -1 : e->pass_slab;
    return AF_OK;
}

int af_mixed_stats(af_engine *e, double *last_div, int32_t *cycles)
{
    STAGE_BEGIN(e);
    if (e->cfg.Precision != AF_PRECISION_MIXED)
        return fail(AF_ERR_ARG, "aeroflux: engine is not in mixed precision");
    double d = 0;
    int c = 0;
    HIP_TRY(hipMemcpyAsync(e->red_host, e->mixed.last_res, sizeof(double), hipMemcpyDeviceToHost,
                           e->stream),
            "read mixed stats");
    if (int rc = sync_stream(e, "mixed stats")) return rc;
    d = *e->red_host;
    HIP_TRY(hipMemcpyAsync(e->flag_host, e->mixed.cycles, sizeof(int), hipMemcpyDeviceToHost,
                           e->stream),
            "read mixed stats");
    if (int rc = sync_stream(e, "mixed stats")) return rc;
    c = *e->flag_host;
    if (last_div) *last_div = d;
    if (cycles) *cycles = c;
    return AF_OK;
}

int af_time_kernel(af_engine *e, int kernel, int reps, double *avg_ms)
{
    STAGE_BEGIN(e);
    if (reps < 1 || !avg_ms) return fail(AF_ERR_ARG, "aeroflux: bad timing arguments");
    const Desc &d = e->cur;
    const DField P = dfield(e, d.f[AF_FIELD_P]), D = dfield(e, d.f[AF_FIELD_DIV]);
    const bool mixed = e->cfg.Precision == AF_PRECISION_MIXED;
    const bool split = e->tile64.k == 0 && !mixed;
    SplitArgs<float> a32;
    if (mixed) {
        a32.red = e->mixed.e_red;
        a32.black = e->mixed.e_black;
        a32.rhs_red = e->mixed.r_red;
        a32.rhs_black = e->mixed.r_black;
        a32.NX = e->G.NX; a32.NY = e->G.NY; a32.NZ = e->G.NZ; a32.NXH = (e->G.NX + 1) / 2;
        a32.g = e->G.g;
        a32.h2 = 0;
        a32.slab = e->pass_slab;
    }
    auto one = [&]() -> hipError_t {
        switch (kernel) {
        case 0:
            if (mixed) return launch_split_pass<float>(a32, AF_COLOR_RED, 0, e->stream);
            if (split) return launch_split_pass<double>(split_args(e), AF_COLOR_RED, 0, e->stream);
            if (e->tile64.k > 0) {
                TbArgs<double> a = tb_args(e, P, D);
                a.dst = e->pbuf;
                return launch_rbgs_tb<double>(a, e->tile64.k, e->tile64.rows, false, e->stream);
            }
            return launch_rbgs(P, D, e->G, e->h2, AF_COLOR_RED, true, e->stream);
        case 1: return enq_sweep(e, d, AF_AXIS_X, e->cfg.DT * 0.5, true);
        case 4: return enq_sweep(e, d, AF_AXIS_Y, e->cfg.DT * 0.5, true);
        case 5: return enq_sweep(e, d, AF_AXIS_Z, e->cfg.DT * 0.5, true);
        case 2: return enq_divergence(e, d);
        case 3: return enq_gradient(e, d);
        case 6: return enq_solve(e, d);
        default: return hipErrorInvalidValue;
        }
    };
    if (kernel < 0 || kernel > 6) return fail(AF_ERR_ARG, "aeroflux: bad kernel id %d", kernel);
    HIP_TRY(one(), "warm-up launch");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), "create event");
    HIP_TRY(hipEventCreate(&b), "create event");
    hipError_t err = hipEventRecord(a, e->stream);
    for (int r = 0; r < reps && err == hipSuccess; r++) err = one();
    if (err == hipSuccess) err = hipEventRecord(b, e->stream);
    if (err == hipSuccess) err = hipEventSynchronize(b);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (err != hipSuccess) return hip_fail(err, "time kernel");
    *avg_ms = (double)ms / reps;
    return AF_OK;
}

}  // extern "C"
