/*
 * aeroflux_oracle.c -- CPU restatement of AeroFlux's per-step data plane.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to check results or to
 * time the reference's algorithm; the product library (aeroflux_amd/) never
 * links or calls it.
 *
 * It restates, function by function and in the same floating-point operation
 * order, the Go reference:
 *   Config.Normalize        AeroFlex/aeroflux.go:47-113
 *   partition               AeroFlex/aeroflux.go:218-232
 *   dispatch (join=barrier) AeroFlex/aeroflux.go:313-324
 *   Step                    AeroFlex/aeroflux.go:348-409
 *   fieldLayout / NewMAC    AeroFlex/grid.go:79-134
 *   advectDiffuseAxis       AeroFlex/stencil.go:21-25
 *   advectComponent         AeroFlex/stencil.go:44-119
 *   faceVelocity            AeroFlex/stencil.go:128-155
 *   sampleTrilinear         AeroFlex/stencil.go:169-225
 *   commitVelocity          AeroFlex/stencil.go:230-235
 *   applyVelocityBC         AeroFlex/stencil.go:246-315
 *   MaxVelocity / CFL       AeroFlex/stencil.go:319-338
 *   project                 AeroFlex/project.go:12-34
 *   computeDivergence       AeroFlex/project.go:42-62
 *   solvePressureRedBlack   AeroFlex/project.go:73-100
 *   gaussSeidelSweep        AeroFlex/project.go:103-130
 *   applyPressureBC         AeroFlex/project.go:139-163
 *   subtractGradient        AeroFlex/project.go:172-205
 *   MaxDivergence           AeroFlex/project.go:210-232
 *   KineticEnergy           AeroFlex/project.go:238-258
 *
 * Build with -O2 -ffp-contract=off (no fused multiply-add, no fast-math):
 * "the Go reference" is taken to be its unfused IEEE-754 binary64 evaluation
 * (amd64 at GOAMD64=v1 has no FMA instructions).
 *
 * Fault path.  On amd64 a NaN departure coordinate reaches int(NaN), an
 * out-of-range index, and the slice bounds check panics inside the worker;
 * runGuarded turns that into an error (aeroflux.go:286-296, arena.go:230-243)
 * and Step returns it right after the failing dispatch (aeroflux.go:376-378).
 * The restatement stops that shard's loop at the faulting face, records the
 * fault, and makes Step return right after that dispatch.
 *
 * Pinning: every function here is checked by tests/test_oracle.py against
 * what the mathematics fixes (closed forms on linear, quadratic, uniform and
 * eigenmode fields, the summation-by-parts identity, planted extrema and
 * NaNs, the stated Step sequence) and the reference's own test invariants;
 * each planted-mistake build (-DORC_MUTANT=n) fails at least one closed
 * form.  An independent numpy restatement is a further cross-check.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_CACHE_LINE 128 /* layout.go:13 */

/* Planted-mistake builds for the mutation tests (tests/test_oracle.py):
 * -DORC_MUTANT=n compiles one deliberate transcription error so the tests
 * can show their checks catch it.  The default build defines none. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#if ORC_MUTANT == 1
#define NEG(v) (0 - (v))
#else
#define NEG(v) (-(v))
#endif

typedef struct {
    int64_t NX, NY, NZ, Ghost;
    double H, Viscosity, Density, DT;
    int64_t Workers, GaussSeidelIters;
} orc_config;

typedef struct {
    int64_t off; /* byte offset of (0,0,0) in the arena */
    int64_t nx, ny, nz, sy, sz, count;
} orc_field;

enum { F_P = 0, F_U, F_V, F_W, F_DIV, F_TMP0, F_TMP1, F_TMP2, F_COUNT };
enum { AX_X = 0, AX_Y, AX_Z };
enum { C_U = 0, C_V, C_W };
enum { OP_SWEEP = 0, OP_DIV, OP_GRAD, OP_RED, OP_BLACK, OP_STOP };

typedef struct {
    int64_t z0, z1;
    int id;
} orc_shard;

typedef struct orc_engine orc_engine;

typedef struct {
    orc_engine *e;
    int id;
} orc_worker_arg;

struct orc_engine {
    orc_config cfg;
    int64_t NX, NY, NZ, G;
    double H;
    orc_field f[F_COUNT];
    int64_t bytes;
    unsigned char *arena;

    int nworkers;
    orc_shard *shards;
    pthread_t *threads;
    orc_worker_arg *args;
    pthread_barrier_t bar_start, bar_end;
    int started;

    /* Phase parameters carried to workers through the engine, as
     * aeroflux.go:152-158 does. */
    int op;
    int cur_axis;
    double cur_dt;
    int *shard_fault; /* one slot per worker, drained by dispatch */
    int step_fault;
};

/* ------------------------------------------------------------------ config */

static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* Config.Normalize, aeroflux.go:47-113.  Returns 0 or an error code; the
 * message (prefixed "aeroflux: ") is written into err. */
int orc_normalize(orc_config *c, char *err, size_t n)
{
    if (c->NX <= 0 || c->NY <= 0 || c->NZ <= 0) {
        snprintf(err, n, "aeroflux: interior extents must be positive, got %ldx%ldx%ld",
                 (long)c->NX, (long)c->NY, (long)c->NZ);
        return 1;
    }
    if (c->Ghost == 0) c->Ghost = 1;
    if (c->Ghost < 1) {
        snprintf(err, n, "aeroflux: ghost thickness must be >= 1, got %ld", (long)c->Ghost);
        return 1;
    }
    if (c->H == 0) c->H = 1.0 / (double)max64(c->NX, max64(c->NY, c->NZ));
    if (c->H <= 0 || isnan(c->H) || isinf(c->H)) {
        snprintf(err, n, "aeroflux: cell size must be finite and positive, got %g", c->H);
        return 1;
    }
    if (c->Density == 0) c->Density = 1.0;
    if (c->Density <= 0 || isnan(c->Density) || isinf(c->Density)) {
        snprintf(err, n, "aeroflux: density must be finite and positive, got %g", c->Density);
        return 1;
    }
    if (c->Viscosity < 0 || isnan(c->Viscosity) || isinf(c->Viscosity)) {
        snprintf(err, n, "aeroflux: viscosity must be finite and non-negative, got %g", c->Viscosity);
        return 1;
    }
    if (c->DT == 0) {
        if (c->Viscosity > 0)
            c->DT = 0.5 * c->H * c->H / (6 * c->Viscosity);
        else
            c->DT = 0.25 * c->H;
    }
    if (c->DT <= 0 || isnan(c->DT) || isinf(c->DT)) {
        snprintf(err, n, "aeroflux: timestep must be finite and positive, got %g", c->DT);
        return 1;
    }
    if (c->Workers == 0) c->Workers = (int64_t)sysconf(_SC_NPROCESSORS_ONLN);
    if (c->Workers < 1) {
        snprintf(err, n, "aeroflux: worker count must be >= 1, got %ld", (long)c->Workers);
        return 1;
    }
    if (c->Workers > c->NZ) c->Workers = c->NZ;
    if (c->GaussSeidelIters == 0) c->GaussSeidelIters = 24;
    if (c->GaussSeidelIters < 1) {
        snprintf(err, n, "aeroflux: gauss-seidel iterations must be >= 1, got %ld",
                 (long)c->GaussSeidelIters);
        return 1;
    }
    if (c->Viscosity > 0) {
        double limit = c->H * c->H / (6 * c->Viscosity);
        if (c->DT > limit) {
            snprintf(err, n,
                     "aeroflux: dt %g exceeds explicit diffusion stability limit %g (h=%g, nu=%g); reduce dt or h",
                     c->DT, limit, c->H, c->Viscosity);
            return 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ layout */

/* fieldLayout, grid.go:79-97 */
static orc_field field_layout(int64_t *off, int64_t nx, int64_t ny, int64_t nz)
{
    orc_field f;
    f.off = *off;
    f.nx = nx; f.ny = ny; f.nz = nz;
    f.sy = nx; f.sz = nx * ny;
    f.count = nx * ny * nz;
    int64_t sz = f.count * 8;
    int64_t r = sz % ORC_CACHE_LINE;
    if (r != 0) sz += ORC_CACHE_LINE - r;
    *off += sz;
    return f;
}

/* NewMAC field order, grid.go:114-128 */
static int64_t mac_layout(orc_field *f, int64_t nx, int64_t ny, int64_t nz, int64_t g)
{
    int64_t g2 = 2 * g, cx = nx + g2, cy = ny + g2, cz = nz + g2, off = 0;
    f[F_P] = field_layout(&off, cx, cy, cz);
    f[F_U] = field_layout(&off, cx + 1, cy, cz);
    f[F_V] = field_layout(&off, cx, cy + 1, cz);
    f[F_W] = field_layout(&off, cx, cy, cz + 1);
    f[F_DIV] = field_layout(&off, cx, cy, cz);
    f[F_TMP0] = field_layout(&off, cx + 1, cy, cz);
    f[F_TMP1] = field_layout(&off, cx, cy + 1, cz);
    f[F_TMP2] = field_layout(&off, cx, cy, cz + 1);
    return off;
}

/* MACBytes, grid.go:138-144 */
int64_t orc_mac_bytes(int64_t nx, int64_t ny, int64_t nz, int64_t g)
{
    orc_field f[F_COUNT];
    if (nx < 1 || ny < 1 || nz < 1 || g < 1) return 0;
    return mac_layout(f, nx, ny, nz, g);
}

static inline int64_t IDX(const orc_field *f, int64_t x, int64_t y, int64_t z)
{
    return x + y * f->sy + z * f->sz; /* Field.Index, grid.go:24-26 */
}

static inline double *DATA(orc_engine *e, int id)
{
    return (double *)(e->arena + e->f[id].off);
}

/* ------------------------------------------------------------- advection */

/* faceVelocity, stencil.go:128-155 */
static inline void face_velocity(orc_engine *e, int c, int64_t x, int64_t y, int64_t z,
                                 double *vx, double *vy, double *vz)
{
    const orc_field *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W];
    const double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);
    switch (c) {
    case C_U:
        *vx = u[IDX(U, x, y, z)];
        *vy = 0.25 * (v[IDX(V, x - 1, y, z)] + v[IDX(V, x, y, z)] +
                      v[IDX(V, x - 1, y + 1, z)] + v[IDX(V, x, y + 1, z)]);
        *vz = 0.25 * (w[IDX(W, x - 1, y, z)] + w[IDX(W, x, y, z)] +
                      w[IDX(W, x - 1, y, z + 1)] + w[IDX(W, x, y, z + 1)]);
        break;
    case C_V:
#if ORC_MUTANT == 2
        *vx = 0.25 * (u[IDX(U, x, y - 1, z)] + u[IDX(U, x - 1, y - 1, z)] +
#else
        *vx = 0.25 * (u[IDX(U, x, y - 1, z)] + u[IDX(U, x + 1, y - 1, z)] +
#endif
                      u[IDX(U, x, y, z)] + u[IDX(U, x + 1, y, z)]);
        *vy = v[IDX(V, x, y, z)];
        *vz = 0.25 * (w[IDX(W, x, y - 1, z)] + w[IDX(W, x, y, z)] +
                      w[IDX(W, x, y - 1, z + 1)] + w[IDX(W, x, y, z + 1)]);
        break;
    default:
        *vx = 0.25 * (u[IDX(U, x, y, z - 1)] + u[IDX(U, x + 1, y, z - 1)] +
                      u[IDX(U, x, y, z)] + u[IDX(U, x + 1, y, z)]);
        *vy = 0.25 * (v[IDX(V, x, y, z - 1)] + v[IDX(V, x, y + 1, z - 1)] +
                      v[IDX(V, x, y, z)] + v[IDX(V, x, y + 1, z)]);
        *vz = w[IDX(W, x, y, z)];
        break;
    }
}

/* sampleTrilinear, stencil.go:169-225.  Returns 0 and stores the sample, or
 * returns 1 for the case the reference turns into a bounds panic: a NaN
 * coordinate survives every clamp comparison and int(NaN) is out of range. */
static inline int sample_trilinear(const orc_field *f, const double *d, double px, double py,
                                   double pz, double *out)
{
    double maxX = (double)(f->nx - 1);
    double maxY = (double)(f->ny - 1);
    double maxZ = (double)(f->nz - 1);

    if (px < 0) px = 0; else if (px > maxX) px = maxX;
    if (py < 0) py = 0; else if (py > maxY) py = maxY;
    if (pz < 0) pz = 0; else if (pz > maxZ) pz = maxZ;

    if (isnan(px) || isnan(py) || isnan(pz)) return 1;

    int64_t x0 = (int64_t)px, y0 = (int64_t)py, z0 = (int64_t)pz;
    if (x0 >= f->nx - 1) x0 = f->nx - 2;
    if (y0 >= f->ny - 1) y0 = f->ny - 2;
    if (z0 >= f->nz - 1) z0 = f->nz - 2;

    double tx = px - (double)x0;
    double ty = py - (double)y0;
    double tz = pz - (double)z0;

    int64_t i000 = IDX(f, x0, y0, z0);
    int64_t i100 = i000 + 1;
    int64_t i010 = i000 + f->sy;
    int64_t i110 = i010 + 1;
    int64_t i001 = i000 + f->sz;
    int64_t i101 = i001 + 1;
    int64_t i011 = i001 + f->sy;
    int64_t i111 = i011 + 1;

#if ORC_MUTANT == 6
    double c00 = d[i000] * tx + d[i100] * (1 - tx);
#else
    double c00 = d[i000] * (1 - tx) + d[i100] * tx;
#endif
    double c10 = d[i010] * (1 - tx) + d[i110] * tx;
    double c01 = d[i001] * (1 - tx) + d[i101] * tx;
    double c11 = d[i011] * (1 - tx) + d[i111] * tx;

    double c0 = c00 * (1 - ty) + c10 * ty;
    double c1 = c01 * (1 - ty) + c11 * ty;

#if ORC_MUTANT == 13
    if (ty == 0) c0 = c00;
    if (tz == 0) {
        *out = c0;
        return 0;
    }
#endif
    *out = c0 * (1 - tz) + c1 * tz;
    return 0;
}

/* advectComponent, stencil.go:44-119.  Returns 1 on a contained fault. */
static int advect_component(orc_engine *e, orc_shard s, int src_id, int dst_id, int c, int ax,
                            double dt)
{
    const orc_field *src = &e->f[src_id], *dst = &e->f[dst_id];
    double h = e->H;
    double invH = 1.0 / h;
    double nu = e->cfg.Viscosity;
    const double *sd = DATA(e, src_id);
    double *dd = DATA(e, dst_id);

    int64_t x0 = e->G, x1 = e->G + e->NX;
    int64_t y0 = e->G, y1 = e->G + e->NY;
    int64_t z0 = s.z0, z1 = s.z1;
    switch (c) {
    case C_U: x1++; break;
    case C_V: y1++; break;
    default:
        if (s.z1 == e->G + e->NZ) z1++;
        break;
    }

    double lam = nu * dt * invH * invH;

    for (int64_t z = z0; z < z1; z++) {
        for (int64_t y = y0; y < y1; y++) {
            int64_t base = IDX(src, 0, y, z);
            int64_t dbase = IDX(dst, 0, y, z);
            for (int64_t x = x0; x < x1; x++) {
                int64_t i = base + x;
                double vx, vy, vz;
                face_velocity(e, c, x, y, z, &vx, &vy, &vz);

                double px = (double)x, py = (double)y, pz = (double)z;
                switch (ax) {
                case AX_X: px -= dt * vx * invH; break;
                case AX_Y: py -= dt * vy * invH; break;
                default: pz -= dt * vz * invH; break;
                }

                double adv;
                if (sample_trilinear(src, sd, px, py, pz, &adv)) return 1;

                double lap = 0;
                if (lam != 0) {
                    int64_t im, ip;
                    switch (ax) {
                    case AX_X: im = i - 1; ip = i + 1; break;
                    case AX_Y: im = i - src->sy; ip = i + src->sy; break;
                    default: im = i - src->sz; ip = i + src->sz; break;
                    }
#if ORC_MUTANT == 3
                    lap = lam * (sd[im] - 2 * sd[i]);
                    (void)ip;
#else
                    lap = lam * (sd[im] - 2 * sd[i] + sd[ip]);
#endif
                }
                dd[dbase + x] = adv + lap;
            }
        }
    }
    return 0;
}

/* advectDiffuseAxis, stencil.go:21-25.  Like the Go worker, a fault stops
 * this shard's remaining work (the panic unwinds the whole call). */
static int advect_diffuse_axis(orc_engine *e, orc_shard s, int ax, double dt)
{
    if (advect_component(e, s, F_U, F_TMP0, C_U, ax, dt)) return 1;
    if (advect_component(e, s, F_V, F_TMP1, C_V, ax, dt)) return 1;
    if (advect_component(e, s, F_W, F_TMP2, C_W, ax, dt)) return 1;
    return 0;
}

/* commitVelocity, stencil.go:230-235: descriptor swap, no copy. */
void orc_commit_velocity(orc_engine *e)
{
    orc_field t;
    t = e->f[F_U]; e->f[F_U] = e->f[F_TMP0]; e->f[F_TMP0] = t;
    t = e->f[F_V]; e->f[F_V] = e->f[F_TMP1]; e->f[F_TMP1] = t;
    t = e->f[F_W]; e->f[F_W] = e->f[F_TMP2]; e->f[F_TMP2] = t;
}

/* applyVelocityBC, stencil.go:246-315 (serial, driver thread). */
void orc_apply_velocity_bc(orc_engine *e)
{
    int64_t g = e->G;
    const orc_field *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W];
    double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);

    int64_t ux0 = g, ux1 = g + e->NX;
    int64_t y0 = g, y1 = g + e->NY;
    int64_t z0 = g, z1 = g + e->NZ;

    for (int64_t z = z0; z < z1; z++)
        for (int64_t y = y0; y < y1; y++) {
            u[IDX(U, ux0, y, z)] = 0;
            u[IDX(U, ux1, y, z)] = 0;
        }
    for (int64_t z = z0; z < z1; z++)
        for (int64_t x = g; x < g + e->NX; x++) {
            v[IDX(V, x, y0, z)] = 0;
            v[IDX(V, x, y1, z)] = 0;
        }
    for (int64_t y = y0; y < y1; y++)
        for (int64_t x = g; x < g + e->NX; x++) {
            w[IDX(W, x, y, z0)] = 0;
            w[IDX(W, x, y, z1)] = 0;
        }

    for (int64_t z = z0; z < z1; z++)
        for (int64_t x = ux0; x <= ux1; x++) {
            u[IDX(U, x, y0 - 1, z)] = NEG(u[IDX(U, x, y0, z)]);
            u[IDX(U, x, y1, z)] = NEG(u[IDX(U, x, y1 - 1, z)]);
        }
#if ORC_MUTANT == 15
    for (int64_t y = y0; y < y1; y++)
#else
    for (int64_t y = y0 - 1; y <= y1; y++)
#endif
        for (int64_t x = ux0; x <= ux1; x++) {
            u[IDX(U, x, y, z0 - 1)] = NEG(u[IDX(U, x, y, z0)]);
            u[IDX(U, x, y, z1)] = NEG(u[IDX(U, x, y, z1 - 1)]);
        }
    for (int64_t z = z0; z < z1; z++)
        for (int64_t y = y0; y <= y1; y++) {
            v[IDX(V, g - 1, y, z)] = NEG(v[IDX(V, g, y, z)]);
            v[IDX(V, g + e->NX, y, z)] = NEG(v[IDX(V, g + e->NX - 1, y, z)]);
        }
    for (int64_t y = y0; y <= y1; y++)
        for (int64_t x = g - 1; x <= g + e->NX; x++) {
            v[IDX(V, x, y, z0 - 1)] = NEG(v[IDX(V, x, y, z0)]);
            v[IDX(V, x, y, z1)] = NEG(v[IDX(V, x, y, z1 - 1)]);
        }
    for (int64_t z = z0; z <= z1; z++)
        for (int64_t y = y0; y < y1; y++) {
            w[IDX(W, g - 1, y, z)] = NEG(w[IDX(W, g, y, z)]);
            w[IDX(W, g + e->NX, y, z)] = NEG(w[IDX(W, g + e->NX - 1, y, z)]);
        }
    for (int64_t z = z0; z <= z1; z++)
        for (int64_t x = g - 1; x <= g + e->NX; x++) {
            w[IDX(W, x, y0 - 1, z)] = NEG(w[IDX(W, x, y0, z)]);
            w[IDX(W, x, y1, z)] = NEG(w[IDX(W, x, y1 - 1, z)]);
        }
}

/* ------------------------------------------------------------ projection */

/* computeDivergence, project.go:42-62 */
static void compute_divergence(orc_engine *e, orc_shard s)
{
    const orc_field *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W], *D = &e->f[F_DIV];
    const double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);
    double *dv = DATA(e, F_DIV);
#if ORC_MUTANT == 4
    double scale = e->cfg.Density / (e->cfg.DT * 0.5 * e->H);
#else
    double scale = e->cfg.Density / (e->cfg.DT * e->H);
#endif
    int64_t g = e->G;
    for (int64_t z = s.z0; z < s.z1; z++)
        for (int64_t y = g; y < g + e->NY; y++)
            for (int64_t x = g; x < g + e->NX; x++) {
                double du = u[IDX(U, x + 1, y, z)] - u[IDX(U, x, y, z)];
                double dvv = v[IDX(V, x, y + 1, z)] - v[IDX(V, x, y, z)];
                double dw = w[IDX(W, x, y, z + 1)] - w[IDX(W, x, y, z)];
                dv[IDX(D, x, y, z)] = scale * (du + dvv + dw);
            }
}

/* gaussSeidelSweep, project.go:103-130 */
static void gauss_seidel_sweep(orc_engine *e, orc_shard s, int col)
{
    const orc_field *P = &e->f[F_P], *D = &e->f[F_DIV];
    double *p = DATA(e, F_P);
    const double *dv = DATA(e, F_DIV);
    double h2 = e->H * e->H;
    int64_t g = e->G, sy = P->sy, sz = P->sz;
    for (int64_t z = s.z0; z < s.z1; z++)
        for (int64_t y = g; y < g + e->NY; y++) {
            int64_t start = g;
            if (((start + y + z) & 1) != col) start++;
            int64_t base = IDX(P, 0, y, z);
            for (int64_t x = start; x < g + e->NX; x += 2) {
                int64_t i = base + x;
                double sum = p[i - 1] + p[i + 1] + p[i - sy] + p[i + sy] + p[i - sz] + p[i + sz];
#if ORC_MUTANT == 5
                p[i] = (sum + h2 * dv[IDX(D, x, y, z)]) / 6.0;
#else
                p[i] = (sum - h2 * dv[IDX(D, x, y, z)]) / 6.0;
#endif
            }
        }
}

/* applyPressureBC, project.go:139-163 (serial, driver thread). */
void orc_apply_pressure_bc(orc_engine *e)
{
    const orc_field *P = &e->f[F_P];
    double *p = DATA(e, F_P);
    int64_t g = e->G, x1 = g + e->NX, y1 = g + e->NY, z1 = g + e->NZ;
    for (int64_t z = g; z < z1; z++)
        for (int64_t y = g; y < y1; y++) {
#if ORC_MUTANT == 8
            p[IDX(P, g - 1, y, z)] = p[IDX(P, g + 1, y, z)];
#else
            p[IDX(P, g - 1, y, z)] = p[IDX(P, g, y, z)];
#endif
            p[IDX(P, x1, y, z)] = p[IDX(P, x1 - 1, y, z)];
        }
    for (int64_t z = g; z < z1; z++)
        for (int64_t x = g - 1; x <= x1; x++) {
            p[IDX(P, x, g - 1, z)] = p[IDX(P, x, g, z)];
            p[IDX(P, x, y1, z)] = p[IDX(P, x, y1 - 1, z)];
        }
    for (int64_t y = g - 1; y <= y1; y++)
        for (int64_t x = g - 1; x <= x1; x++) {
            p[IDX(P, x, y, g - 1)] = p[IDX(P, x, y, g)];
            p[IDX(P, x, y, z1)] = p[IDX(P, x, y, z1 - 1)];
        }
}

/* subtractGradient, project.go:172-205 */
static void subtract_gradient(orc_engine *e, orc_shard s)
{
    const orc_field *P = &e->f[F_P], *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W];
    const double *p = DATA(e, F_P);
    double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);
#if ORC_MUTANT == 7
    double scale = e->cfg.DT / e->cfg.Density;
#else
    double scale = e->cfg.DT / (e->cfg.Density * e->H);
#endif
    int64_t g = e->G;
    for (int64_t z = s.z0; z < s.z1; z++) {
        for (int64_t y = g; y < g + e->NY; y++)
            for (int64_t x = g; x <= g + e->NX; x++)
                u[IDX(U, x, y, z)] -= scale * (p[IDX(P, x, y, z)] - p[IDX(P, x - 1, y, z)]);
        for (int64_t y = g; y <= g + e->NY; y++)
            for (int64_t x = g; x < g + e->NX; x++)
                v[IDX(V, x, y, z)] -= scale * (p[IDX(P, x, y, z)] - p[IDX(P, x, y - 1, z)]);
    }
    int64_t zhi = s.z1;
    if (s.z1 == g + e->NZ) zhi++;
    for (int64_t z = s.z0; z < zhi; z++)
        for (int64_t y = g; y < g + e->NY; y++)
            for (int64_t x = g; x < g + e->NX; x++)
                w[IDX(W, x, y, z)] -= scale * (p[IDX(P, x, y, z)] - p[IDX(P, x, y, z - 1)]);
}

/* ------------------------------------------------------------ worker pool */

static int run_shard(orc_engine *e, orc_shard s)
{
    switch (e->op) {
    case OP_SWEEP: return advect_diffuse_axis(e, s, e->cur_axis, e->cur_dt);
    case OP_DIV: compute_divergence(e, s); return 0;
    case OP_GRAD: subtract_gradient(e, s); return 0;
    case OP_RED: gauss_seidel_sweep(e, s, 0); return 0;
    case OP_BLACK: gauss_seidel_sweep(e, s, 1); return 0;
    }
    return 0;
}

static void *worker_main(void *p)
{
    orc_worker_arg *a = (orc_worker_arg *)p;
    orc_engine *e = a->e;
    for (;;) {
        pthread_barrier_wait(&e->bar_start);
        if (e->op == OP_STOP) break;
        e->shard_fault[a->id] = run_shard(e, e->shards[a->id]);
        pthread_barrier_wait(&e->bar_end);
    }
    return NULL;
}

/* dispatch, aeroflux.go:313-324: every shard runs op, the join is the
 * barrier, and every worker's error slot is drained. */
static int dispatch(orc_engine *e, int op)
{
    e->op = op;
    if (e->nworkers == 1) {
        e->shard_fault[0] = run_shard(e, e->shards[0]);
    } else {
        pthread_barrier_wait(&e->bar_start);
        pthread_barrier_wait(&e->bar_end);
    }
    int first = 0;
    for (int i = 0; i < e->nworkers; i++)
        if (e->shard_fault[i]) first = 1;
    if (first) e->step_fault = 1;
    return first;
}

/* partition, aeroflux.go:218-232 */
static void partition(orc_engine *e)
{
    int n = e->nworkers;
    int64_t base = e->NZ / n, rem = e->NZ % n, z = e->G;
    for (int i = 0; i < n; i++) {
        int64_t cnt = base + (i < rem ? 1 : 0);
        e->shards[i].z0 = z;
        e->shards[i].z1 = z + cnt;
        e->shards[i].id = i;
        z += cnt;
    }
}

/* ------------------------------------------------------------ public API */

/* New, aeroflux.go:176-213.  The config must already be normalised. */
orc_engine *orc_new(const orc_config *cfg)
{
    orc_engine *e = (orc_engine *)calloc(1, sizeof(orc_engine));
    if (!e) return NULL;
    e->cfg = *cfg;
    e->NX = cfg->NX; e->NY = cfg->NY; e->NZ = cfg->NZ; e->G = cfg->Ghost; e->H = cfg->H;
    e->bytes = mac_layout(e->f, e->NX, e->NY, e->NZ, e->G);
    size_t alloc = (size_t)((e->bytes + 4095) / 4096 * 4096);
    if (posix_memalign((void **)&e->arena, 4096, alloc) != 0) {
        free(e);
        return NULL;
    }
    memset(e->arena, 0, alloc); /* MAC.Zero, grid.go:158-163 */

    e->nworkers = (int)cfg->Workers;
    e->shards = (orc_shard *)calloc((size_t)e->nworkers, sizeof(orc_shard));
    e->shard_fault = (int *)calloc((size_t)e->nworkers, sizeof(int));
    partition(e);
    if (e->nworkers > 1) {
        e->threads = (pthread_t *)calloc((size_t)e->nworkers, sizeof(pthread_t));
        e->args = (orc_worker_arg *)calloc((size_t)e->nworkers, sizeof(orc_worker_arg));
        pthread_barrier_init(&e->bar_start, NULL, (unsigned)e->nworkers + 1);
        pthread_barrier_init(&e->bar_end, NULL, (unsigned)e->nworkers + 1);
        for (int i = 0; i < e->nworkers; i++) {
            e->args[i].e = e;
            e->args[i].id = i;
            pthread_create(&e->threads[i], NULL, worker_main, &e->args[i]);
        }
        e->started = 1;
    }
    return e;
}

/* Close, aeroflux.go:419-439 */
void orc_free(orc_engine *e)
{
    if (!e) return;
    if (e->started) {
        e->op = OP_STOP;
        pthread_barrier_wait(&e->bar_start);
        for (int i = 0; i < e->nworkers; i++) pthread_join(e->threads[i], NULL);
        pthread_barrier_destroy(&e->bar_start);
        pthread_barrier_destroy(&e->bar_end);
    }
    free(e->threads);
    free(e->args);
    free(e->shards);
    free(e->shard_fault);
    free(e->arena);
    free(e);
}

/* Field descriptor (current, after any descriptor swaps) and raw pointer. */
void orc_field_desc(orc_engine *e, int id, int64_t out[7])
{
    const orc_field *f = &e->f[id];
    out[0] = f->off; out[1] = f->nx; out[2] = f->ny; out[3] = f->nz;
    out[4] = f->sy; out[5] = f->sz; out[6] = f->count;
}

double *orc_field_ptr(orc_engine *e, int id) { return DATA(e, id); }
int64_t orc_arena_bytes(orc_engine *e) { return e->bytes; }
unsigned char *orc_arena_ptr(orc_engine *e) { return e->arena; }

/* One dispatch of advectDiffuseAxis (no BC, no commit). */
int orc_sweep_axis(orc_engine *e, int ax, double dt)
{
    e->cur_axis = ax;
    e->cur_dt = dt;
    return dispatch(e, OP_SWEEP);
}

void orc_compute_divergence(orc_engine *e) { dispatch(e, OP_DIV); }
void orc_gs_sweep(orc_engine *e, int col) { dispatch(e, col ? OP_BLACK : OP_RED); }
void orc_subtract_gradient(orc_engine *e) { dispatch(e, OP_GRAD); }

/* solvePressureRedBlack, project.go:73-100 */
void orc_solve_pressure(orc_engine *e)
{
    double *p = DATA(e, F_P);
    for (int64_t i = 0; i < e->f[F_P].count; i++) p[i] = 0;
    for (int64_t it = 0; it < e->cfg.GaussSeidelIters; it++) {
        orc_apply_pressure_bc(e);
        dispatch(e, OP_RED);
        orc_apply_pressure_bc(e);
        dispatch(e, OP_BLACK);
    }
    orc_apply_pressure_bc(e);
}

/* project, project.go:12-34 */
int orc_project(orc_engine *e)
{
    orc_apply_velocity_bc(e);
    dispatch(e, OP_DIV);
    orc_solve_pressure(e);
    dispatch(e, OP_GRAD);
    orc_apply_velocity_bc(e);
    return 0;
}

/* Step, aeroflux.go:348-409.  Returns 0, or 1 for a contained fault (the
 * state is then partially updated, as in the reference). */
int orc_step(orc_engine *e)
{
    static const int fwd[3] = {AX_X, AX_Y, AX_Z};
    static const int rev[3] = {AX_Z, AX_Y, AX_X};
    e->step_fault = 0;
    e->cur_dt = e->cfg.DT * 0.5;
    for (int k = 0; k < 3; k++) {
        orc_apply_velocity_bc(e);
        e->cur_axis = fwd[k];
        if (dispatch(e, OP_SWEEP)) return 1;
        orc_commit_velocity(e);
    }
    orc_project(e);
    for (int k = 0; k < 3; k++) {
        orc_apply_velocity_bc(e);
#if ORC_MUTANT == 11
        e->cur_axis = fwd[k];
#else
        e->cur_axis = rev[k];
#endif
        if (dispatch(e, OP_SWEEP)) return 1;
        orc_commit_velocity(e);
    }
    orc_project(e);
    return e->step_fault;
}

/* MaxDivergence, project.go:210-232 */
double orc_max_divergence(orc_engine *e)
{
    const orc_field *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W];
    const double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);
    int64_t g = e->G;
    double invH = 1.0 / e->H, mx = 0;
    for (int64_t z = g; z < g + e->NZ; z++)
        for (int64_t y = g; y < g + e->NY; y++)
            for (int64_t x = g; x < g + e->NX; x++) {
                double d = invH * ((u[IDX(U, x + 1, y, z)] - u[IDX(U, x, y, z)]) +
                                   (v[IDX(V, x, y + 1, z)] - v[IDX(V, x, y, z)]) +
                                   (w[IDX(W, x, y, z + 1)] - w[IDX(W, x, y, z)]));
                double a = fabs(d);
#if ORC_MUTANT == 12
                if (a > mx || isnan(a)) mx = a;
#else
                if (a > mx) mx = a;
#endif
            }
    return mx;
}

/* KineticEnergy, project.go:238-258 */
double orc_kinetic_energy(orc_engine *e)
{
    const orc_field *U = &e->f[F_U], *V = &e->f[F_V], *W = &e->f[F_W];
    const double *u = DATA(e, F_U), *v = DATA(e, F_V), *w = DATA(e, F_W);
    int64_t g = e->G;
    double vol = e->H * e->H * e->H, sum = 0;
    for (int64_t z = g; z < g + e->NZ; z++)
        for (int64_t y = g; y < g + e->NY; y++)
            for (int64_t x = g; x < g + e->NX; x++) {
#if ORC_MUTANT == 9
                double uc = 0.5 * (u[IDX(U, x - 1, y, z)] + u[IDX(U, x, y, z)]);
#else
                double uc = 0.5 * (u[IDX(U, x, y, z)] + u[IDX(U, x + 1, y, z)]);
#endif
                double vc = 0.5 * (v[IDX(V, x, y, z)] + v[IDX(V, x, y + 1, z)]);
                double wc = 0.5 * (w[IDX(W, x, y, z)] + w[IDX(W, x, y, z + 1)]);
                sum += uc * uc + vc * vc + wc * wc;
            }
    return 0.5 * sum * vol;
}

/* MaxVelocity, stencil.go:319-331 (whole padded U, V, W arrays). */
double orc_max_velocity(orc_engine *e)
{
    double mx = 0;
    static const int ids[3] = {F_U, F_V, F_W};
    for (int k = 0; k < 3; k++) {
        const double *d = DATA(e, ids[k]);
#if ORC_MUTANT == 10
        for (int64_t i = e->f[ids[k]].sz; i < e->f[ids[k]].count - e->f[ids[k]].sz; i++) {
#else
        for (int64_t i = 0; i < e->f[ids[k]].count; i++) {
#endif
            double a = fabs(d[i]);
            if (a > mx) mx = a;
        }
    }
    return mx;
}

/* CFL, stencil.go:336-338 */
double orc_cfl(orc_engine *e)
{
#if ORC_MUTANT == 14
    return orc_max_velocity(e) * (e->cfg.DT * 0.5) / e->cfg.H;
#else
    return orc_max_velocity(e) * e->cfg.DT / e->cfg.H;
#endif
}
