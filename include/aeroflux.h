/*
 * aeroflux.h -- C ABI of the MI355X-native AeroFlux step data plane.
 *
 * This is the drop-in boundary beneath the reference's Go API: a cgo shim
 * keeps `aeroflux.New(Config) / Step() / MaxDivergence() / Close()` unchanged
 * and calls these functions.  Every entry point names the reference function
 * it replaces (file:line in the AeroFlex Go package).  Signatures use only
 * plain C types: fixed-width integers, doubles, and host pointers; device
 * addresses cross as uint64_t, mirroring the reference's uintptr discipline
 * (accelerate_darwin.go:21-24).
 *
 * Threading: one engine is driven from one host thread at a time (as
 * Step/project/diagnostics are called from one goroutine in the reference).
 * Each engine owns one HIP stream at default priority; no host threads are
 * created.  Errors are integer codes; af_last_error() returns the message of
 * the most recent failure on the calling thread, prefixed "aeroflux: ".
 */
#ifndef AEROFLUX_H
#define AEROFLUX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes */
enum {
    AF_OK = 0,
    AF_ERR_CONFIG = 1,      /* Config.Normalize rejection (aeroflux.go:47-113) */
    AF_ERR_FAULT = 2,       /* contained device fault: non-finite departure
                               coordinate (the reference's guard-page /
                               bounds-panic path, arena.go:230-243) */
    AF_ERR_OOM = 3,         /* device arena allocation failed */
    AF_ERR_HIP = 4,         /* HIP runtime error */
    AF_ERR_ARG = 5,         /* bad argument (field id, size, null) */
    AF_ERR_CLOSED = 6,      /* engine already closed */
    AF_ERR_UNSUPPORTED = 7  /* feature absent on this platform */
};

/* Field ids in arena order, NewMAC (grid.go:119-127). */
enum {
    AF_FIELD_P = 0,
    AF_FIELD_U = 1,
    AF_FIELD_V = 2,
    AF_FIELD_W = 3,
    AF_FIELD_DIV = 4,
    AF_FIELD_TMP0 = 5,
    AF_FIELD_TMP1 = 6,
    AF_FIELD_TMP2 = 7,
    AF_FIELD_COUNT = 8
};

enum { AF_AXIS_X = 0, AF_AXIS_Y = 1, AF_AXIS_Z = 2 };      /* stencil.go:7-11 */
enum { AF_COLOR_RED = 0, AF_COLOR_BLACK = 1 };              /* layout.go:30-33 */

/* Pressure-solve precision (no counterpart in the reference, whose solve is
 * FP64 only).  AF_PRECISION_FP64 is the reference algorithm, bit for bit. */
enum {
    AF_PRECISION_FP64 = 0,
    AF_PRECISION_MIXED = 1  /* FP32 red-black smoothing inside an FP64
                               defect-correction loop */
};

/* Config, aeroflux.go:14-41.  The first block mirrors the Go struct field
 * for field (Go int -> int64_t, bool -> int32_t).  The trailing block holds
 * MI355X-build extensions; their zero values select reference behaviour. */
typedef struct af_config {
    int64_t NX, NY, NZ;
    int64_t Ghost;
    double H;
    double Viscosity;
    double Density;
    double DT;
    int64_t Workers;          /* CPU worker count in the reference; kept and
                                 normalised identically, unused on the GPU */
    int64_t GaussSeidelIters;
    int32_t UseSparseSolver;  /* accepted and ignored, as by the reference's
                                 non-Darwin build: the red-black solve runs
                                 (project.go:19, sparse_fallback.go:20-24);
                                 the GPU direct solve is switched on by
                                 af_enable_direct_solver */
    int32_t Lock;             /* accepted; device memory is always resident */

    /* --- extensions --- */
    int32_t Precision;        /* AF_PRECISION_* */
    int32_t MixedMaxCycles;   /* mixed: max FP64 residual checks per
                                 checked solve (0 -> 4): the first, one
                                 after the continued FP32 iteration, then
                                 one per defect-correction cycle */
    int32_t MixedSweeps;      /* mixed: FP32 sweep pairs of the first cycle
                                 (0 -> GaussSeidelIters) */
    int32_t MixedFixSweeps;   /* mixed: FP32 sweep pairs of the continued
                                 iteration and of each defect-correction
                                 cycle (0 -> 2) */
    double MixedTarget;       /* mixed: max|div| the step leaves.  0 -> each
                                 projection runs exactly one cycle.  > 0 ->
                                 the step's second projection (and a
                                 stand-alone af_project) checks the FP64
                                 residual after the first cycle and, while
                                 the max|div| it implies exceeds this,
                                 continues the FP32 iteration by
                                 MixedFixSweeps pairs once, then runs FP64
                                 defect-correction cycles; the step's first
                                 projection runs one cycle */
    int32_t TileK;            /* FP64 red-black schedule: 0 -> one pass per
                                 launch on the colour-split or the arena
                                 layout, whichever is faster on this grid
                                 (timed at New; split at 128^3 and up);
                                 -2 -> colour-split layout; -1 -> arena
                                 layout; 2, 4, 6 -> that many passes per
                                 launch, temporally blocked, arena layout.
                                 All give the same bits. */
    int32_t TileRows;         /* blocked: rows per thread, 1 or 2 (0 -> tuned) */
    int32_t TileZ;            /* blocked: z-planes per workgroup (0 -> tuned) */
    int32_t SweepShape;       /* advection workgroup shape: 0 -> tuned, k ->
                                 shape k-1 (see af_sweep_shape) */
    int32_t SweepSlab;        /* advection block order: 0 -> tuned, -1 ->
                                 plane by plane, k -> y-slabs of k rows */
    int32_t PassSlab;         /* colour-split pass order (FP64 split and
                                 mixed solves): 0 -> tuned, -1 -> plane by
                                 plane, k -> y-slabs of k workgroup rows
                                 (4 grid rows each) */
} af_config;

/* Field, grid.go:12-21 (descriptor after any commitVelocity swaps). */
typedef struct af_field {
    int64_t Off;   /* byte offset of element (0,0,0) in the arena */
    int64_t NX, NY, NZ;
    int64_t SY, SZ;
    int64_t Count;
} af_field;

typedef struct af_engine af_engine;

/* ------------------------------------------------------------ lifecycle */

/* Config.Normalize, aeroflux.go:47-113: fills defaults, rejects bad
 * geometry.  Workers defaults to the host's online CPU count (GOMAXPROCS). */
int af_config_normalize(af_config *cfg);

/* MACBytes, grid.go:138-144: arena footprint, 0 for an invalid geometry. */
int64_t af_mac_bytes(int64_t nx, int64_t ny, int64_t nz, int64_t ghost);

/* New, aeroflux.go:176-213: normalises a copy of *cfg, allocates the zeroed
 * HBM arena on the current HIP device, and prepares the per-step graph. */
int af_new(const af_config *cfg, af_engine **out);

/* Close, aeroflux.go:419-439: releases device memory; idempotent.  The
 * handle stays valid (closed) until af_free. */
int af_close(af_engine *e);

/* Frees a handle (closing it first if needed).  No reference counterpart:
 * the Go engine struct is garbage-collected. */
void af_free(af_engine *e);

/* Copy of the engine's normalised config (Engine.Cfg, aeroflux.go:134). */
int af_get_config(const af_engine *e, af_config *out);

/* ------------------------------------------------------------- hot path */

/* Step, aeroflux.go:348-409: one Strang-palindrome step X,Y,Z,[project],
 * Z,Y,X,[project], replayed as one hipGraph.  Blocks until done; returns
 * AF_ERR_FAULT if any departure coordinate was NaN (state then unspecified,
 * as after a contained fault in the reference). */
int af_step(af_engine *e);

/* n Steps back to back.  With first_fault_step null: one host
 * synchronisation at the end, AF_ERR_FAULT if any step faulted.  Non-null:
 * the fault flag is read after every step (one synchronisation per step, as
 * n Step calls), the call stops at the first faulting step and
 * *first_fault_step receives its index.  No reference counterpart (a loop
 * of Step calls). */
int af_step_n(af_engine *e, int64_t n, int64_t *first_fault_step);

/* ---------------------------------------------------------- diagnostics */

int af_max_divergence(af_engine *e, double *out);   /* project.go:210-232 */
int af_kinetic_energy(af_engine *e, double *out);   /* project.go:238-258 */
int af_max_velocity(af_engine *e, double *out);     /* stencil.go:319-331 */
int af_cfl(af_engine *e, double *out);              /* stencil.go:336-338 */

/* af_max_divergence without any allocation, for callers that cannot
 * allocate device or pinned memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the
 * reference.  u, v, w are the device addresses of element (0,0,0) of
 * fields laid out as NewMAC lays out U, V and W for this grid (for an
 * engine: af_field_device_ptr); the result, the same bits as
 * af_max_divergence on those fields, is written to the device double at
 * `out` when `stream` (a hipStream_t as an integer, 0 = the null stream, of
 * the current device) reaches it: the call only enqueues two kernels and
 * does not synchronise.  Reduction partials live in a fixed 64 KiB device
 * array per GPU, held until the process ends, so grids need
 * ceil(NX/64) * ceil(NY/4) <= 8192 (AF_ERR_ARG otherwise).  Calls that
 * overlap on one GPU share that array and race. */
int af_max_divergence_noalloc(uint64_t u, uint64_t v, uint64_t w, int64_t nx, int64_t ny,
                              int64_t nz, int64_t ghost, double h, uint64_t out,
                              uint64_t stream);

/* EnableSparseSolver, sparse_fallback.go:17-19: always fails here, with the
 * reference's non-Darwin behaviour. */
int af_enable_sparse_solver(af_engine *e);

/* Direct pressure solve on the GPU, for the system the reference's Darwin
 * build factors in EnableSparseSolver (sparse_darwin.go:19-125) and solves
 * in solvePressureDirect (:127-172): the interior 7-point Neumann Laplacian
 * with the first cell's diagonal raised by 1/H^2, right-hand side -Div,
 * followed by applyPressureBC.  Solved spectrally (orthonormal DCT-II along
 * each axis), exact up to rounding.  Once enabled every projection uses it
 * in place of the red-black sweeps (project.go:19-27).  The transforms run
 * as FFT passes when every extent is a power of two in 8..512, otherwise as
 * FP64 GEMMs against DCT matrices.  No counterpart in the reference's Linux
 * build. */
int af_enable_direct_solver(af_engine *e);
/* The same with the transform chosen: 0 = as above, 1 = FFT passes
 * (AF_ERR_UNSUPPORTED for other extents), 2 = dense GEMMs.  Switching
 * method releases the other one's buffers. */
int af_enable_direct_solver_method(af_engine *e, int32_t method);
/* *on = 0 when the red-black solve runs, else the method in use (1 FFT,
 * 2 dense). */
int af_direct_solver_enabled(const af_engine *e, int32_t *on);

/* --------------------------------------------------------- field access */

/* Field descriptor (MAC.P/U/..., grid.go:60-67) after descriptor swaps. */
int af_field_desc(const af_engine *e, int field_id, af_field *out);

/* MAC.Data (grid.go:147-149) cannot alias HBM; these copy a whole padded
 * field between host memory and the arena.  count must equal Field.Count. */
int af_field_read(af_engine *e, int field_id, double *host, int64_t count);
int af_field_write(af_engine *e, int field_id, const double *host, int64_t count);

/* Device address of a field's element (0,0,0), for device-side consumers. */
int af_field_device_ptr(const af_engine *e, int field_id, uint64_t *out);

/* Total arena bytes (MAC.Bytes, grid.go:128). */
int64_t af_arena_bytes(const af_engine *e);

/* ------------------------------------------------- snapshot and restore */
/* No counterpart in the reference (it has no engine snapshot format).  A
 * snapshot is this 128-byte header followed by the eight fields in field-id
 * order (P, U, V, W, Div, Tmp0, Tmp1, Tmp2), each Field.Count doubles of
 * the padded field, with nothing between them; little-endian.  It holds the
 * whole state a step reads, so restoring it into an engine (or a batch slot)
 * with the same grid and stepping on gives the same bits as never having
 * stopped.  Checksum = sum mod 2^64 over the payload's 64-bit words w_i
 * (i from 0 across all fields) of m(w_i ^ (i * 0x9E3779B97F4A7C15)), m the
 * SplitMix64 output function (z ^= z >> 30; z *= 0xBF58476D1CE4E5B9;
 * z ^= z >> 27; z *= 0x94D049BB133111EB; z ^= z >> 31). */
typedef struct af_snapshot_header {
    char Magic[8];             /* 'A','F','S','N','A','P',0,1 */
    uint32_t Version;          /* 1 */
    uint32_t HeaderBytes;      /* 128 */
    int64_t NX, NY, NZ, Ghost; /* must match the engine restored into */
    double H, DT, Viscosity, Density; /* recorded only */
    int64_t PayloadBytes;      /* 8 * sum of the eight Field.Count */
    uint64_t Checksum;
    uint8_t Reserved[32];      /* zero */
} af_snapshot_header;

/* Bytes a snapshot of this engine takes (header + payload). */
int af_snapshot_bytes(const af_engine *e, int64_t *bytes);
/* Writes a snapshot into dst (host or device memory of at least
 * af_snapshot_bytes bytes; pinned host memory copies fastest). */
int af_snapshot(af_engine *e, void *dst, int64_t bytes);
/* Restores a snapshot.  The header is checked (magic, version, grid, sizes)
 * before the arena is touched: AF_ERR_ARG leaves the engine as it was.  A
 * payload whose checksum does not match is rejected with AF_ERR_ARG after
 * zeroing the eight fields.  Config values other than the grid (DT,
 * Viscosity, Precision ...) are the engine's own, not the snapshot's. */
int af_restore(af_engine *e, const void *src, int64_t bytes);

/* ----------------------------------------- stage entry points (parity) */
/* These mirror the engine's unexported methods that the reference's own
 * in-package tests call (engine_test.go:121, :334). */

int af_apply_velocity_bc(af_engine *e);                 /* stencil.go:246-315 */
int af_sweep_axis(af_engine *e, int axis, double dt);   /* stencil.go:21-25, one dispatch */
int af_commit_velocity(af_engine *e);                   /* stencil.go:230-235 */
int af_compute_divergence(af_engine *e);                /* project.go:42-62 */
int af_gs_sweep(af_engine *e, int color);               /* project.go:103-130 */
int af_apply_pressure_bc(af_engine *e);                 /* project.go:139-163 */
int af_solve_pressure(af_engine *e);                    /* project.go:73-100 */
int af_subtract_gradient(af_engine *e);                 /* project.go:172-205 */
int af_project(af_engine *e);                           /* project.go:12-34 */

/* -------------------------------------------------------- measurement */

/* Average device time (ms, HIP events on the engine's stream) of `reps`
 * back-to-back launches of one kernel of the step, launched exactly as the
 * step launches it.  kernel: 0 = one launch of the red-black kernel the
 * solve uses (one colour pass on the colour-split layout by default -- an
 * FP32 pass for mixed precision -- one pass on the arena layout for
 * TileK = -1, TileK passes for TileK > 0),
 * 1 = advection sweep (X axis), 2 = divergence, 3 = gradient subtract,
 * 4 = advection sweep (Y axis), 5 = advection sweep (Z axis),
 * 6 = one whole pressure solve from Div, as the projection runs it (the
 * direct solve once af_enable_direct_solver was called).
 * No reference
 * counterpart (measurement only); it overwrites P and the Tmp fields. */
int af_time_kernel(af_engine *e, int kernel, int reps, double *avg_ms);

/* Attainable streaming bandwidth (GB/s, bytes moved / time) of one access
 * pattern over arrays of n doubles each (n even), timed over `reps`
 * back-to-back launches after one warm-up: kind 0 = read one array, 1 =
 * copy, 2 = two reads and one write (a colour pass's pattern), 3 = four
 * reads and three writes (the gradient's).  The arrays hold pseudo-random
 * values in [-1, 1] (all-zero data moves measurably faster).  Allocates and
 * frees its arrays.  No reference counterpart (measurement only). */
int af_measure_stream(int kind, int64_t n, int reps, double *gbps);

/* Mixed precision only: max|div| implied by the last FP64 residual check
 * (0 if MixedTarget == 0) and the cycles the last solve ran (the residual
 * checks of a checked solve; 1 for a one-cycle solve).  No reference
 * counterpart. */
int af_mixed_stats(af_engine *e, double *last_div, int32_t *cycles);

/* Tile shape the engine's red-black solve runs with (after tuning):
 * passes per launch (-1: one pass per launch), rows per thread, z-chunk.
 * No reference counterpart. */
int af_solver_tile(const af_engine *e, int32_t *k, int32_t *rows, int32_t *zchunk);

/* Advection sweep workgroup shape in use: index, rows per workgroup and z
 * positions per thread (64 lanes along x; LDS-staged shapes 8-13: the
 * largest z-chunk a workgroup walks; 8-11 stage one box of the three
 * fields, 12-13 each field's own box).  Chosen by timing every shape -- and
 * for the staged shapes two z-chunk counts -- on the engine's grid at New
 * unless SweepShape fixes it.  No reference counterpart. */
int af_sweep_shape(const af_engine *e, int32_t *shape, int32_t *rows, int32_t *zper);

/* Rows per y-slab of the advection block order in use (-1: plane by plane).
 * No reference counterpart. */
int af_sweep_slab(const af_engine *e, int32_t *rows);

/* Workgroup rows per y-slab of the colour-split passes' order in use (-1:
 * plane by plane).  No reference counterpart. */
int af_pass_slab(const af_engine *e, int32_t *rows);

/* ------------------------------------------------------------- batching */
/* Many engines of uneven sizes stepped together.  No reference counterpart:
 * the reference steps one Engine per Step call (aeroflux.go:348-409); a
 * batch steps each of its engines exactly as that Step does.  A batch owns
 * one HBM pool of a fixed budget; each engine added to it is a slot holding
 * New's zeroed arena (grid.go:79-134, same field order and offsets) plus its
 * solve buffers.  One batched step steps every resident slot once: each
 * stage of the Strang palindrome is one persistent launch over a work-list
 * of the tiles of all slots, and the step is replayed as one hipGraph.
 * Membership changes between steps update the work-list, not the graph.
 * Every slot's results are bit-identical to the same engine stepped alone
 * with af_step.  Calls on one batch must come from one host thread at a
 * time; each batch owns one HIP stream at default priority. */
typedef struct af_batch af_batch;

typedef struct af_batch_config {
    int32_t Precision;         /* AF_PRECISION_* of every slot */
    int32_t MaxSlots;          /* slot capacity (0 -> 1024) */
    int64_t MemoryBudget;      /* bytes of HBM for slots (0 -> 1 GiB) */
    /* Solve schedule shared by every slot; a slot's normalised config must
     * carry the same values (0 -> the Config.Normalize defaults). */
    int64_t GaussSeidelIters;
    int32_t MixedMaxCycles;
    int32_t MixedSweeps;
    int32_t MixedFixSweeps;
    int32_t MixedChecked;      /* mixed: 0 -> one cycle (slots need
                                  MixedTarget == 0); 1 -> cycles until each
                                  slot's MixedTarget (> 0) is met */
    int32_t SweepShape;        /* global-gather advection shape: 0 -> tuned;
                                  k (1..8) -> shape k-1 (af_sweep_shape) */
    int32_t Workgroups;        /* persistent workgroups per launch (0 -> tuned;
                                  else a multiple of 8) */
    int32_t SweepStaged;       /* LDS-staged advection for slots whose x
                                  extent fills 64-wide tiles to >= 3/4:
                                  0 -> tuned; -1 -> off (every slot
                                  gathers); k (1..8) -> variant k-1, i.e.
                                  (rows, z-planes) of (4,16), (4,32), (8,16),
                                  (8,32) per workgroup staging one box of
                                  all three fields, then (8,16), (8,32),
                                  (4,16), (4,32) staging each field's own
                                  box */
    int32_t ServePolicy;       /* af_serve: 0 -> admission first come first
                                  served, every resident job advances each
                                  batched step; 1 -> least remaining work
                                  (cells x steps left) admitted first,
                                  preempting: a resident job is snapshotted
                                  to host memory and requeued when a queued
                                  job that cannot be placed has less than
                                  1/PreemptRatio of its remaining work; each
                                  job is preempted at most once; 2 ->
                                  admission as 0, but each batched step
                                  advances only the resident jobs with the
                                  least remaining work, up to StepCellBudget
                                  interior cells (at least one job), so short
                                  jobs finish first while the GPU stays
                                  full; 3 -> both: least remaining work
                                  admitted first (no preemption) and stepped
                                  first */
    double PreemptRatio;       /* policy 1 (0 -> 4) */
    int64_t SnapshotBudget;    /* policy 1: host bytes that preempted jobs'
                                  snapshots may hold at once (0 -> 32 GiB) */
    int64_t StepCellBudget;    /* policies 2, 3: interior cells one batched step
                                  advances (0 -> 32 Mi cells) */
    int64_t SolveGroupBytes;   /* pressure solve blocking: slots are split
                                  into groups of about equal solve-buffer bytes
                                  (split solution and right-hand side), at most
                                  this many per group where slot sizes allow,
                                  and the solve runs group after group so
                                  that a group's buffers stay in the
                                  last-level cache across its passes
                                  (0 -> 224 MiB; -1 -> one group).  Results
                                  do not depend on it */
    int32_t StepMaxWait;       /* policies 2, 3: a resident job passed over by
                                  this many consecutive batched steps goes
                                  first in the next one, longest wait first
                                  (0 -> no bound: a large job may wait while
                                  smaller ones keep it out of the budget) */
} af_batch_config;

typedef struct af_batch_info {
    int32_t Slots;             /* resident slots */
    int32_t SweepShape;        /* advection workgroup shape in use (index) */
    int32_t SweepWorkgroups;   /* persistent workgroups: advection sweeps */
    int32_t PassWorkgroups;    /* persistent workgroups: red-black passes */
    int32_t OtherWorkgroups;   /* persistent workgroups: every other stage */
    int32_t Tuned;             /* 1 once the autotuner has run */
    int64_t Cells;             /* interior cells over resident slots */
    int64_t BytesInUse;        /* pool bytes held by slots */
    int64_t BytesLargestFree;  /* largest free pool block */
    int64_t MemoryBudget;
    int32_t SweepStaged;       /* LDS-staged variant in use (-1: none) */
    int32_t StagedSlots;       /* resident slots that use it */
    int32_t SolveGroups;       /* solve groups of the last work-list built */
} af_batch_info;

int af_batch_new(const af_batch_config *cfg, af_batch **out);
/* Releases the pool and every slot; idempotent.  Handle stays valid. */
int af_batch_close(af_batch *b);
void af_batch_free(af_batch *b);
int af_batch_get_info(const af_batch *b, af_batch_info *out);

/* Pool bytes one slot of this config takes (arena + solve buffers), 0 if
 * the config does not normalise.  For admission decisions. */
int64_t af_batch_slot_bytes(const af_config *cfg, int32_t precision);

/* Adds an engine: normalises a copy of *cfg (as New), reserves its block in
 * the pool and zeroes it (MAC.Zero, grid.go:158-163).  AF_ERR_OOM if no pool
 * block fits; AF_ERR_CONFIG if the config does not match the batch's
 * precision and solve schedule.  *slot receives the slot id. */
int af_batch_add(af_batch *b, const af_config *cfg, int32_t *slot);
/* Returns a slot's block to the pool (Close for one engine). */
int af_batch_remove(af_batch *b, int32_t slot);

/* n batched steps, one host synchronisation at the end.  Slots whose fault
 * flag was raised (a NaN departure coordinate, as af_step) are written to
 * faulted[0 .. min(*nfaulted, cap)); returns AF_ERR_FAULT if any. */
int af_batch_step_n(af_batch *b, int64_t n, int32_t *faulted, int32_t cap, int32_t *nfaulted);

/* Per-slot counterparts of the engine calls. */
int af_batch_field_desc(const af_batch *b, int32_t slot, int field_id, af_field *out);
int af_batch_field_read(af_batch *b, int32_t slot, int field_id, double *host, int64_t count);
int af_batch_field_write(af_batch *b, int32_t slot, int field_id, const double *host,
                         int64_t count);
int af_batch_apply_velocity_bc(af_batch *b, int32_t slot);          /* stencil.go:246-315 */
int af_batch_max_divergence(af_batch *b, int32_t slot, double *out); /* project.go:210-232 */
int af_batch_kinetic_energy(af_batch *b, int32_t slot, double *out); /* project.go:238-258 */
int af_batch_mixed_stats(af_batch *b, int32_t slot, double *last_div, int32_t *cycles);
/* Snapshot and restore of one slot, in the engine's format (above): a
 * slot's snapshot restores into an engine of the same grid and the other
 * way round. */
int af_batch_snapshot_bytes(const af_batch *b, int32_t slot, int64_t *bytes);
int af_batch_snapshot(af_batch *b, int32_t slot, void *dst, int64_t bytes);
int af_batch_restore(af_batch *b, int32_t slot, const void *src, int64_t bytes);

/* Runs the autotuner now on the resident slots (it otherwise runs at the
 * first step): advection shape and persistent workgroup counts, each timed
 * on batched launches.  Results of later steps are unaffected; until the
 * next step the Tmp fields hold the timing sweeps' output. */
int af_batch_tune(af_batch *b);

/* Average device time (ms) of `reps` launches of one batched stage over all
 * resident slots: 0 = red-black colour pass, 1, 5, 6 = advection sweep along
 * X, Y, Z, 2 = divergence, 3 = gradient subtract, 4 = velocity BC, 7 =
 * pressure ghost fill, 8 = merge of the solve into P, 9 = FP64 residual of
 * the mixed-precision solve (mixed batches only).  Measurement only: 0, 1,
 * 4-7 and 9 leave the state a step reads unchanged (the sweeps write the
 * Tmp fields, which a step rewrites before reading; 9 writes solve buffers
 * the next divergence rewrites); 2 rewrites Div, 3 changes the velocity and
 * 8 rewrites P's interior from the last solve's buffers (which it already
 * equals unless the batch iterates checked cycles). */
int af_batch_time_stage(af_batch *b, int stage, int reps, double *avg_ms);

/* -------------------------------------------------------------- serving */
/* A request loop over a batch (no reference counterpart).  Step-N jobs of
 * different sizes arrive over time; each is admitted as a slot as soon as
 * the pool can place it (in the order of the batch's ServePolicy: first
 * come first served, or least remaining work first), gets a seeded
 * band-limited initial velocity on the device, and is stepped with the
 * other resident jobs by the batched step graph (continuous batching: the
 * resident set changes between steps, while earlier steps still run; under
 * policies 2 and 3 each step advances a least-work subset of them).  When
 * a job's N steps are done its final MaxDivergence and KineticEnergy are
 * recorded and its block returns to the pool.  Every job's result equals
 * the same engine stepped alone (af_new, af_init_band_limited, af_step_n). */
typedef struct af_serve_job {
    int64_t NX, NY, NZ;
    int64_t Steps;             /* N */
    double Viscosity;
    double DT;                 /* 0 -> Config.Normalize default */
    double H;                  /* 0 -> 1 / max(NX, NY, NZ) */
    uint64_t Seed;             /* initial condition */
    double Arrival;            /* seconds after the start of the run */
    double MixedTarget;        /* checked mixed batches (MixedChecked = 1):
                                  the job's max|div| target, > 0 (see
                                  af_config.MixedTarget); otherwise 0 */
} af_serve_job;

typedef struct af_serve_result {
    double Admitted;           /* seconds after start: pool block granted */
    double Done;               /* seconds after start: last step seen complete */
    double MaxDivergence;      /* after the last step (project.go:210-232) */
    double KineticEnergy;      /* after the last step (project.go:238-258) */
    int32_t Status;            /* AF_OK, AF_ERR_FAULT, or AF_ERR_OOM if the
                                  job can never fit the pool */
    int32_t Slot;
    int64_t FirstStep;         /* index of the batched step that first advanced
                                  the job (0-based; -1 if never) */
    int64_t DoneStep;          /* index of the batched step that ran its last
                                  step (-1 if never) */
} af_serve_result;

typedef struct af_serve_stats {
    double Wall;               /* seconds from start to the last completion */
    int64_t Steps;             /* batched steps replayed */
    int64_t CellUpdates;       /* sum over jobs of NX*NY*NZ*N */
    int32_t MaxResident;       /* peak concurrent jobs */
    int32_t Preemptions;       /* jobs snapshotted out and requeued */
    int64_t PeakBytes;         /* peak pool bytes in use */
} af_serve_stats;

/* Progress of a served run: statistics so far (Wall = seconds since the
 * start, Steps = batched steps enqueued, CellUpdates = work of the jobs
 * retired) and the number of jobs complete. */
typedef void (*af_serve_progress)(const af_serve_stats *so_far, int64_t jobs_done, void *user);

/* Runs `n` jobs (sorted by Arrival) through batch `b`, which must be empty.
 * Initial conditions: af_init_band_limited(seed, kmax, slope, rms).
 * realtime = 1: a job is visible only from its arrival time on the wall
 * clock; 0: all jobs are queued at the start.  Blocks until every job is
 * done; out[i] receives job i's result.  If `progress` is not null it is
 * called on the calling thread between batched steps, at most once per
 * `progress_interval` seconds. */
int af_serve(af_batch *b, const af_serve_job *jobs, int64_t n, int32_t kmax, double slope,
             double rms, int32_t realtime, af_serve_result *out, af_serve_stats *stats,
             af_serve_progress progress, double progress_interval, void *user);

/* ------------------------------------------------- initial conditions */
/* Seeded band-limited random velocity, built on the GPU (no reference
 * counterpart: the reference's tests seed fields through Mesh.Data,
 * engine_test.go:302-335).  Zeroes U, V and W, writes
 *   sum over integer k with 1 <= |k| <= kmax of
 *   |k|^slope cos(2 pi k.x / L + phi_{c,k})
 * at each component's interior faces (L = N*H per axis; phases are 2 pi
 * times SplitMix64 uniforms from `seed`), scales the three components so
 * their mean square over all interior faces is rms^2, and applies the
 * velocity BC (stencil.go:246-315).  The same field as the host generator
 * aeroflux_amd.ic.band_limited up to summation order. */
int af_init_band_limited(af_engine *e, uint64_t seed, int32_t kmax, double slope, double rms);
int af_batch_init_band_limited(af_batch *b, int32_t slot, uint64_t seed, int32_t kmax,
                               double slope, double rms);

/* Message of the last failure on this thread ("" if none). */
const char *af_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AEROFLUX_H */
