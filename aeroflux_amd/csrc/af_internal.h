// af_internal.h -- shared declarations between the engine (af_engine.hip)
// and the kernel launchers (af_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace af {

// One field as the kernels see it: base pointer and padded geometry.
// Element counts are checked on the host to fit int32 (max 2^31-1 per field).
struct DField {
    double *d;
    int sy, sz;
    int nx, ny, nz;
};

// Interior extents and ghost width.
struct Geom {
    int NX, NY, NZ, g;
};

struct SweepArgs {
    DField U, V, W;        // current velocity descriptors (sources)
    DField T0, T1, T2;     // destinations (Tmp0/1/2)
    Geom G;
    double dt, invH, lam;
    int *flag;             // set to 1 on a NaN departure coordinate
    int shape = 4;         // workgroup shape, 0 .. sweep_shape_count()-1
    int slab = 1 << 30;    // rows per y-slab of the block order (>= NY+1: plane order)
    int bc = 0;            // also apply applyVelocityBC to the output (body::store_face)
    int zchunks = 0;       // LDS-staged shapes: z-chunks per tile column (0: the shape's own)
};

// Number of sweep shapes: gathering shapes (rows per block x z per thread)
// followed by the LDS-staged variants.
int sweep_shape_count();
int sweep_lds_variant_count();
hipError_t launch_sweep_lds(int axis, const SweepArgs &a, int variant, hipStream_t s);
int sweep_lds_chunks(const Geom &G, int variant);
int sweep_lds_default_chunks(const Geom &G, int variant);

// Launchers.  All enqueue on `s` and return the hipGetLastError() code.
hipError_t launch_velocity_bc(DField U, DField V, DField W, Geom G, hipStream_t s);
hipError_t launch_sweep(int axis, const SweepArgs &a, hipStream_t s);
// split: 0 -> Div only; 1 -> also the FP64 colour-split copy (red, black
// are double arrays); 2 -> also the FP32 split right-hand side -(h2 * div)
// of the mixed-precision solve (float arrays) and its first FP32 red pass
// from zero into `first` (float, the correction's red array); 3 -> as 1, and
// also the solve's first colour pass from zero into `first` (double, the
// solution's red array): each red cell's value depends only on its own
// right-hand side.  store_div false (split 2 or 3 only) skips the store to
// D, for a Div that nothing reads before the next divergence overwrites it.
hipError_t launch_divergence(DField U, DField V, DField W, DField D, Geom G, double scale,
                             hipStream_t s, int split = 0, void *red = nullptr,
                             void *black = nullptr, double h2 = 0, void *first = nullptr,
                             bool store_div = true);
// Red-black colour pass.  fused_bc: a neighbour outside the interior reads the
// cell's own value (the Neumann mirror applied just before the pass).
hipError_t launch_rbgs(DField P, DField D, Geom G, double h2, int color, bool fused_bc,
                       hipStream_t s);
// First red pass of a solve from P == 0: writes every interior cell.
hipError_t launch_rbgs_first(DField P, DField D, Geom G, double h2, hipStream_t s);
// Pressure ghost fill (Neumann mirror); outer layers (ghost > 1) are zeroed
// when zero_outer is set.  Gated launches (gate != nullptr, here and in the
// gradients below) run only when *gate == run_if, read on the device.
hipError_t launch_pressure_bc(DField P, Geom G, bool zero_outer, hipStream_t s,
                              const int *gate = nullptr, int run_if = 1);
hipError_t launch_zero(double *d, int64_t n, hipStream_t s);
// velocity_bc: also apply applyVelocityBC to the updated velocity (the BC
// that follows subtractGradient in project), so that launch can be skipped.
hipError_t launch_gradient(DField U, DField V, DField W, DField P, Geom G, double scale,
                           hipStream_t s, bool velocity_bc = false, const int *gate = nullptr,
                           int run_if = 1);
// subtractGradient with the velocity BC, P read from the colour-split
// solution (red, black: the FP64 solution, or the mixed solve's FP32
// correction after its one cycle) instead of P: the values of merging it
// into P and applying the pressure BC first.  With P, the kernel also
// writes the solution to P's interior (the merge; P's ghost fill is the
// caller's).
template <typename TS>
hipError_t launch_gradient_split(DField U, DField V, DField W, const TS *red, const TS *black,
                                 Geom G, double scale, hipStream_t s, const DField *P = nullptr,
                                 const int *gate = nullptr, int run_if = 1);
hipError_t launch_set_int(int *p, int v, hipStream_t s);

// Reductions: two launches, fixed tree, no atomics.  `partials` must hold
// reduce_partials_needed(G) doubles; the result lands in *out (device).
int reduce_partials_needed(Geom G);
// Partials launch_max_divergence itself writes (<= reduce_partials_needed).
int max_divergence_partials(Geom G);
hipError_t launch_max_divergence(DField U, DField V, DField W, Geom G, double invH,
                                 double *partials, double *out, hipStream_t s);
hipError_t launch_kinetic_energy(DField U, DField V, DField W, Geom G, double *partials,
                                 double *out, hipStream_t s);
hipError_t launch_max_abs3(const double *u, int64_t nu, const double *v, int64_t nv,
                           const double *w, int64_t nw, double *partials, double *out,
                           hipStream_t s);

}  // namespace af
