// af_mixed.h -- mixed-precision pressure solve: FP32 red-black smoothing
// inside an FP64 defect-correction loop.  No reference counterpart (the
// reference solve, project.go:73-100, is FP64 only); the FP64 path is its
// oracle.
#pragma once

#include "af_internal.h"

namespace af {

struct MixedState {
    float *e_red = nullptr, *e_black = nullptr;  // FP32 correction, colour-split
    float *r_red = nullptr, *r_black = nullptr;  // FP32 right-hand side / residual, split
    double *partials = nullptr;                  // residual-max partials
    int *done = nullptr;                         // device: set once the target is met
    double *last_res = nullptr;                  // device: last max|div| implied by the residual
    int *cycles = nullptr;                       // device: cycles run since the last reset
    // device: a defect-correction cycle merged the solution into P (else
    // the whole solution is the correction e in the split arrays)
    int *merged = nullptr;
    int np = 0;        // residual partials (one per 64 x 4 column)
    int np_merge = 0;  // first-cycle merge + residual partials (column x z-chunk)
};

struct MixedCtx {
    DField P, D;       // FP64 pressure (output) and divergence (input)
    Geom G;
    double h2;         // H*H
    int sweeps;        // FP32 sweep pairs in the first cycle
    int fix_sweeps;    // FP32 sweep pairs in each later cycle
    int cycles;        // maximum cycles
    // max|div| target after the projection (0: one cycle).  A checked solve
    // runs the first cycle's sweeps, then checks the FP64 residual of p = e;
    // while the target is not met it continues the FP32 iteration by
    // fix_sweeps pairs once, then runs FP64 defect-correction cycles.
    double target;
    double res_to_div; // DT/(rho*H^2): max|div| = res_to_div * max|residual|
    // The split right-hand side is already in r_red/r_black, and the first
    // cycle's first red pass in e_red (k_divergence<2>).
    bool rhs_ready = false;
    // false leaves P unwritten (no merge, no pressure BC) for a caller that
    // reads the correction e, the whole of the solution, from the split
    // arrays -- unless a checked solve ran a defect-correction cycle, which
    // merges the solution into P and fills P's ghosts (`merged` raised).
    bool merge = true;
    int slab = 1 << 30;      // workgroup y-slab order of the FP32 passes (SplitArgs::slab)
};

hipError_t init_mixed(MixedState &m, Geom G, hipStream_t s);
void release_mixed(MixedState &m);
// Enqueue one mixed-precision solve writing the FP64 P (ghosts filled).
hipError_t enqueue_mixed_solve(MixedState &m, const MixedCtx &c, hipStream_t s);

}  // namespace af
