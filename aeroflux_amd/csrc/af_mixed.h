// af_mixed.h -- mixed-precision pressure solve: FP32 red-black smoothing
// inside an FP64 defect-correction loop.  No reference counterpart (the
// reference solve, project.go:73-100, is FP64 only); the FP64 path is its
// oracle.
#pragma once

#include "af_internal.h"

namespace af {

struct Tile {
    int k = -1;     // colour passes per blocked launch; <= 0: one pass per launch
    int rows = 2;   // rows per thread
    int cz = 16;    // z-planes per workgroup
};

struct MixedState {
    float *e[2] = {nullptr, nullptr};  // FP32 correction (ping-pong), P's padded layout
    float *r = nullptr;                // FP32 right-hand side / residual
    double *partials = nullptr;        // residual-max partials
    int *done = nullptr;               // device: set once the target is met
    double *last_res = nullptr;        // device: last max|div| implied by the residual
    int *cycles = nullptr;             // device: cycles run since the last reset
    int np = 0;
};

struct MixedCtx {
    DField P, D;       // FP64 pressure (output) and divergence (input)
    Geom G;
    double h2;         // H*H
    int sweeps;        // FP32 sweep pairs in the first cycle
    int fix_sweeps;    // FP32 sweep pairs in each later cycle
    int cycles;        // maximum cycles
    double target;     // max|div| target after the projection (0: one cycle)
    double res_to_div; // DT/(rho*H^2): max|div| = res_to_div * max|residual|
    Tile tile;         // FP32 red-black tiling
};

hipError_t init_mixed(MixedState &m, Geom G, DField P, hipStream_t s);
void release_mixed(MixedState &m);
// Enqueue one mixed-precision solve writing the FP64 P (ghosts filled).  The
// FP32 right-hand side m.r must already hold -(H^2 div), which the
// divergence kernel writes in mixed mode.
hipError_t enqueue_mixed_solve(MixedState &m, const MixedCtx &c, hipStream_t s);

}  // namespace af
