// af_direct.h -- direct pressure solve on the GPU: the pinned Neumann system
// that the reference's Darwin build factors with a sparse Cholesky
// (EnableSparseSolver / solvePressureDirect, sparse_darwin.go:19-172),
// solved spectrally.
#pragma once

#include <rocblas/rocblas.h>

#include "af_internal.h"

namespace af {

// How the transforms run: FFT passes (every extent a power of two in
// 8 .. 512, af_dct.hip) or dense GEMMs against N x N DCT matrices (rocBLAS).
enum DirectMethod { DIRECT_AUTO = 0, DIRECT_FFT = 1, DIRECT_DENSE = 2 };

struct DirectState {
    int nx = 0, ny = 0, nz = 0;
    double h2 = 0;
    int method = 0;                                      // DIRECT_FFT or DIRECT_DENSE once set up
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;  // dense: orthonormal DCT-II, column-major
    double *lx = nullptr, *ly = nullptr, *lz = nullptr;  // 1-D eigenvalues (divided by H^2)
    double *w0 = nullptr, *w1 = nullptr;                 // compact interior work arrays
    double *partials = nullptr;                          // sum partials
    int npartials = 0;
    double *sum = nullptr;                               // device: S = sum of the rhs
    double *shift = nullptr;                             // FFT: device H^2 S - q_0
    // FFT tables per axis: twiddles, post/pre-twiddles ((re, im) pairs),
    // inverse scales and C(k, 0).
    double *tw[3] = {}, *pw[3] = {}, *wc[3] = {}, *iq[3] = {}, *c0[3] = {};
    rocblas_handle blas = nullptr;
    bool ready() const { return method != 0; }
};

// Whether the FFT passes cover grid G.
bool direct_fft_supported(const Geom &G);

// Allocates the transforms and work arrays for grid G (interior NX x NY x NZ)
// for `method` (DIRECT_AUTO picks FFT where supported); the dense path binds
// the BLAS handle to stream s and runs each transform once, so later enqueues
// (also under stream capture) allocate nothing.
hipError_t init_direct(DirectState &d, const Geom &G, double H, int method, hipStream_t s,
                       int *blas_status);
void release_direct(DirectState &d);

// Enqueues the solve of (L + e0 e0^T / H^2) p = -Div over the interior (L the
// 7-point Neumann Laplacian over H^2, e0 the first cell), writes p to P's
// interior and fills P's first ghost layer (applyPressureBC).
hipError_t enqueue_direct_solve(DirectState &d, const DField &P, const DField &Div, const Geom &G,
                                hipStream_t s, int *blas_status);

// The FFT form with the projection's divergence folded into its first pass:
// computeDivergence(U, V, W) with `scale` is written to Div (bit-identical
// to the divergence kernel) and solved for in one go.  FFT method only.
hipError_t enqueue_direct_project_solve(DirectState &d, const DField &P, const DField &Div,
                                        const DField &U, const DField &V, const DField &W,
                                        double scale, const Geom &G, hipStream_t s);

}  // namespace af
