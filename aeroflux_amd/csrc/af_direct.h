// af_direct.h -- direct pressure solve on the GPU: the pinned Neumann system
// that the reference's Darwin build factors with a sparse Cholesky
// (EnableSparseSolver / solvePressureDirect, sparse_darwin.go:19-172),
// solved spectrally.
#pragma once

#include <rocblas/rocblas.h>

#include "af_internal.h"

namespace af {

struct DirectState {
    int nx = 0, ny = 0, nz = 0;
    double h2 = 0;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;  // orthonormal DCT-II, column-major
    double *lx = nullptr, *ly = nullptr, *lz = nullptr;  // 1-D eigenvalues (divided by H^2)
    double *w0 = nullptr, *w1 = nullptr;                 // compact interior work arrays
    double *partials = nullptr;                          // sum partials
    double *sum = nullptr;                               // device: S = sum of the rhs
    rocblas_handle blas = nullptr;
    bool ready() const { return blas != nullptr; }
};

// Allocates the transforms and work arrays for grid G (interior NX x NY x NZ),
// binds the BLAS handle to stream s and runs each transform once, so later
// enqueues (also under stream capture) allocate nothing.
hipError_t init_direct(DirectState &d, const Geom &G, double H, hipStream_t s, int *blas_status);
void release_direct(DirectState &d);

// Enqueues the solve of (L + e0 e0^T / H^2) p = -Div over the interior (L the
// 7-point Neumann Laplacian over H^2, e0 the first cell), writes p to P's
// interior and fills P's first ghost layer (applyPressureBC).
hipError_t enqueue_direct_solve(DirectState &d, const DField &P, const DField &Div, const Geom &G,
                                hipStream_t s, int *blas_status);

}  // namespace af
