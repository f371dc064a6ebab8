package aeroflux

// #include "aeroflux.h"
import "C"

// EnableSparseSolver keeps the reference's non-Darwin behaviour
// (sparse_fallback.go:17-19): it always fails with the same message, and
// the red-black solve stays in use.
func (e *Engine) EnableSparseSolver() error {
	return codeErr(C.af_enable_sparse_solver(e.h))
}

// EnableDirectSolver has no counterpart in the reference's Linux build: it
// switches every later projection to an exact spectral solve, on the GPU, of
// the pinned Neumann system the Darwin EnableSparseSolver factors
// (sparse_darwin.go:19-172).
func (e *Engine) EnableDirectSolver() error {
	return codeErr(C.af_enable_direct_solver(e.h))
}
