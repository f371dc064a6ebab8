package aeroflux

// #include "aeroflux.h"
import "C"

import "math"

// scalar runs one GPU reduction after uploading pending mirrors.  The
// reference's diagnostics return a bare float64, so a failure (a closed
// engine, a device error) reads as NaN.
func (e *Engine) scalar(fn func(*C.af_engine, *C.double) C.int) float64 {
	if e.Mesh.flush() != nil {
		return math.NaN()
	}
	var v C.double
	if fn(e.h, &v) != C.AF_OK {
		return math.NaN()
	}
	return float64(v)
}

// MaxDivergence is the largest |div u| over interior cells, NaN cells
// skipped (project.go:210-232), as a fixed-order device reduction.
func (e *Engine) MaxDivergence() float64 {
	return e.scalar(func(h *C.af_engine, v *C.double) C.int { return C.af_max_divergence(h, v) })
}

// KineticEnergy is 0.5 * H^3 * sum of squared cell-centre velocities
// (project.go:238-258); rows are summed in the reference's order and
// combined by a fixed tree, within 1e-12 relative of the serial sum.
func (e *Engine) KineticEnergy() float64 {
	return e.scalar(func(h *C.af_engine, v *C.double) C.int { return C.af_kinetic_energy(h, v) })
}

// project is the unexported projection the reference's in-package tests
// call (engine_test.go:121; project.go:12-34).
func (e *Engine) project() error {
	if err := e.Mesh.flush(); err != nil {
		return err
	}
	return codeErr(C.af_project(e.h))
}
