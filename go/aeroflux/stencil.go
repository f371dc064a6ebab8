package aeroflux

// #include "aeroflux.h"
import "C"

// MaxVelocity is the largest |component| over the whole padded U, V and W
// arrays (stencil.go:319-331).
func (e *Engine) MaxVelocity() float64 {
	return e.scalar(func(h *C.af_engine, v *C.double) C.int { return C.af_max_velocity(h, v) })
}

// CFL is MaxVelocity * DT / H (stencil.go:336-338).
func (e *Engine) CFL() float64 {
	return e.scalar(func(h *C.af_engine, v *C.double) C.int { return C.af_cfl(h, v) })
}

// applyVelocityBC is the unexported boundary fill the reference's tests
// call after seeding (engine_test.go:334; stencil.go:246-315).
func (e *Engine) applyVelocityBC() error {
	if err := e.Mesh.flush(); err != nil {
		return err
	}
	return codeErr(C.af_apply_velocity_bc(e.h))
}
