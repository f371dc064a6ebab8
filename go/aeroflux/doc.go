// Package aeroflux is the reference's Go API (AeroFlex/aeroflux.go:14-439,
// grid.go:8-193, project.go:210-258, stencil.go:319-338) over the MI355X
// step data plane: every call crosses a thin cgo shim into libaeroflux.so
// (include/aeroflux.h), whose hand-written HIP kernels run the step on the
// GPU.  Callers keep aeroflux.New(Config), Step(), MaxDivergence() and
// Close().
//
// What changes for a caller:
//
//   - MAC.Data cannot alias device memory.  It returns a host mirror that is
//     refreshed from the GPU on every call and uploaded again before the next
//     call that runs GPU work (Step, the diagnostics), because the caller may
//     have written through it, as the reference's tests do when they seed a
//     field.  That upload costs one host-to-device copy of the field (138 MB
//     at 256^3).  A caller that only reads uses ReadData, which never marks
//     the mirror for upload.  A slice held across Step goes stale; call Data
//     again to see the new values.
//   - A contained device fault (a NaN departure coordinate, the reference's
//     guard-page trigger) comes back from Step as an error that wraps
//     ErrDeviceFault.
//   - EnableSparseSolver keeps the non-Darwin reference's error;
//     EnableDirectSolver switches on the GPU's exact spectral solve of the
//     same pinned system instead.
//
// No Go pointer is kept by C, and the library starts no threads; an Engine
// must not be used from two goroutines at once, as in the reference.
package aeroflux
