package aeroflux

// #include "aeroflux.h"
import "C"

// Config mirrors the reference's Config (aeroflux.go:14-41) field for field;
// the GPU keeps the same meanings.  Workers is normalised as the reference
// normalises it but does not change the GPU's schedule; Lock is accepted
// (device memory is always resident); UseSparseSolver is accepted and, as
// on the reference's non-Darwin build, leaves the red-black solve in place.
type Config struct {
	NX, NY, NZ       int
	Ghost            int
	H                float64
	Viscosity        float64
	Density          float64
	DT               float64
	Workers          int
	GaussSeidelIters int
	UseSparseSolver  bool
	Lock             bool
}

func b32(v bool) C.int32_t {
	if v {
		return 1
	}
	return 0
}

func (c *Config) toC() C.af_config {
	var x C.af_config
	x.NX, x.NY, x.NZ = C.int64_t(c.NX), C.int64_t(c.NY), C.int64_t(c.NZ)
	x.Ghost = C.int64_t(c.Ghost)
	x.H, x.Viscosity = C.double(c.H), C.double(c.Viscosity)
	x.Density, x.DT = C.double(c.Density), C.double(c.DT)
	x.Workers, x.GaussSeidelIters = C.int64_t(c.Workers), C.int64_t(c.GaussSeidelIters)
	x.UseSparseSolver, x.Lock = b32(c.UseSparseSolver), b32(c.Lock)
	return x
}

func (c *Config) fromC(x *C.af_config) {
	c.NX, c.NY, c.NZ, c.Ghost = int(x.NX), int(x.NY), int(x.NZ), int(x.Ghost)
	c.H, c.Viscosity, c.Density, c.DT = float64(x.H), float64(x.Viscosity), float64(x.Density), float64(x.DT)
	c.Workers, c.GaussSeidelIters = int(x.Workers), int(x.GaussSeidelIters)
}

// Normalize fills defaults and validates with the reference's checks, in its
// order and with its messages (aeroflux.go:47-113).
func (c *Config) Normalize() error {
	x := c.toC()
	rc := C.af_config_normalize(&x)
	if rc == C.AF_OK {
		c.fromC(&x)
	}
	return codeErr(rc)
}

// Engine owns one GPU engine: the HBM arena in the reference's MAC layout
// and the per-step graph (aeroflux.go:133-172).
type Engine struct {
	Cfg  Config
	Mesh *MAC
	h    *C.af_engine
}

// New mirrors aeroflux.go:176-213: normalise, allocate the zeroed arena (on
// the current HIP device), lay out the grid.
func New(cfg Config) (*Engine, error) {
	if err := cfg.Normalize(); err != nil {
		return nil, err
	}
	x := cfg.toC()
	var h *C.af_engine
	if err := codeErr(C.af_new(&x, &h)); err != nil {
		return nil, err
	}
	e := &Engine{Cfg: cfg, h: h}
	e.Mesh = newMAC(e)
	return e, nil
}

func (e *Engine) handle() *C.af_engine { return e.h }

// Step advances one timestep, the Strang palindrome X, Y, Z, project, Z, Y,
// X, project of aeroflux.go:348-409, replayed on the GPU as one graph.
// Mirrors written through Data are uploaded first.
func (e *Engine) Step() error {
	if err := e.Mesh.flush(); err != nil {
		return err
	}
	return codeErr(C.af_step(e.h))
}

// StepN runs n steps with one host synchronisation (no reference
// counterpart: a loop of Step calls with the same results).
func (e *Engine) StepN(n int) error {
	if err := e.Mesh.flush(); err != nil {
		return err
	}
	return codeErr(C.af_step_n(e.h, C.int64_t(n), nil))
}

// Close releases the engine's device memory (aeroflux.go:419-439).  It is
// idempotent; mirrors obtained from Data stay valid host slices but no
// longer track anything.
func (e *Engine) Close() error {
	if e.h == nil {
		return nil
	}
	err := codeErr(C.af_close(e.h))
	C.af_free(e.h)
	e.h = nil
	return err
}
