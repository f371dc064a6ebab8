package aeroflux

// #include "aeroflux.h"
import "C"

// Precision of the pressure solve (no reference counterpart).
const (
	// PrecisionFP64 is the reference algorithm, bit for bit.
	PrecisionFP64 = int(C.AF_PRECISION_FP64)
	// PrecisionMixed smooths with FP32 red-black passes inside an FP64
	// defect-correction loop; divergence, residual and gradient stay FP64.
	PrecisionMixed = int(C.AF_PRECISION_MIXED)
)

// Options are the build's extensions to Config (no reference counterpart).
// Their zero values select the reference's behaviour, so
// NewWithOptions(cfg, Options{}) is New(cfg).
type Options struct {
	// Precision is PrecisionFP64 (default) or PrecisionMixed.
	Precision int
	// MixedSweeps is the first cycle's FP32 sweep pairs (0: GaussSeidelIters).
	MixedSweeps int
	// MixedFixSweeps is the sweep pairs of the continued FP32 iteration and
	// of each defect-correction cycle (0: 2).
	MixedFixSweeps int
	// MixedMaxCycles bounds the FP64 residual checks of a checked solve
	// (0: 4).
	MixedMaxCycles int
	// MixedTarget is the max|div| a step leaves: 0 runs one cycle per
	// projection; > 0 checks the step's second projection and iterates
	// until the FP64 residual bounds max|div| by it (or the checks run out).
	MixedTarget float64
}

// NewWithOptions is New with the build's extensions set.
func NewWithOptions(cfg Config, opt Options) (*Engine, error) {
	if err := cfg.Normalize(); err != nil {
		return nil, err
	}
	x := cfg.toC()
	x.Precision = C.int32_t(opt.Precision)
	x.MixedSweeps = C.int32_t(opt.MixedSweeps)
	x.MixedFixSweeps = C.int32_t(opt.MixedFixSweeps)
	x.MixedMaxCycles = C.int32_t(opt.MixedMaxCycles)
	x.MixedTarget = C.double(opt.MixedTarget)
	var h *C.af_engine
	if err := codeErr(C.af_new(&x, &h)); err != nil {
		return nil, err
	}
	e := &Engine{Cfg: cfg, h: h}
	e.Mesh = newMAC(e)
	return e, nil
}

// MixedStats reports, for a mixed-precision engine, the max|div| implied by
// the last solve's last FP64 residual check (0 without a target) and the
// checks that solve ran (1 for a one-cycle solve).
func (e *Engine) MixedStats() (float64, int, error) {
	var d C.double
	var n C.int32_t
	if err := codeErr(C.af_mixed_stats(e.h, &d, &n)); err != nil {
		return 0, 0, err
	}
	return float64(d), int(n), nil
}
