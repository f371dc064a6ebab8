package aeroflux

/*
#cgo CFLAGS: -I${SRCDIR}/../../include
#cgo LDFLAGS: -L${SRCDIR}/../../aeroflux_amd -laeroflux -Wl,-rpath,${SRCDIR}/../../aeroflux_amd
#include <stdlib.h>
#include "aeroflux.h"
*/
import "C"

import (
	"errors"
	"fmt"
)

// ErrDeviceFault is wrapped by the error Step returns when the GPU raised
// its fault flag: a departure coordinate was NaN, the condition under which
// the reference's sampleTrilinear leaves the arena and its guard page
// (arena.go:230-243) turns the fault into "aeroflux: guard-page fault
// contained: ...".  As there, the state after the error is unspecified.
var ErrDeviceFault = errors.New("aeroflux: device fault contained")

// codeErr turns a C return code into the Go error the reference would
// return: nil for AF_OK, the library's "aeroflux: ..." message otherwise,
// with ErrDeviceFault wrapped for a contained fault.
func codeErr(rc C.int) error {
	if rc == C.AF_OK {
		return nil
	}
	msg := C.GoString(C.af_last_error())
	if msg == "" {
		msg = fmt.Sprintf("aeroflux: error code %d", int(rc))
	}
	if rc == C.AF_ERR_FAULT {
		return fmt.Errorf("%s: %w", msg, ErrDeviceFault)
	}
	return errors.New(msg)
}

// RecoverFault keeps the reference's exported helper (arena.go:222-243):
// deferred in a goroutine, it converts a runtime-error panic into an error
// stored in *dst and lets every other panic propagate.  The GPU path never
// panics for a fault -- the device flag reports it -- so only callers that
// used the helper in their own workers still need it.
func RecoverFault(dst *error) {
	r := recover()
	if r == nil {
		return
	}
	type runtimeError interface {
		error
		RuntimeError()
	}
	if re, ok := r.(runtimeError); ok {
		*dst = fmt.Errorf("aeroflux: guard-page fault contained: %w", re)
		return
	}
	panic(r)
}
