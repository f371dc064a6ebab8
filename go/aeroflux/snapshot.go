package aeroflux

/*
#include <stdlib.h>
#include "aeroflux.h"
*/
import "C"

// Snapshot returns the engine's whole state in the library's snapshot
// format (include/aeroflux.h: a 128-byte header and the eight padded
// fields).  No reference counterpart; Restore puts it back into an engine
// or batch slot of the same grid, and stepping on gives the same bits as
// never having stopped.
func (e *Engine) Snapshot() ([]byte, error) {
	if err := e.Mesh.flush(); err != nil { // edits through Data views first
		return nil, err
	}
	var n C.int64_t
	if err := codeErr(C.af_snapshot_bytes(e.h, &n)); err != nil {
		return nil, err
	}
	buf := C.malloc(C.size_t(n)) // C memory: no Go pointer crosses the boundary
	defer C.free(buf)
	if err := codeErr(C.af_snapshot(e.h, buf, n)); err != nil {
		return nil, err
	}
	return C.GoBytes(buf, C.int(n)), nil
}

// Restore loads a snapshot taken from an engine or slot of the same grid.
// Pending writes through Data views are dropped: they would overwrite the
// restored state at the next Step.
func (e *Engine) Restore(b []byte) error {
	if len(b) == 0 {
		return codeErr(C.af_restore(e.h, nil, 0))
	}
	buf := C.CBytes(b)
	defer C.free(buf)
	if err := codeErr(C.af_restore(e.h, buf, C.int64_t(len(b)))); err != nil {
		return err
	}
	for id := range e.Mesh.dirty {
		e.Mesh.dirty[id] = false
	}
	return nil
}
