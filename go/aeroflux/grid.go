package aeroflux

// #include "aeroflux.h"
import "C"

import (
	"errors"
	"unsafe"
)

// Field describes one padded field of the arena exactly as the reference's
// Field does (grid.go:12-21): byte offset, padded extents, strides, count.
type Field struct {
	Off        int
	NX, NY, NZ int
	SY, SZ     int
	Count      int
	id         C.int
}

// Index returns the linear element index of padded (x, y, z) (grid.go:24-26).
func (f *Field) Index(x, y, z int) int { return x + y*f.SY + z*f.SZ }

// ByteOffset returns the arena byte offset of padded (x, y, z) (grid.go:29-31).
func (f *Field) ByteOffset(x, y, z int) int { return f.Off + f.Index(x, y, z)*8 }

// MAC is the staggered grid over the engine's arena (grid.go:52-73): the
// same eight fields in the same order and layout (NewMAC, grid.go:101-134).
type MAC struct {
	NX, NY, NZ, Ghost                 int
	H                                 float64
	P, U, V, W, Div, Tmp0, Tmp1, Tmp2 Field
	Bytes                             int

	e      *Engine
	mirror [fieldCount][]float64
	dirty  [fieldCount]bool
}

// fieldCount is AF_FIELD_COUNT: P, U, V, W, Div, Tmp0, Tmp1, Tmp2.
const fieldCount = 8

func newMAC(e *Engine) *MAC {
	m := &MAC{NX: e.Cfg.NX, NY: e.Cfg.NY, NZ: e.Cfg.NZ, Ghost: e.Cfg.Ghost, H: e.Cfg.H, e: e}
	m.Bytes = int(C.af_arena_bytes(e.h))
	m.refreshDescriptors()
	return m
}

func (m *MAC) fields() [fieldCount]*Field {
	return [fieldCount]*Field{&m.P, &m.U, &m.V, &m.W, &m.Div, &m.Tmp0, &m.Tmp1, &m.Tmp2}
}

// refreshDescriptors reads the current descriptor of every role.  A Step
// returns U, V and W to their own buffers (six commitVelocity swaps), so
// they change only under the stage-level entry points.
func (m *MAC) refreshDescriptors() {
	for id, f := range m.fields() {
		var d C.af_field
		C.af_field_desc(m.e.h, C.int(id), &d)
		*f = Field{Off: int(d.Off), NX: int(d.NX), NY: int(d.NY), NZ: int(d.NZ),
			SY: int(d.SY), SZ: int(d.SZ), Count: int(d.Count), id: C.int(id)}
	}
}

var errClosed = errors.New("aeroflux: engine is closed")

func (m *MAC) read(f *Field) []float64 {
	if m.e.h == nil {
		return nil
	}
	buf := m.mirror[f.id]
	if len(buf) != f.Count {
		buf = make([]float64, f.Count)
		m.mirror[f.id] = buf
	}
	// A pending write through the mirror wins over the device copy.
	if !m.dirty[f.id] {
		C.af_field_read(m.e.h, f.id, (*C.double)(unsafe.Pointer(&buf[0])), C.int64_t(len(buf)))
	}
	return buf
}

// Data returns the field as a host slice (MAC.Data, grid.go:147-149).  The
// reference returns a live view of the arena; device memory cannot be
// aliased, so this is a mirror: refreshed from the GPU now, and uploaded
// again before the next GPU work because the caller may write through it.
// For read-only use, ReadData avoids that upload.
func (m *MAC) Data(f *Field) []float64 {
	buf := m.read(f)
	if buf != nil {
		m.dirty[f.id] = true
	}
	return buf
}

// ReadData returns the field's current values without marking the mirror
// for upload (no reference counterpart).  Writes through the returned slice
// are not seen by the engine.
func (m *MAC) ReadData(f *Field) []float64 { return m.read(f) }

// flush uploads the mirrors written through Data, once.
func (m *MAC) flush() error {
	if m.e.h == nil {
		return errClosed
	}
	for id := range m.dirty {
		if !m.dirty[id] {
			continue
		}
		buf := m.mirror[id]
		rc := C.af_field_write(m.e.h, C.int(id), (*C.double)(unsafe.Pointer(&buf[0])), C.int64_t(len(buf)))
		if err := codeErr(rc); err != nil {
			return err
		}
		m.dirty[id] = false
	}
	return nil
}

// Ptr returns the device address of element (x, y, z) of f (grid.go:152-155
// hands a host address to Accelerate; here the arena lives in HBM).
func (m *MAC) Ptr(f *Field, x, y, z int) uintptr {
	var base C.uint64_t
	if C.af_field_device_ptr(m.e.h, f.id, &base) != C.AF_OK {
		return 0
	}
	return uintptr(base) + uintptr(f.Index(x, y, z)*8)
}

// Zero clears every field (grid.go:158-163), on the device.
func (m *MAC) Zero() error {
	for _, f := range m.fields() {
		buf := make([]float64, f.Count)
		rc := C.af_field_write(m.e.h, f.id, (*C.double)(unsafe.Pointer(&buf[0])), C.int64_t(len(buf)))
		if err := codeErr(rc); err != nil {
			return err
		}
		m.mirror[f.id], m.dirty[f.id] = nil, false
	}
	return nil
}

// InteriorX, InteriorY, InteriorZ return half-open padded interior ranges
// of the cell-centred fields (grid.go:166-172).
func (m *MAC) InteriorX() (lo, hi int) { return m.Ghost, m.Ghost + m.NX }
func (m *MAC) InteriorY() (lo, hi int) { return m.Ghost, m.Ghost + m.NY }
func (m *MAC) InteriorZ() (lo, hi int) { return m.Ghost, m.Ghost + m.NZ }

// CellCount is the number of interior cells (grid.go:176).
func (m *MAC) CellCount() int { return m.NX * m.NY * m.NZ }

// CellIndex maps 0-based interior cell coordinates to the Poisson system's
// index, x fastest (grid.go:181-183).
func (m *MAC) CellIndex(i, j, k int) int { return i + j*m.NX + k*m.NX*m.NY }

// MACBytes is the arena footprint of a grid (grid.go:138-144).
func MACBytes(nx, ny, nz, ghost int) int {
	return int(C.af_mac_bytes(C.int64_t(nx), C.int64_t(ny), C.int64_t(nz), C.int64_t(ghost)))
}
