// af_snapshot.h -- arena snapshot and restore, shared by the engine
// (af_engine.hip) and the batched engine (af_batch.hip).  Not part of the C
// ABI; the format is af_snapshot_header in include/aeroflux.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/aeroflux.h"

namespace af {

// Bytes of a snapshot of an arena with these eight field descriptors.
int64_t snapshot_bytes(const af_field *f);

// Writes the header and the eight fields (field-id order, each Field.Count
// doubles) of the arena at `base` to dst (host or device memory, at least
// snapshot_bytes(f) bytes).  `acc` is one device uint64 of scratch.
// Synchronises `s`.  Returns AF_OK or an error code (message recorded).
int snapshot_save(const af_field *f, unsigned char *base, const af_config &c, void *dst,
                  int64_t bytes, unsigned long long *acc, hipStream_t s);

// Checks src's header against the arena's geometry before touching the
// arena, then copies the fields in and verifies the checksum; on a checksum
// mismatch the eight fields are zeroed and AF_ERR_ARG is returned.
// Synchronises `s`.
int snapshot_load(const af_field *f, unsigned char *base, const af_config &c, const void *src,
                  int64_t bytes, unsigned long long *acc, hipStream_t s);

}  // namespace af
