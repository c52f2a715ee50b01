// host_copy.h -- large host -> device copies through pinned staging buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace nido {

// Copies `bytes` from pageable host memory to device memory on stream s and
// waits for the copy.  Large copies go through two pinned buffers (host
// copies into one split across threads while the other is in flight); a
// pageable hipMemcpy stages through one thread.  Returns the first HIP error.
hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);

// The same for n ids below 2^31 stored as int64 on the host, sent as int32
// (narrowed while staging): half the bytes over the host link.  Pinned
// staging from the same byte threshold as staged_h2d (2^28 ids by default).
hipError_t staged_h2d_ids(int* dst, const int64_t* src, int64_t n, hipStream_t s);

// Pinned staging from min_bytes (bytes crossing the host link) in chunks of
// chunk_bytes; 0 restores a default (1 GiB, 64 MiB).
int set_staging(int64_t min_bytes, int64_t chunk_bytes);

// out[i] = off[i] - off[0] for i <= n, split across threads.
void rebase_offsets(const int64_t* off, int64_t n, int64_t* out);

}  // namespace nido
