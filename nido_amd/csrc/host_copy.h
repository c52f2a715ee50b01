// host_copy.h -- large host -> device copies through pinned staging buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <vector>

namespace nido {

// Checks of the arcs made while they pass through the staging buffers: arc
// j of graph g (gstart[g] <= j < gstart[g + 1], relative to the first arc)
// needs its column in [vptr[g], vptr[g + 1]) and a finite weight of
// magnitude below 2^62.  Records the first arc that breaks a rule, whether
// any weight is negative, and each graph's sum of |w| in FP64 (a bound with
// room to spare: the exact totals are taken on the device).
struct ArcCheck {
    const int64_t* gstart = nullptr;
    const int64_t* vptr = nullptr;
    int G = 1;
    std::atomic<int64_t> bad{INT64_MAX};
    std::atomic<bool> neg{false};
    std::vector<double> wabs;   // G
    std::mutex mu;
};

// Whether an upload of this many bytes goes through the pinned staging
// buffers (where the ArcCheck variants check on the way).
bool staged_path(size_t bytes);

// Copies `bytes` from pageable host memory to device memory on stream s and
// waits for the copy.  Large copies go through two pinned buffers (host
// copies into one split across threads while the other is in flight); a
// pageable hipMemcpy stages through one thread.  Returns the first HIP error.
hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);

// Copies `bytes` from device memory to pageable host memory on stream s and
// waits for the copy.  From the staging threshold the copy goes through two
// pinned buffers: while one chunk crosses the host link, the chunk before it
// is copied out of the other by all host threads (which also take the
// destination's first-touch page faults in parallel).
hipError_t staged_d2h(void* dst, const void* src, size_t bytes, hipStream_t s);

// The same for n ids below 2^31 stored as int64 on the host, sent as int32
// (narrowed while staging): half the bytes over the host link.  Pinned
// staging from the same byte threshold as staged_h2d (2^26 ids by default).
hipError_t staged_h2d_ids(int* dst, const int64_t* src, int64_t n, hipStream_t s, ArcCheck* chk = nullptr);

// The ids of a handle whose graphs all have at most 2^16 vertices, sent as
// 16-bit ids local to their graph (a quarter of the int64 bytes; the device
// adds each graph's first vertex back), through the staging buffers and
// checked by chk on the way.
hipError_t staged_h2d_local_ids(uint16_t* dst, const int64_t* src, int64_t n, hipStream_t s, ArcCheck& chk);

// Weights through the staging buffers, checked by chk on the way (the
// caller has made sure staged_path holds for n weights).
hipError_t staged_h2d_weights(double* dst, const double* src, int64_t n, hipStream_t s, ArcCheck& chk);

// Pinned staging from min_bytes (bytes crossing the host link) in chunks of
// chunk_bytes; 0 restores a default (256 MiB, 64 MiB).
int set_staging(int64_t min_bytes, int64_t chunk_bytes);

// out[i] = off[i] - off[0] for i <= n, split across threads.
void rebase_offsets(const int64_t* off, int64_t n, int64_t* out);

}  // namespace nido
