// host_copy.cpp -- large host -> device copies through pinned staging buffers.
#include "host_copy.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "../../include/nido_amd.h"
#include "errors.h"

namespace nido {

namespace {
// Copies of at least kMinBytes go through two pinned buffers of kChunkBytes
// each (below 1 GiB the runtime's own staging is as fast); both can be
// lowered so that tests reach the pinned path with small inputs.
constexpr int64_t kDefaultMinBytes = int64_t(1) << 30;
constexpr int64_t kDefaultChunkBytes = int64_t(64) << 20;
std::atomic<int64_t> g_min_bytes{kDefaultMinBytes};
std::atomic<int64_t> g_chunk_bytes{kDefaultChunkBytes};
}  // namespace

hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const size_t kChunk = (size_t)g_chunk_bytes.load();
    if (bytes < (size_t)g_min_bytes.load()) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipHostMalloc(reinterpret_cast<void**>(&buf[i]), kChunk);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
    }
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    for (size_t at = 0, i = 0; e == hipSuccess && at < bytes; at += kChunk, i ^= 1) {
        const size_t n = std::min(kChunk, bytes - at);
        e = hipEventSynchronize(done[i]);      // the buffer's previous copy has landed
        if (e != hipSuccess) break;
        constexpr size_t kPiece = size_t(1) << 20;
        const int64_t pieces = (int64_t)((n + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < pieces; ++p) {
            const size_t o = (size_t)p * kPiece;
            std::memcpy(buf[i] + o, in + at + o, std::min(kPiece, n - o));
        }
        e = hipMemcpyAsync(out + at, buf[i], n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i], s);
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    for (int i = 0; i < 2; ++i) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (buf[i]) (void)hipHostFree(buf[i]);
    }
    return e;
}

hipError_t staged_h2d_ids(int* dst, const int64_t* src, int64_t n, hipStream_t s)
{
    const int64_t kChunk = g_chunk_bytes.load() / (int64_t)sizeof(int);   // ids per pinned buffer
    if (n * (int64_t)sizeof(int) < g_min_bytes.load()) {   // one pageable copy, no pinned allocation
        std::vector<int> t((size_t)n);
        // threads only for graphs larger than a serving request, whose
        // requests already run side by side
#pragma omp parallel for schedule(static) if (n >= (int64_t(1) << 26))
        for (int64_t j = 0; j < n; ++j) t[(size_t)j] = (int)src[j];
        hipError_t e = hipMemcpyAsync(dst, t.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    int* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipHostMalloc(reinterpret_cast<void**>(&buf[i]), sizeof(int) * (size_t)kChunk);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
    }
    for (int64_t at = 0, i = 0; e == hipSuccess && at < n; at += kChunk, i ^= 1) {
        const int64_t m = std::min(kChunk, n - at);
        e = hipEventSynchronize(done[i]);      // the buffer's previous copy has landed
        if (e != hipSuccess) break;
        int* b = buf[i];
        const int64_t* in = src + at;
#pragma omp parallel for schedule(static) if (m >= (int64_t(1) << 20))
        for (int64_t j = 0; j < m; ++j) b[j] = (int)in[j];
        e = hipMemcpyAsync(dst + at, b, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i], s);
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    for (int i = 0; i < 2; ++i) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (buf[i]) (void)hipHostFree(buf[i]);
    }
    return e;
}

int set_staging(int64_t min_bytes, int64_t chunk_bytes)
{
    if (min_bytes < 0 || chunk_bytes < 0 || (chunk_bytes > 0 && (chunk_bytes < 4096 || chunk_bytes % 4096)))
        return set_error(NIDO_ERR_INVALID, "min_bytes >= 0; chunk_bytes 0 or a multiple of 4096");
    g_min_bytes.store(min_bytes ? min_bytes : kDefaultMinBytes);
    g_chunk_bytes.store(chunk_bytes ? chunk_bytes : kDefaultChunkBytes);
    return NIDO_OK;
}

void rebase_offsets(const int64_t* off, int64_t n, int64_t* out)
{
    const int64_t base = off[0];
#pragma omp parallel for schedule(static) if (n >= (int64_t(1) << 22))
    for (int64_t i = 0; i <= n; ++i) out[i] = off[i] - base;
}

}  // namespace nido

extern "C" int nido_set_upload_staging(int64_t min_bytes, int64_t chunk_bytes)
{
    return nido::set_staging(min_bytes, chunk_bytes);
}
