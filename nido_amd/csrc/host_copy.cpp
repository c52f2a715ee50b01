// host_copy.cpp -- large host -> device copies through pinned staging buffers.
#include "host_copy.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/nido_amd.h"
#include "errors.h"

namespace nido {

namespace {
// Copies of at least kMinBytes go through two pinned buffers of kChunkBytes
// each (a pageable copy of 388 MB crossed at 26 GB/s, the staged copies at
// ~51 GB/s; below a few chunks there is little to overlap); both can be
// lowered so that tests reach the pinned path with small inputs.
constexpr int64_t kDefaultMinBytes = int64_t(256) << 20;
constexpr int64_t kDefaultChunkBytes = int64_t(64) << 20;
std::atomic<int64_t> g_min_bytes{kDefaultMinBytes};
std::atomic<int64_t> g_chunk_bytes{kDefaultChunkBytes};

// A pair of pinned staging buffers with their events, taken from a
// process-wide pool and given back when the copy is done: pinning and
// unpinning 128 MiB costs milliseconds per copy.  The pool keeps at most
// kPoolPairs idle pairs (concurrent copies beyond that pin their own and
// release them); the pairs live until the process ends.
constexpr size_t kPoolPairs = 2;
struct Pair {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    size_t bytes = 0;
    int device = -1;
};
std::mutex g_pool_mu;
std::vector<Pair>* g_pool = new std::vector<Pair>();   // never destroyed: no HIP calls at exit

void free_pair(Pair& p)
{
    for (int i = 0; i < 2; ++i) {
        if (p.done[i]) (void)hipEventDestroy(p.done[i]);
        if (p.buf[i]) (void)hipHostFree(p.buf[i]);
        p.done[i] = nullptr;
        p.buf[i] = nullptr;
    }
}

class Staging {
  public:
    explicit Staging(size_t bytes)
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        {
            std::lock_guard<std::mutex> lk(g_pool_mu);
            for (size_t i = 0; i < g_pool->size(); ++i)
                if ((*g_pool)[i].bytes == bytes && (*g_pool)[i].device == dev) {
                    p_ = (*g_pool)[i];
                    g_pool->erase(g_pool->begin() + (ptrdiff_t)i);
                    return;
                }
        }
        p_.bytes = bytes;
        p_.device = dev;
        for (int i = 0; i < 2 && err_ == hipSuccess; ++i) {
            err_ = hipHostMalloc(reinterpret_cast<void**>(&p_.buf[i]), bytes);
            if (err_ == hipSuccess) err_ = hipEventCreateWithFlags(&p_.done[i], hipEventDisableTiming);
        }
    }
    ~Staging()
    {
        if (err_ == hipSuccess && reuse_) {
            std::lock_guard<std::mutex> lk(g_pool_mu);
            if (g_pool->size() < kPoolPairs) {
                g_pool->push_back(p_);
                return;
            }
        }
        free_pair(p_);
    }
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    hipError_t status() const { return err_; }
    // Marks a copy that ended without error: only then does the pair go
    // back to the pool.
    hipError_t finish(hipError_t e)
    {
        reuse_ = e == hipSuccess;
        return e;
    }
    // The caller's stream has finished every copy out of or into the
    // buffers before the pair goes back (each copy ends with a stream
    // synchronisation), so the events may be recorded again at once.
    char* buf(int i) const { return p_.buf[i]; }
    hipEvent_t done(int i) const { return p_.done[i]; }

  private:
    Pair p_;
    hipError_t err_ = hipSuccess;
    bool reuse_ = false;
};

void atomic_min(std::atomic<int64_t>& a, int64_t x)
{
    int64_t cur = a.load();
    while (x < cur && !a.compare_exchange_weak(cur, x)) {
    }
}

// Calls fn(g, j0, j1) for the pieces of arcs [j0, j1) inside graph g, in
// order (graphs without arcs skipped).
template <typename Fn>
void by_graph(const ArcCheck& chk, int64_t j0, int64_t j1, Fn&& fn)
{
    if (j0 >= j1) return;
    int g = (int)(std::upper_bound(chk.gstart, chk.gstart + chk.G + 1, j0) - chk.gstart) - 1;
    for (int64_t j = j0; j < j1; ++g) {
        const int64_t e = std::min(j1, chk.gstart[g + 1]);
        if (e > j) fn(g, j, e);
        j = std::max(j, e);
    }
}

// Arcs [at, at + m) of src narrowed into b, their columns checked.
void narrow_checked(int* b, const int64_t* src, int64_t at, int64_t m, ArcCheck& chk)
{
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
        by_graph(chk, at + m * tid / nt, at + m * (tid + 1) / nt, [&](int g, int64_t j0, int64_t j1) {
            const int64_t lo = chk.vptr[g], hi = chk.vptr[g + 1];
            bool out = false;
            for (int64_t j = j0; j < j1; ++j) {
                const int64_t c = src[j];
                b[j - at] = (int)c;
                out |= c < lo || c >= hi;
            }
            if (out)
                for (int64_t j = j0; j < j1; ++j)
                    if (src[j] < lo || src[j] >= hi) {
                        atomic_min(chk.bad, j);
                        break;
                    }
        });
    }
}

// Arcs [at, at + m) of src as 16-bit ids local to their graph (column minus
// the graph's first vertex; every graph has at most 2^16 vertices), their
// columns checked.
void narrow_local_checked(uint16_t* b, const int64_t* src, int64_t at, int64_t m, ArcCheck& chk)
{
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
        by_graph(chk, at + m * tid / nt, at + m * (tid + 1) / nt, [&](int g, int64_t j0, int64_t j1) {
            const int64_t lo = chk.vptr[g], hi = chk.vptr[g + 1];
            bool out = false;
            for (int64_t j = j0; j < j1; ++j) {
                const int64_t c = src[j];
                b[j - at] = (uint16_t)(c - lo);
                out |= c < lo || c >= hi;
            }
            if (out)
                for (int64_t j = j0; j < j1; ++j)
                    if (src[j] < lo || src[j] >= hi) {
                        atomic_min(chk.bad, j);
                        break;
                    }
        });
    }
}

// Weights [at, at + m) of src copied into b, checked, |w| summed per graph.
void weights_checked(double* b, const double* src, int64_t at, int64_t m, ArcCheck& chk)
{
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
        by_graph(chk, at + m * tid / nt, at + m * (tid + 1) / nt, [&](int g, int64_t j0, int64_t j1) {
            bool bad = false, neg = false;
            double sa = 0.0;
            for (int64_t j = j0; j < j1; ++j) {
                const double x = src[j];
                b[j - at] = x;
                const double ax = std::fabs(x);
                bad |= !(ax < 0x1p62);   // NaN, infinite or too large
                neg |= x < 0.0;
                sa += ax < 0x1p62 ? ax : 0.0;
            }
            if (bad)
                for (int64_t j = j0; j < j1; ++j)
                    if (!(std::fabs(src[j]) < 0x1p62)) {
                        atomic_min(chk.bad, j);
                        break;
                    }
            if (neg) chk.neg.store(true);
            std::lock_guard<std::mutex> lk(chk.mu);
            chk.wabs[(size_t)g] += sa;
        });
    }
}
}  // namespace

bool staged_path(size_t bytes)
{
    return bytes >= (size_t)g_min_bytes.load();
}

hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const size_t kChunk = (size_t)g_chunk_bytes.load();
    if (bytes < (size_t)g_min_bytes.load()) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    Staging st(kChunk);
    hipError_t e = st.status();
    char* const buf[2] = {st.buf(0), st.buf(1)};
    const hipEvent_t done[2] = {st.done(0), st.done(1)};
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
    return st.finish(e);
}

hipError_t staged_d2h(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const size_t kChunk = (size_t)g_chunk_bytes.load();
    if (bytes < (size_t)g_min_bytes.load()) {
        hipError_t e = bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    Staging st(kChunk);
    hipError_t e = st.status();
    char* const buf[2] = {st.buf(0), st.buf(1)};
    const hipEvent_t done[2] = {st.done(0), st.done(1)};
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t k) -> hipError_t {   // chunk k into buffer k & 1
        const size_t at = k * kChunk, n = std::min(kChunk, bytes - at);
        hipError_t r = hipMemcpyAsync(buf[k & 1], in + at, n, hipMemcpyDeviceToHost, s);
        return r == hipSuccess ? hipEventRecord(done[k & 1], s) : r;
    };
    for (size_t k = 0; k < std::min<size_t>(2, nchunks) && e == hipSuccess; ++k) e = issue(k);
    for (size_t k = 0; k < nchunks && e == hipSuccess; ++k) {
        e = hipEventSynchronize(done[k & 1]);   // chunk k has landed in its buffer
        if (e != hipSuccess) break;
        const size_t at = k * kChunk, n = std::min(kChunk, bytes - at);
        constexpr size_t kPiece = size_t(1) << 20;
        const int64_t pieces = (int64_t)((n + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < pieces; ++p) {
            const size_t o = (size_t)p * kPiece;
            std::memcpy(out + at + o, buf[k & 1] + o, std::min(kPiece, n - o));
        }
        if (k + 2 < nchunks) e = issue(k + 2);   // the buffer is free again
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    return st.finish(e);
}

hipError_t staged_h2d_ids(int* dst, const int64_t* src, int64_t n, hipStream_t s, ArcCheck* chk)
{
    const int64_t kChunk = g_chunk_bytes.load() / (int64_t)sizeof(int);   // ids per pinned buffer
    if (n * (int64_t)sizeof(int) < g_min_bytes.load()) {   // one pageable copy, no pinned allocation
        std::vector<int> t((size_t)n);
        if (chk) {
            narrow_checked(t.data(), src, 0, n, *chk);
        } else {
            // threads only for graphs larger than a serving request, whose
            // requests already run side by side
#pragma omp parallel for schedule(static) if (n >= (int64_t(1) << 26))
            for (int64_t j = 0; j < n; ++j) t[(size_t)j] = (int)src[j];
        }
        hipError_t e = hipMemcpyAsync(dst, t.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    Staging st(sizeof(int) * (size_t)kChunk);
    hipError_t e = st.status();
    int* const buf[2] = {reinterpret_cast<int*>(st.buf(0)), reinterpret_cast<int*>(st.buf(1))};
    const hipEvent_t done[2] = {st.done(0), st.done(1)};
    for (int64_t at = 0, i = 0; e == hipSuccess && at < n; at += kChunk, i ^= 1) {
        const int64_t m = std::min(kChunk, n - at);
        e = hipEventSynchronize(done[i]);      // the buffer's previous copy has landed
        if (e != hipSuccess) break;
        int* b = buf[i];
        const int64_t* in = src + at;
        if (chk) {
            narrow_checked(b, src, at, m, *chk);
        } else {
#pragma omp parallel for schedule(static) if (m >= (int64_t(1) << 20))
            for (int64_t j = 0; j < m; ++j) b[j] = (int)in[j];
        }
        e = hipMemcpyAsync(dst + at, b, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i], s);
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    return st.finish(e);
}

hipError_t staged_h2d_local_ids(uint16_t* dst, const int64_t* src, int64_t n, hipStream_t s, ArcCheck& chk)
{
    const int64_t kChunk = g_chunk_bytes.load() / (int64_t)sizeof(uint16_t);
    Staging st(sizeof(uint16_t) * (size_t)kChunk);
    hipError_t e = st.status();
    uint16_t* const buf[2] = {reinterpret_cast<uint16_t*>(st.buf(0)), reinterpret_cast<uint16_t*>(st.buf(1))};
    const hipEvent_t done[2] = {st.done(0), st.done(1)};
    for (int64_t at = 0, i = 0; e == hipSuccess && at < n; at += kChunk, i ^= 1) {
        const int64_t m = std::min(kChunk, n - at);
        e = hipEventSynchronize(done[i]);      // the buffer's previous copy has landed
        if (e != hipSuccess) break;
        narrow_local_checked(buf[i], src, at, m, chk);
        e = hipMemcpyAsync(dst + at, buf[i], sizeof(uint16_t) * (size_t)m, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i], s);
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    return st.finish(e);
}

hipError_t staged_h2d_weights(double* dst, const double* src, int64_t n, hipStream_t s, ArcCheck& chk)
{
    const int64_t kChunk = g_chunk_bytes.load() / (int64_t)sizeof(double);
    chk.wabs.assign((size_t)chk.G, 0.0);
    Staging st(sizeof(double) * (size_t)kChunk);
    hipError_t e = st.status();
    double* const buf[2] = {reinterpret_cast<double*>(st.buf(0)), reinterpret_cast<double*>(st.buf(1))};
    const hipEvent_t done[2] = {st.done(0), st.done(1)};
    for (int64_t at = 0, i = 0; e == hipSuccess && at < n; at += kChunk, i ^= 1) {
        const int64_t m = std::min(kChunk, n - at);
        e = hipEventSynchronize(done[i]);      // the buffer's previous copy has landed
        if (e != hipSuccess) break;
        weights_checked(buf[i], src, at, m, chk);
        e = hipMemcpyAsync(dst + at, buf[i], sizeof(double) * (size_t)m, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[i], s);
    }
    const hipError_t w = hipStreamSynchronize(s);
    if (e == hipSuccess) e = w;
    return st.finish(e);
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
