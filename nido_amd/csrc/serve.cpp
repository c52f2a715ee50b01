// serve.cpp -- a Louvain request loop over the C ABI: continuous batching on
// one GPU.
//
// Requests (one graph each) queue in arrival order.  A request is admitted
// when one of `slots` workers is free and its estimated HBM footprint fits
// in what is left of the budget; each worker owns a HIP stream and runs the
// whole LouvainGPU::run (louvain_gpu.cpp:6-139) for its request there, with
// every phase's state resident on the device, so admitted requests overlap
// on the GPU.  A finished request frees its slot and its share of the
// budget, and the next ones are admitted at once.  With `group` > 1 a worker
// that takes a small request (up to kGroupArcs arcs) also takes the small
// requests queued right behind it, up to `group` of them and while their
// footprints fit, and runs them together in one batched handle
// (nido_batch_create / nido_batch_run): one launch sequence serves them all,
// where one request would leave most of the GPU idle.  A request's result
// does not depend on what runs beside it, in a group or not.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nido_amd.h"
#include "errors.h"
#include "host_check.h"

using nido::set_error;

namespace {

constexpr int64_t kGroupArcs = int64_t(1) << 23;        // requests up to this size are grouped
constexpr int64_t kGroupTotalArcs = int64_t(1) << 26;   // arcs of one group, at most
constexpr int64_t kMaxVertices = (int64_t(1) << 31) - 1;  // vertices of one handle, below this

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    int64_t nv = 0;
    std::vector<int64_t> off, col;
    std::vector<double> w;
    const int64_t* boff = nullptr;   // borrowed input (NIDO_SERVE_BORROW)
    const int64_t* bcol = nullptr;
    const double* bw = nullptr;
    int64_t ne = 0;
    int64_t footprint = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    int state = 0;            // 0 queued, 1 running, 2 done
    int rc = NIDO_OK;
    std::string err;
    nido_run_stats st{};
    std::vector<int64_t> clustering;
};

}  // namespace

struct nido_server {
    nido_serve_config cfg{};
    int device = 0;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> reqs;
    int64_t next_id = 1;
    int64_t budget_free = 0;
    int64_t running = 0, completed = 0, peak_running = 0;
    bool stop = false;

    void worker(int slot);
    void run_one(Request* r, hipStream_t st);
    void run_group(const std::vector<Request*>& grp, hipStream_t st);
};

namespace {

const int64_t* req_off(const Request* r) { return r->boff ? r->boff : r->off.data(); }
const int64_t* req_col(const Request* r) { return r->boff ? r->bcol : r->col.data(); }
const double* req_w(const Request* r) { return r->boff ? r->bw : r->w.data(); }

}  // namespace

void nido_server::run_one(Request* r, hipStream_t st)
{
    nido_graph* g = nullptr;
    int rc = nido_graph_create(r->nv, r->ne, req_off(r), req_col(r), req_w(r), cfg.nbatches, st, &g);
    if (rc == NIDO_OK) rc = nido_louvain_run(g, cfg.max_loops, cfg.tol, cfg.tol_phase, 0, 0, &r->st);
    if (rc == NIDO_OK) {
        r->clustering.resize((size_t)r->nv);
        rc = r->nv ? nido_get_clustering(g, r->clustering.data()) : NIDO_OK;
    }
    if (rc != NIDO_OK) r->err = nido_last_error();
    if (g) nido_graph_destroy(g);
    r->rc = rc;
}

// The group's graphs side by side in one CSR (graph i owns vertices
// [vptr[i], vptr[i+1])), run to completion together; each request then
// takes its own graph's statistics and clustering.
void nido_server::run_group(const std::vector<Request*>& grp, hipStream_t st)
{
    const int G = (int)grp.size();
    std::vector<int64_t> vptr((size_t)G + 1, 0);
    int64_t ne = 0;
    for (int i = 0; i < G; ++i) {
        vptr[(size_t)i + 1] = vptr[(size_t)i] + grp[(size_t)i]->nv;
        ne += grp[(size_t)i]->ne;
    }
    std::vector<int64_t> off((size_t)vptr[(size_t)G] + 1), col((size_t)ne);
    std::vector<double> w((size_t)ne);
    int64_t e = 0;
    for (int i = 0; i < G; ++i) {
        const Request* r = grp[(size_t)i];
        const int64_t* ro = req_off(r);
        const int64_t* rcol = req_col(r);
        const double* rw = req_w(r);
        const int64_t v0 = vptr[(size_t)i];
        for (int64_t v = 0; v < r->nv; ++v) off[(size_t)(v0 + v)] = e + (ro[v] - ro[0]);
        for (int64_t j = 0; j < r->ne; ++j) {
            col[(size_t)(e + j)] = rcol[ro[0] + j] + v0;
            w[(size_t)(e + j)] = rw[ro[0] + j];
        }
        e += r->ne;
    }
    off[(size_t)vptr[(size_t)G]] = e;
    nido_graph* g = nullptr;
    nido_run_stats all{};
    int rc = nido_batch_create(G, vptr.data(), off.data(), col.data(), w.data(), cfg.nbatches, st, &g);
    if (rc == NIDO_ERR_INVALID || rc == NIDO_ERR_CAPACITY || rc == NIDO_ERR_NOMEM) {
        // every request was checked on its own at submission, so this is the
        // group's doing (its size or memory): run the members one at a time,
        // each with the result it gets alone
        std::vector<int64_t>().swap(off);
        std::vector<int64_t>().swap(col);
        std::vector<double>().swap(w);
        for (Request* r : grp) run_one(r, st);
        return;
    }
    if (rc == NIDO_OK) rc = nido_batch_run(g, cfg.max_loops, cfg.tol, cfg.tol_phase, &all);
    const std::string err = rc != NIDO_OK ? std::string(nido_last_error()) : std::string();
    for (int i = 0; i < G; ++i) {
        Request* r = grp[(size_t)i];
        int ri = rc;
        if (ri == NIDO_OK) ri = nido_batch_graph_stats(g, i, &r->st);
        if (ri == NIDO_OK) {
            r->clustering.resize((size_t)r->nv);
            ri = r->nv ? nido_batch_clustering(g, i, r->clustering.data()) : NIDO_OK;
        }
        r->rc = ri;
        if (ri != NIDO_OK) r->err = rc != NIDO_OK ? err : std::string(nido_last_error());
    }
    if (g) nido_graph_destroy(g);
}

void nido_server::worker(int slot)
{
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<Request*> grp;
        int64_t fp = 0;
        {
            std::unique_lock<std::mutex> lock(mu);
            // a request fits once enough running ones finish (its footprint
            // is checked against the whole budget at submission); on stop
            // the queue is still drained
            cv_work.wait(lock, [&] {
                return (!queue.empty() && queue.front()->footprint <= budget_free) || (stop && queue.empty());
            });
            if (queue.empty()) return;
            // the head request, then (if it is small) the small ones right
            // behind it, in arrival order, while they fit
            int64_t arcs = 0, verts = 0;
            do {
                Request* r = queue.front();
                if (!grp.empty() && (r->ne > kGroupArcs || fp + r->footprint > budget_free ||
                                     arcs + r->ne > kGroupTotalArcs || verts + r->nv >= kMaxVertices))
                    break;
                queue.pop_front();
                grp.push_back(r);
                fp += r->footprint;
                arcs += r->ne;
                verts += r->nv;
                r->state = 1;
            } while (grp.front()->ne <= kGroupArcs && (int)grp.size() < cfg.group && !queue.empty());
            budget_free -= fp;
            running += (int64_t)grp.size();
            if (running > peak_running) peak_running = running;
        }
        const double t0 = now_s();
        for (Request* r : grp) r->t_start = t0;
        try {
            if (grp.size() == 1) run_one(grp[0], streams[(size_t)slot]);
            else run_group(grp, streams[(size_t)slot]);
        } catch (const std::exception& e) {
            // a host copy of the group or a clustering did not fit (or any
            // other failure a worker thread must not let escape): the
            // requests fail, and their slot and budget share are freed below
            for (Request* r : grp) {
                r->rc = NIDO_ERR_NOMEM;
                r->err = std::string("out of host memory: ") + e.what();
                std::vector<int64_t>().swap(r->clustering);
            }
        }
        for (Request* r : grp) {   // the input is no longer needed
            std::vector<int64_t>().swap(r->off);
            std::vector<int64_t>().swap(r->col);
            std::vector<double>().swap(r->w);
        }
        {
            std::lock_guard<std::mutex> lock(mu);
            const double t1 = now_s();
            for (Request* r : grp) {
                r->t_end = t1;
                r->state = 2;
            }
            budget_free += fp;
            running -= (int64_t)grp.size();
            completed += (int64_t)grp.size();
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int64_t nido_serve_footprint(int64_t nv, int64_t ne)
{
    // resident CSR + 32-bit mirrors + per-vertex state + long-row and
    // coarsening work space (bounded by a few bytes per arc)
    return 160 * (nv + 1) + 96 * ne + (64ll << 20);
}

int nido_serve_create(const nido_serve_config* cfg, nido_server** out)
{
    if (!cfg || !out) return set_error(NIDO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 64) return set_error(NIDO_ERR_INVALID, "slots must be in [1, 64]");
    if (cfg->nbatches < 1) return set_error(NIDO_ERR_INVALID, "nbatches must be >= 1");
    if (cfg->hbm_budget <= 0) return set_error(NIDO_ERR_INVALID, "hbm_budget must be > 0");
    if (cfg->group < 0 || cfg->group > 4096) return set_error(NIDO_ERR_INVALID, "group must be in [0, 4096]");
    nido_server* s = new (std::nothrow) nido_server;
    if (!s) return set_error(NIDO_ERR_NOMEM, "out of host memory");
    s->cfg = *cfg;
    s->budget_free = cfg->hbm_budget;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return set_error(NIDO_ERR_HIP, "no HIP device");
    }
    for (int i = 0; i < cfg->slots; ++i) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (auto x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return set_error(NIDO_ERR_HIP, "stream creation failed");
        }
        s->streams.push_back(st);
    }
    for (int i = 0; i < cfg->slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return NIDO_OK;
}

int nido_serve_submit(nido_server* s, int64_t nv, const int64_t* off, const int64_t* col, const double* w,
                      int flags, int64_t* id)
{
    if (!s || !off || !id || nv < 0) return set_error(NIDO_ERR_INVALID, "invalid argument");
    if (nv >= kMaxVertices) return set_error(NIDO_ERR_CAPACITY, "nv must be below 2^31-1");
    const int64_t ne = off[nv] - off[0];
    if (ne < 0 || (ne > 0 && (!col || !w))) return set_error(NIDO_ERR_INVALID, "invalid CSR");
    {
        // the whole check of nido_graph_create, on this request alone: a
        // malformed request is refused here, so it never fails a group
        const int64_t whole[2] = {0, nv};
        const int rc = nido::check_csr(nv, off, col, w, 1, whole, nullptr);
        if (rc != NIDO_OK) return rc;
    }
    Request* r = new (std::nothrow) Request;
    if (!r) return set_error(NIDO_ERR_NOMEM, "out of host memory");
    r->nv = nv;
    r->footprint = nido_serve_footprint(nv, ne);
    if (r->footprint > s->cfg.hbm_budget) {
        delete r;
        return set_error(NIDO_ERR_CAPACITY, "request needs ~%lld bytes, more than the whole budget",
                         (long long)nido_serve_footprint(nv, ne));
    }
    r->ne = ne;
    if (flags & NIDO_SERVE_BORROW) {
        r->boff = off;
        r->bcol = col;
        r->bw = w;
    } else {
        r->off.resize((size_t)nv + 1);
        for (int64_t i = 0; i <= nv; ++i) r->off[(size_t)i] = off[i] - off[0];
        r->col.assign(col + off[0], col + off[0] + ne);
        r->w.assign(w + off[0], w + off[0] + ne);
    }
    {
        std::lock_guard<std::mutex> lock(s->mu);
        r->id = s->next_id++;
        r->t_submit = now_s();
        s->reqs[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    *id = r->id;
    return NIDO_OK;
}

int nido_serve_wait(nido_server* s, int64_t id, nido_run_stats* st, double* latency_s)
{
    if (!s) return set_error(NIDO_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(NIDO_ERR_INVALID, "unknown request %lld", (long long)id);
    Request* r = it->second;
    s->cv_done.wait(lock, [&] { return r->state == 2; });
    if (r->rc != NIDO_OK) return set_error(r->rc, "request %lld failed: %s", (long long)id, r->err.c_str());
    if (st) *st = r->st;
    if (latency_s) *latency_s = r->t_end - r->t_submit;
    return NIDO_OK;
}

int nido_serve_clustering(nido_server* s, int64_t id, int64_t* comm)
{
    if (!s || !comm) return set_error(NIDO_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(NIDO_ERR_INVALID, "unknown request %lld", (long long)id);
    if (it->second->state != 2) return set_error(NIDO_ERR_STATE, "request %lld has not finished", (long long)id);
    std::memcpy(comm, it->second->clustering.data(), sizeof(int64_t) * it->second->clustering.size());
    return NIDO_OK;
}

int nido_serve_release(nido_server* s, int64_t id)
{
    if (!s) return set_error(NIDO_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(NIDO_ERR_INVALID, "unknown request %lld", (long long)id);
    if (it->second->state != 2) return set_error(NIDO_ERR_STATE, "request %lld has not finished", (long long)id);
    delete it->second;
    s->reqs.erase(it);
    return NIDO_OK;
}

int nido_serve_status(nido_server* s, int64_t* queued, int64_t* running, int64_t* completed, int64_t* peak_running)
{
    if (!s) return set_error(NIDO_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    if (queued) *queued = (int64_t)s->queue.size();
    if (running) *running = s->running;
    if (completed) *completed = s->completed;
    if (peak_running) *peak_running = s->peak_running;
    return NIDO_OK;
}

int nido_serve_destroy(nido_server* s)
{
    if (!s) return NIDO_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;   // queued requests still run; workers exit once the queue is empty
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto x : s->streams) (void)hipStreamDestroy(x);
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
    return NIDO_OK;
}

}  // extern "C"
