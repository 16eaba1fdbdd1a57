// serve.cpp -- a continuously batching query server over device-resident
// libraries.
//
// Requests (1..n queries against one library, top-k each) are queued by
// bs_server_submit.  A worker thread keeps up to max_inflight batches on the
// GPU, each on its own stream with its own workspace.  Whenever a slot is free
// it takes the oldest queued request and merges every other queued request for
// the same library (up to a query cap) into one batch: one prep + Gram + top-k
// launch sequence for all of them.  A batch is admitted only while the device
// memory it needs fits in the server's HBM budget; requests that arrive while
// others run are admitted as slots and memory free up, so their work overlaps
// on the device.  Every pair is computed from the two conformations' own
// panels and constants, so a query's result does not depend on what it was
// batched with.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "library.hpp"

using namespace bs;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    size_t lib = 0;
    const double* q = nullptr;
    int n_q = 0, k = 0;
    uint32_t* idx = nullptr;
    float* rmsd = nullptr;
    Clock::time_point t_submit, t_done;
    bs_status status = BS_ERR_NOT_READY;
    std::string error;
};
using ReqPtr = std::shared_ptr<Request>;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    QueryWorkspace ws;
    std::vector<ReqPtr> reqs;
    std::vector<int> offset;  // first query of each request within the batch
    int n_q = 0, k = 0;
    size_t lib = 0;
    size_t reserved = 0;
    bool busy = false;
    // pinned staging
    double* h_q = nullptr;
    size_t h_q_cap = 0;  // doubles
    uint32_t* h_out = nullptr;
    size_t h_out_cap = 0;  // uint32 words
};

}  // namespace

struct bs_server {
    bs_context* ctx = nullptr;
    std::vector<bs_library*> libs;
    size_t budget = 0;
    size_t reserved = 0;
    int max_batch_queries = 512;
    int merge_queries = 512;  // queries merged into one batch (the size slots are reserved for)
    std::deque<Slot> slots;  // deque: slots are never moved
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> queue;
    std::unordered_map<uint64_t, ReqPtr> tickets;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;

    void run();
    bool admit_locked(Slot& s, std::vector<ReqPtr>& batch);
    void launch(Slot& s);
    void finish(Slot& s);
};

namespace {

void ensure_pinned(double*& p, size_t& cap, size_t need)
{
    if (need <= cap) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    const size_t want = need + need / 4;
    BS_HIP(hipHostMalloc((void**)&p, want * sizeof(double), hipHostMallocDefault));
    cap = want;
}

void ensure_pinned(uint32_t*& p, size_t& cap, size_t need)
{
    if (need <= cap) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    const size_t want = need + need / 4;
    BS_HIP(hipHostMalloc((void**)&p, want * sizeof(uint32_t), hipHostMallocDefault));
    cap = want;
}

}  // namespace

// Picks the batch for a free slot: the oldest request plus every queued
// request for the same library that still fits under the query cap.  Returns
// false when the memory it needs does not fit in the budget right now.
bool bs_server::admit_locked(Slot& s, std::vector<ReqPtr>& batch)
{
    batch.clear();
    const size_t lib = queue.front()->lib;
    int n_q = 0, k = 0;
    for (const auto& r : queue) {
        if (r->lib != lib) continue;
        if (!batch.empty() && n_q + r->n_q > merge_queries) break;
        batch.push_back(r);
        n_q += r->n_q;
        k = std::max(k, r->k);
    }
    const bs_library* L = libs[lib];
    const size_t need = query_workspace_bytes(n_q, L->P.n, L->P.L, k);
    if (budget != 0 && reserved + need > budget) {
        bool any_busy = false;
        for (auto& o : slots) any_busy |= o.busy;
        if (any_busy) return false;  // wait for memory to free up
        // nothing in flight and still too large: the oldest request cannot run
        ReqPtr r = queue.front();
        queue.pop_front();
        r->status = BS_ERR_OUT_OF_BUDGET;
        r->error = "request needs more device memory than the server budget";
        r->t_done = Clock::now();
        cv_done.notify_all();
        batch.clear();
        return true;
    }
    for (const auto& r : batch) queue.erase(std::find(queue.begin(), queue.end(), r));
    s.reqs = batch;
    s.offset.clear();
    int off = 0;
    for (const auto& r : batch) {
        s.offset.push_back(off);
        off += r->n_q;
    }
    s.n_q = n_q;
    s.k = k;
    s.lib = lib;
    s.reserved = need;
    reserved += need;
    s.busy = true;
    return true;
}

void bs_server::launch(Slot& s)
{
    bs_library* lib = libs[s.lib];
    const int L = lib->P.L;
    ensure_pinned(s.h_q, s.h_q_cap, (size_t)s.n_q * L * 3);
    for (size_t r = 0; r < s.reqs.size(); ++r)
        std::memcpy(s.h_q + (size_t)s.offset[r] * L * 3, s.reqs[r]->q,
                    (size_t)s.reqs[r]->n_q * L * 3 * sizeof(double));
    ensure_pinned(s.h_out, s.h_out_cap, (size_t)s.n_q * s.k * 2);
    enqueue_query(lib, s.ws, s.h_q, s.n_q, s.k, s.stream, nullptr);
    BS_HIP(hipMemcpyAsync(s.h_out, s.ws.topk.get(), (size_t)s.n_q * s.k * 8,
                          hipMemcpyDeviceToHost, s.stream));
    BS_HIP(hipEventRecord(s.done, s.stream));
}

void bs_server::finish(Slot& s)
{
    const uint32_t* idx = s.h_out;
    const float* val = reinterpret_cast<const float*>(s.h_out + (size_t)s.n_q * s.k);
    const auto now = Clock::now();
    for (size_t r = 0; r < s.reqs.size(); ++r) {
        Request& q = *s.reqs[r];
        for (int a = 0; a < q.n_q; ++a) {
            const size_t src = (size_t)(s.offset[r] + a) * s.k;
            std::memcpy(q.idx + (size_t)a * q.k, idx + src, q.k * sizeof(uint32_t));
            std::memcpy(q.rmsd + (size_t)a * q.k, val + src, q.k * sizeof(float));
        }
        q.t_done = now;
        q.status = BS_OK;
    }
    s.reqs.clear();
    reserved -= s.reserved;
    s.reserved = 0;
    s.busy = false;
}

void bs_server::run()
{
    (void)hipSetDevice(ctx->device);
    std::vector<ReqPtr> batch;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        // completions
        bool any_busy = false;
        for (auto& s : slots) {
            if (!s.busy) continue;
            const hipError_t e = hipEventQuery(s.done);
            if (e == hipSuccess) {
                finish(s);
                cv_done.notify_all();
            } else if (e != hipErrorNotReady) {
                for (auto& r : s.reqs) {
                    r->status = BS_ERR_HIP;
                    r->error = hipGetErrorString(e);
                    r->t_done = Clock::now();
                }
                s.reqs.clear();
                reserved -= s.reserved;
                s.reserved = 0;
                s.busy = false;
                cv_done.notify_all();
            } else {
                any_busy = true;
            }
        }
        // admissions
        for (auto& s : slots) {
            if (s.busy || queue.empty()) continue;
            if (!admit_locked(s, batch)) break;
            if (!s.busy) continue;  // the request was rejected
            lk.unlock();
            try {
                launch(s);
            } catch (const Error& err) {
                lk.lock();
                for (auto& r : s.reqs) {
                    r->status = err.code;
                    r->error = err.what();
                    r->t_done = Clock::now();
                }
                s.reqs.clear();
                reserved -= s.reserved;
                s.reserved = 0;
                s.busy = false;
                cv_done.notify_all();
                continue;
            }
            lk.lock();
            any_busy = true;
        }
        if (stop && queue.empty() && !any_busy) break;
        if (any_busy) {
            // poll the in-flight batches; new arrivals also wake us
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        } else {
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        }
    }
}

extern "C" {

bs_status bs_server_create(bs_context* ctx, bs_library* const* libs, size_t n_libs,
                           size_t hbm_budget_bytes, int max_inflight, bs_server** out)
{
    return guard([&] {
        BS_REQUIRE(ctx && libs && out, "null pointer argument");
        BS_REQUIRE(n_libs >= 1, "a server needs at least one library");
        BS_REQUIRE(max_inflight >= 1 && max_inflight <= 16, "max_inflight must be in [1, 16]");
        for (size_t i = 0; i < n_libs; ++i)
            BS_REQUIRE(libs[i] && libs[i]->ctx == ctx, "library %zu is null or from another context", i);
        *out = nullptr;
        BS_HIP(hipSetDevice(ctx->device));
        auto srv = std::make_unique<bs_server>();
        srv->ctx = ctx;
        srv->libs.assign(libs, libs + n_libs);
        srv->budget = hbm_budget_bytes;
        for (int i = 0; i < max_inflight; ++i) srv->slots.emplace_back();
        // Every slot is sized up front for a merged batch (up to 256 queries of
        // the longest library against the largest one, k up to 1024), and
        // requests are merged only up to that size, so that serving does not
        // allocate, and free, device memory between requests (a free
        // synchronises the whole device and stalls every batch in flight).
        int max_n = 0, max_L = 0;
        for (size_t i = 0; i < n_libs; ++i) {
            max_n = std::max(max_n, libs[i]->P.n);
            max_L = std::max(max_L, libs[i]->P.L);
        }
        const int pre_q = std::min(256, srv->max_batch_queries);
        const int pre_k = std::min(1024, max_n);
        for (auto& s : srv->slots) {
            BS_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            BS_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            s.ws.bind(&ctx->budget);
            s.ws.gsc.max_wg = ctx->gsc.max_wg;
            if (hbm_budget_bytes == 0 ||
                query_workspace_bytes(pre_q, max_n, max_L, pre_k) * srv->slots.size() <= hbm_budget_bytes) {
                s.ws.reserve(pre_q, max_n, max_L, pre_k);
                ensure_pinned(s.h_q, s.h_q_cap, (size_t)pre_q * max_L * 3);
                ensure_pinned(s.h_out, s.h_out_cap, (size_t)pre_q * pre_k * 2);
                srv->merge_queries = pre_q;
            }
        }
        bs_server* raw = srv.get();
        srv->worker = std::thread([raw] { raw->run(); });
        *out = srv.release();
    });
}

void bs_server_destroy(bs_server* srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->ctx->device);
    for (auto& s : srv->slots) {
        (void)hipStreamSynchronize(s.stream);
        s.ws.reset();
        if (s.h_q) (void)hipHostFree(s.h_q);
        if (s.h_out) (void)hipHostFree(s.h_out);
        (void)hipEventDestroy(s.done);
        (void)hipStreamDestroy(s.stream);
    }
    delete srv;
}

bs_status bs_server_submit(bs_server* srv, size_t lib_index, const double* q_xyz, size_t n_q,
                           size_t k, uint32_t* idx, float* rmsd, uint64_t* ticket)
{
    return guard([&] {
        BS_REQUIRE(srv && q_xyz && idx && rmsd && ticket, "null pointer argument");
        BS_REQUIRE(lib_index < srv->libs.size(), "library index %zu out of range", lib_index);
        BS_REQUIRE(n_q >= 1 && (int)n_q <= srv->max_batch_queries,
                   "n_q must be in [1, %d]", srv->max_batch_queries);
        const size_t n_lib = (size_t)srv->libs[lib_index]->P.n;
        BS_REQUIRE(k >= 1 && k <= 1024 && k <= n_lib, "k must be in [1, min(1024, library size)]");
        auto r = std::make_shared<Request>();
        r->lib = lib_index;
        r->q = q_xyz;
        r->n_q = (int)n_q;
        r->k = (int)k;
        r->idx = idx;
        r->rmsd = rmsd;
        r->t_submit = Clock::now();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            BS_REQUIRE(!srv->stop, "server is shutting down");
            r->ticket = srv->next_ticket++;
            srv->tickets[r->ticket] = r;
            srv->queue.push_back(r);
        }
        srv->cv_work.notify_all();
        *ticket = r->ticket;
    });
}

static ReqPtr find_ticket(bs_server* srv, uint64_t ticket)
{
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) fail(BS_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    return it->second;
}

bs_status bs_server_poll(bs_server* srv, uint64_t ticket)
{
    bs_status st = BS_OK;
    const bs_status g = guard([&] {
        BS_REQUIRE(srv != nullptr, "server is null");
        std::lock_guard<std::mutex> lk(srv->mu);
        ReqPtr r = find_ticket(srv, ticket);
        st = r->status;
        if (st != BS_OK && st != BS_ERR_NOT_READY) set_last_error(r->error);
    });
    return g != BS_OK ? g : st;
}

bs_status bs_server_wait(bs_server* srv, uint64_t ticket)
{
    bs_status st = BS_OK;
    const bs_status g = guard([&] {
        BS_REQUIRE(srv != nullptr, "server is null");
        std::unique_lock<std::mutex> lk(srv->mu);
        ReqPtr r = find_ticket(srv, ticket);
        srv->cv_done.wait(lk, [&] { return r->status != BS_ERR_NOT_READY; });
        st = r->status;
        if (st != BS_OK) set_last_error(r->error);
    });
    return g != BS_OK ? g : st;
}

bs_status bs_server_cancel(bs_server* srv, uint64_t ticket)
{
    return guard([&] {
        BS_REQUIRE(srv != nullptr, "server is null");
        std::lock_guard<std::mutex> lk(srv->mu);
        ReqPtr r = find_ticket(srv, ticket);
        auto it = std::find(srv->queue.begin(), srv->queue.end(), r);
        if (it == srv->queue.end())
            fail(BS_ERR_NOT_READY, "ticket %llu is already running or finished",
                 (unsigned long long)ticket);
        srv->queue.erase(it);
        r->status = BS_ERR_CANCELLED;
        r->error = "cancelled";
        r->t_done = Clock::now();
        srv->cv_done.notify_all();
    });
}

bs_status bs_server_latency_ms(bs_server* srv, uint64_t ticket, double* ms)
{
    return guard([&] {
        BS_REQUIRE(srv && ms, "null pointer argument");
        std::lock_guard<std::mutex> lk(srv->mu);
        ReqPtr r = find_ticket(srv, ticket);
        if (r->status == BS_ERR_NOT_READY) fail(BS_ERR_NOT_READY, "ticket still running");
        *ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    });
}

}  // extern "C"
