// serve.cpp -- a continuously batching query server over device-resident
// libraries.
//
// Requests (1..n queries against one library, top-k each) are queued by
// bs_server_submit, which copies the queries.  A worker thread keeps up to
// max_inflight batches in flight, one per slot, each slot with its own
// executor (HIP stream, workspace, pinned staging).  Whenever a slot is free
// it takes the oldest queued request and merges every other queued request for
// the same library (up to a query cap) into one batch: one prep + Gram + top-k
// launch sequence for all of them.  A batch is admitted only while the device
// memory it needs fits in the server's HBM budget; requests that arrive while
// others run are admitted as slots and memory free up, so their work overlaps
// on the device.  Every pair is computed from the two conformations' own
// panels and constants, so a query's result does not depend on what it was
// batched with.
//
// The queue, admission, merging and tickets are independent of the device
// (serve.hpp); HipExecutor below is the product's executor.
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "library.hpp"

using namespace bs;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    size_t lib = 0;
    std::vector<double> q;  // the queries, copied at submit
    int n_q = 0, k = 0;
    uint32_t* idx = nullptr;  // caller's result buffers; null once the ticket is released early
    float* rmsd = nullptr;
    bool released = false;    // the caller gave the ticket up: drop the record when it ends
    Clock::time_point t_submit, t_done;
    bs_status status = BS_ERR_NOT_READY;
    std::string error;
};
using ReqPtr = std::shared_ptr<Request>;

struct Slot {
    std::unique_ptr<BatchExecutor> exec;
    std::vector<ReqPtr> reqs;
    std::vector<int> offset;  // first query of each request within the batch
    int n_q = 0, k = 0;
    size_t lib = 0;
    size_t reserved = 0;
    bool busy = false;
};

}  // namespace

struct bs_server {
    ServerSetup cfg;
    size_t reserved = 0;
    std::deque<Slot> slots;  // deque: slots are never moved
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> queue;
    std::unordered_map<uint64_t, ReqPtr> tickets;
    // query copies of finished requests, reused by later submits: a fresh
    // buffer of a few hundred KiB would be new pages, faulted in on the copy
    std::vector<std::vector<double>> q_pool;
    static constexpr size_t kQueryPool = 32;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;

    void run();
    // A request reached its final status: a released ticket's record goes.
    void ended_locked(const ReqPtr& r)
    {
        if (q_pool.size() < kQueryPool && r->q.capacity() > 0) q_pool.push_back(std::move(r->q));
        r->q = std::vector<double>();
        if (r->released) tickets.erase(r->ticket);
    }
    void fail_slot_locked(Slot& s, bs_status code, const char* msg);
    bool admit_locked(Slot& s, std::vector<ReqPtr>& batch);
    void launch(Slot& s);
    void finish_locked(Slot& s);
};

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
        if (!batch.empty() && n_q + r->n_q > cfg.merge_queries) break;
        // every request's k was checked alone at submit, and the check is
        // monotone in k, so the merged k = max k stays in range
        if (!cfg.k_supported(lib, std::max(k, r->k))) continue;
        batch.push_back(r);
        n_q += r->n_q;
        k = std::max(k, r->k);
    }
    const size_t need = cfg.batch_bytes(lib, n_q, k);
    if (cfg.budget != 0 && reserved + need > cfg.budget) {
        bool any_busy = false;
        for (auto& o : slots) any_busy |= o.busy;
        if (any_busy) return false;  // wait for memory to free up
        // nothing in flight and still too large: the oldest request cannot run
        ReqPtr r = queue.front();
        queue.pop_front();
        r->status = BS_ERR_OUT_OF_BUDGET;
        r->error = "request needs more device memory than the server budget";
        r->t_done = Clock::now();
        ended_locked(r);
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

// Stages the batch's queries and starts it (worker thread, without the lock:
// while a slot is busy its request list is touched by this thread only, and
// a request's query copy never changes after submit).
void bs_server::launch(Slot& s)
{
    const size_t per_q = (size_t)cfg.libs[s.lib].L * 3;
    double* h = s.exec->staging((size_t)s.n_q * per_q);
    for (size_t r = 0; r < s.reqs.size(); ++r)
        std::memcpy(h + (size_t)s.offset[r] * per_q, s.reqs[r]->q.data(),
                    (size_t)s.reqs[r]->n_q * per_q * sizeof(double));
    s.exec->launch(s.lib, s.n_q, s.k);
}

void bs_server::finish_locked(Slot& s)
{
    const uint32_t* idx = s.exec->results();
    const float* val = reinterpret_cast<const float*>(idx + (size_t)s.n_q * s.k);
    const auto now = Clock::now();
    for (size_t r = 0; r < s.reqs.size(); ++r) {
        Request& q = *s.reqs[r];
        if (q.idx) {  // a released ticket's buffers may be gone
            for (int a = 0; a < q.n_q; ++a) {
                const size_t src = (size_t)(s.offset[r] + a) * s.k;
                std::memcpy(q.idx + (size_t)a * q.k, idx + src, q.k * sizeof(uint32_t));
                std::memcpy(q.rmsd + (size_t)a * q.k, val + src, q.k * sizeof(float));
            }
        }
        q.t_done = now;
        q.status = BS_OK;
        ended_locked(s.reqs[r]);
    }
    s.reqs.clear();
    reserved -= s.reserved;
    s.reserved = 0;
    s.busy = false;
    cv_done.notify_all();
}

void bs_server::fail_slot_locked(Slot& s, bs_status code, const char* msg)
{
    for (auto& r : s.reqs) {
        r->status = code;
        r->error = msg;
        r->t_done = Clock::now();
        ended_locked(r);
    }
    s.reqs.clear();
    reserved -= s.reserved;
    s.reserved = 0;
    s.busy = false;
    cv_done.notify_all();
}

void bs_server::run()
{
    for (auto& s : slots) s.exec->bind_thread();
    std::vector<ReqPtr> batch;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        // completions: a busy slot's executor is touched by this thread only
        bool any_busy = false;
        for (auto& s : slots) {
            if (!s.busy) continue;
            bool fin = false;
            try {
                fin = s.exec->done();
            } catch (const Error& err) {
                fail_slot_locked(s, err.code, err.what());
                continue;
            }
            if (fin)
                finish_locked(s);
            else
                any_busy = true;
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
                fail_slot_locked(s, err.code, err.what());
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

namespace bs {

bs_server* make_server(ServerSetup setup)
{
    BS_REQUIRE(!setup.libs.empty() && !setup.executors.empty(), "a server needs libraries and slots");
    BS_REQUIRE(setup.batch_bytes && setup.k_supported, "a server needs its sizing functions");
    auto srv = std::make_unique<bs_server>();
    for (auto& e : setup.executors) {
        srv->slots.emplace_back();
        srv->slots.back().exec = std::move(e);
    }
    setup.executors.clear();
    srv->cfg = std::move(setup);
    bs_server* raw = srv.get();
    srv->worker = std::thread([raw] { raw->run(); });
    return srv.release();
}

}  // namespace bs

namespace {

// The product's executor: pinned staging, prep + Gram + FP64 re-evaluation +
// top-k of one merged batch on the slot's own stream, results copied back to
// pinned memory, completion by event.
class HipExecutor final : public BatchExecutor {
public:
    HipExecutor(bs_context* ctx, std::vector<bs_library*> libs) : ctx_(ctx), libs_(std::move(libs))
    {
        BS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        BS_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
        ws_.bind(&ctx->budget);
        ws_.gsc.max_wg = ctx->gsc.max_wg;
    }
    ~HipExecutor() override
    {
        (void)hipSetDevice(ctx_->device);
        if (stream_) (void)hipStreamSynchronize(stream_);
        ws_.reset();
        if (h_q_) (void)hipHostFree(h_q_);
        if (h_out_) (void)hipHostFree(h_out_);
        if (done_) (void)hipEventDestroy(done_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    // Sizes everything up front for batches of up to n_q queries of L atoms
    // against n_lib members with top-k, so that such batches never allocate.
    void reserve(int n_q, int n_lib, int L, int k)
    {
        ws_.reserve(n_q, n_lib, L, k);
        ensure_pinned(h_q_, h_q_cap_, (size_t)n_q * L * 3);
        ensure_pinned(h_out_, h_out_cap_, (size_t)n_q * k * 2);
    }
    void bind_thread() override { (void)hipSetDevice(ctx_->device); }
    double* staging(size_t doubles) override
    {
        ensure_pinned(h_q_, h_q_cap_, doubles);
        return h_q_;
    }
    void launch(size_t lib, int n_q, int k) override
    {
        ensure_pinned(h_out_, h_out_cap_, (size_t)n_q * k * 2);
        enqueue_query(libs_[lib], ws_, h_q_, n_q, k, stream_, nullptr);
        BS_HIP(hipMemcpyAsync(h_out_, ws_.topk.get(), (size_t)n_q * k * 8, hipMemcpyDeviceToHost, stream_));
        BS_HIP(hipEventRecord(done_, stream_));
    }
    bool done() override
    {
        const hipError_t e = hipEventQuery(done_);
        if (e == hipSuccess) return true;
        if (e == hipErrorNotReady) return false;
        fail(BS_ERR_HIP, "%s", hipGetErrorString(e));
    }
    const uint32_t* results() const override { return h_out_; }

private:
    template <class T>
    static void ensure_pinned(T*& p, size_t& cap, size_t need)
    {
        if (need <= cap) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = need + need / 4;
        BS_HIP(hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault));
        cap = want;
    }

    bs_context* ctx_;
    std::vector<bs_library*> libs_;
    hipStream_t stream_ = nullptr;
    hipEvent_t done_ = nullptr;
    QueryWorkspace ws_;
    double* h_q_ = nullptr;
    size_t h_q_cap_ = 0;
    uint32_t* h_out_ = nullptr;
    size_t h_out_cap_ = 0;
};

ReqPtr find_ticket(bs_server* srv, uint64_t ticket)
{
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) fail(BS_ERR_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
    return it->second;
}

}  // namespace

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
        std::vector<bs_library*> lv(libs, libs + n_libs);
        ServerSetup cfg;
        cfg.budget = hbm_budget_bytes;
        int max_n = 0, max_L = 0;
        for (auto* l : lv) {
            cfg.libs.push_back({l->P.n, l->P.L});
            max_n = std::max(max_n, l->P.n);
            max_L = std::max(max_L, l->P.L);
        }
        cfg.batch_bytes = [lv](size_t lib, int n_q, int k) {
            return query_workspace_bytes(n_q, lv[lib]->P.n, lv[lib]->P.L, k);
        };
        cfg.k_supported = [lv](size_t lib, int k) { return topk_supported(lv[lib]->P.n, k); };
        // Every slot is sized up front for a merged batch (up to 256 queries of
        // the longest library against the largest one, k up to 1024), and
        // requests are merged only up to that size, so that serving does not
        // allocate, and free, device memory between requests (a free
        // synchronises the whole device and stalls every batch in flight).
        const int pre_q = std::min(256, cfg.max_batch_queries);
        const int pre_k = std::min(1024, max_n);
        const bool fits = hbm_budget_bytes == 0 ||
                          query_workspace_bytes(pre_q, max_n, max_L, pre_k) * (size_t)max_inflight <=
                              hbm_budget_bytes;
        if (fits) cfg.merge_queries = pre_q;
        for (int i = 0; i < max_inflight; ++i) {
            auto ex = std::make_unique<HipExecutor>(ctx, lv);
            if (fits) ex->reserve(pre_q, max_n, max_L, pre_k);
            cfg.executors.push_back(std::move(ex));
        }
        // the server holds a reference to the context and to each library
        bs_context* c = ctx;
        cfg.on_close = [c, lv] {
            for (auto* l : lv) library_release(l);
            context_release(c);
        };
        *out = make_server(std::move(cfg));
        context_retain(ctx);
        for (auto* l : lv) library_retain(l);
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
    if (srv->worker.joinable()) srv->worker.join();  // the worker finishes every queued request first
    srv->slots.clear();  // executors release their streams, workspaces and staging
    auto on_close = std::move(srv->cfg.on_close);
    delete srv;
    if (on_close) on_close();
}

bs_status bs_server_submit(bs_server* srv, size_t lib_index, const double* q_xyz, size_t n_q,
                           size_t k, uint32_t* idx, float* rmsd, uint64_t* ticket)
{
    return guard([&] {
        BS_REQUIRE(srv && q_xyz && idx && rmsd && ticket, "null pointer argument");
        BS_REQUIRE(lib_index < srv->cfg.libs.size(), "library index %zu out of range", lib_index);
        BS_REQUIRE(n_q >= 1 && (int)n_q <= srv->cfg.max_batch_queries, "n_q must be in [1, %d]",
                   srv->cfg.max_batch_queries);
        const ServedLibrary& lib = srv->cfg.libs[lib_index];
        const size_t n_lib = (size_t)lib.n;
        BS_REQUIRE(k >= 1 && k <= 1024 && k <= n_lib, "k must be in [1, min(1024, library size)]");
        if (!srv->cfg.k_supported(lib_index, (int)k))
            fail(BS_ERR_UNSUPPORTED, "k = %zu is beyond the top-k selection's range for %zu library members", k,
                 n_lib);
        auto r = std::make_shared<Request>();
        r->t_submit = Clock::now();
        r->lib = lib_index;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            if (!srv->q_pool.empty()) {
                r->q = std::move(srv->q_pool.back());
                srv->q_pool.pop_back();
            }
        }
        r->q.assign(q_xyz, q_xyz + n_q * (size_t)lib.L * 3);
        r->n_q = (int)n_q;
        r->k = (int)k;
        r->idx = idx;
        r->rmsd = rmsd;
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
            fail(BS_ERR_NOT_READY, "ticket %llu is already running or finished", (unsigned long long)ticket);
        srv->queue.erase(it);
        r->status = BS_ERR_CANCELLED;
        r->error = "cancelled";
        r->t_done = Clock::now();
        srv->ended_locked(r);
        srv->cv_done.notify_all();
    });
}

bs_status bs_server_release(bs_server* srv, uint64_t ticket)
{
    return guard([&] {
        BS_REQUIRE(srv != nullptr, "server is null");
        std::lock_guard<std::mutex> lk(srv->mu);
        ReqPtr r = find_ticket(srv, ticket);
        if (r->status != BS_ERR_NOT_READY) {
            srv->tickets.erase(ticket);
            return;
        }
        // still queued or running: its results are no longer wanted
        r->released = true;
        r->idx = nullptr;
        r->rmsd = nullptr;
        auto it = std::find(srv->queue.begin(), srv->queue.end(), r);
        if (it != srv->queue.end()) {
            srv->queue.erase(it);
            r->status = BS_ERR_CANCELLED;
            r->error = "released before it ran";
            r->t_done = Clock::now();
            srv->tickets.erase(ticket);
            srv->cv_done.notify_all();
        }
    });
}

size_t bs_server_ticket_count(bs_server* srv)
{
    if (!srv) return 0;
    std::lock_guard<std::mutex> lk(srv->mu);
    return srv->tickets.size();
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
