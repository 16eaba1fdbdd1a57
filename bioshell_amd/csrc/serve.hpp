// serve.hpp -- the serving loop's request queue, admission and ticket
// bookkeeping, apart from the device work of a batch.
//
// A server owns max_inflight slots, each with a BatchExecutor that runs one
// merged batch at a time.  The product's executor (serve.cpp) stages the
// queries in pinned memory and runs prep + Gram + top-k on its own HIP stream;
// any other executor (a host one, for the host sanitizer build of the queue)
// plugs into the same loop.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/bioshell_gpu.h"

namespace bs {

// One library as admission sees it.
struct ServedLibrary {
    int n = 0;  // members
    int L = 0;  // atoms per member
};

// Runs one batch at a time for one slot.  Every call comes from the server's
// worker thread; launch() and done() run without the server's lock.
class BatchExecutor {
public:
    virtual ~BatchExecutor() = default;
    // Called once on the worker thread before anything else (e.g. to make a
    // device current).
    virtual void bind_thread() {}
    // A host buffer of at least `doubles` values for the next batch's queries.
    virtual double* staging(size_t doubles) = 0;
    // Starts the batch: n_q queries of library `lib` (in staging()), top-k each.
    virtual void launch(size_t lib, int n_q, int k) = 0;
    // Whether the launched batch has finished; throws bs::Error if it failed.
    virtual bool done() = 0;
    // The finished batch's results: n_q x k indices, then n_q x k f32 values.
    virtual const uint32_t* results() const = 0;
};

struct ServerSetup {
    std::vector<ServedLibrary> libs;
    size_t budget = 0;         // bytes that the batches in flight may hold together (0: no cap)
    int merge_queries = 512;   // queries merged into one batch at most
    int max_batch_queries = 512;
    // Bytes a batch of n_q queries of library `lib` with top-k needs.
    std::function<size_t(size_t lib, int n_q, int k)> batch_bytes;
    // Whether top-k of k is available for library `lib` (monotone in k).
    std::function<bool(size_t lib, int k)> k_supported;
    std::vector<std::unique_ptr<BatchExecutor>> executors;  // one per slot
    // Runs when the server is destroyed, after its worker and slots are gone.
    std::function<void()> on_close;
};

// Starts a server's worker thread over the given slots.
bs_server* make_server(ServerSetup setup);

}  // namespace bs
