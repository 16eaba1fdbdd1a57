// library.hpp -- device-resident libraries and the query path shared by the
// synchronous query calls and the serving loop.
#pragma once

#include <memory>

#include "engine.hpp"

struct bs_library {
    // the creator's reference plus one per server serving it
    std::atomic<int> refs{1};
    bs_context* ctx = nullptr;  // holds a reference to it
    bs::Prepared P;       // prepared conformations (own reference)
    bs::DevBuf staging;   // raw coordinates (read by the FP64 re-evaluation)
};

namespace bs {
void library_retain(bs_library* lib);
// Drops a reference; the last one frees the library and its context reference.
void library_release(bs_library* lib);
void enqueue_query(bs_library* lib, QueryWorkspace& ws, const double* q_host, int n_q, int k,
                   hipStream_t s, RunRecord* rec);
}  // namespace bs
