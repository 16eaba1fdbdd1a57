// library.hpp -- device-resident libraries and the query path shared by the
// synchronous query calls and the serving loop.
#pragma once

#include <memory>

#include "engine.hpp"

struct bs_library {
    bs_context* ctx = nullptr;
    bs::Prepared P;       // prepared conformations (own reference)
    bs::DevBuf staging;   // raw coordinates (read by the FP64 re-evaluation)
};

namespace bs {
void enqueue_query(bs_library* lib, QueryWorkspace& ws, const double* q_host, int n_q, int k,
                   hipStream_t s, RunRecord* rec);
}  // namespace bs
