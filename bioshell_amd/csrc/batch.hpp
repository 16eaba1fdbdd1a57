// batch.hpp -- batches of independent ensembles and the tile-shape autotuner.
#pragma once

#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace bs {

// One autotuner cell: ensembles with L <= l_max and n <= n_max (first match
// in table order) use Gram tiles of side `tile`.
struct TuneCell {
    int l_max = 0, n_max = 0, tile = 64;
    double ms32 = 0.0, ms64 = 0.0;  // measured when the cell was tuned
};

// Tile side for an ensemble of n conformations of L atoms: the context's
// table for the method, else the built-in rule (32 for n <= 96).
int default_tile(int n, int L);

// Device and pinned host memory of the batched pipeline, plus the
// instantiated graph of the last batch shape.
struct BatchWorkspace {
    DevBuf xyz, out, cent, rot, ref, partial, panels, consts, bits, meta;
    void* h_meta = nullptr;  // pinned: descriptors, prefix sums, work lists
    size_t h_meta_cap = 0;
    void* h_io = nullptr;    // pinned: staging of host inputs and outputs
    size_t h_io_cap = 0;
    hipEvent_t meta_done = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> exec_key;
    std::vector<unsigned char> meta_last;
    void* meta_last_dev = nullptr;
    uint64_t graph_builds = 0, graph_launches = 0;
    void bind(Budget* b);
    void release();
};

}  // namespace bs
