// streamed.cpp -- all-vs-all to a host triangle in T x T blocks, copies
// overlapping compute.
//
// The ensemble is prepared once; the upper triangle of blocks (T conformations
// a side, T a multiple of 64) is then walked row by row.  Block k is computed
// on the context's stream into device slot k % 2 (a packed triangle for
// diagonal blocks, a dense T-stride block otherwise), copied to pinned slot
// k % 2 on the copy stream, and scattered row by row into the caller's packed
// triangle by host threads while blocks k+1 and k+2 are in flight.  Device
// memory holds two blocks of results instead of the whole triangle.
#include <algorithm>
#include <cstring>

#include "engine.hpp"

using namespace bs;

namespace {

// Row spans of block (bi, bj) in the caller's packed triangle.
void scatter_block(int n, int T, int bi, int bj, const float* blk, float* out)
{
    const int r0 = bi * T, c0 = bj * T;
    const int nr = std::min(T, n - r0), nc = std::min(T, n - c0);
    std::vector<Span> spans;
    spans.reserve(nr);
    auto tri = [n](int64_t i, int64_t j) { return i * n - i * (i + 1) / 2 + (j - i - 1); };
    for (int i = 0; i < nr; ++i) {
        const int64_t I = r0 + i;
        if (bi == bj) {
            const int64_t len = nr - 1 - i;
            if (len <= 0) continue;
            const int64_t loc = (int64_t)i * nr - (int64_t)i * (i + 1) / 2;  // (i, i + 1)
            spans.push_back({out + tri(I, I + 1), blk + loc, (size_t)len * 4});
        } else {
            spans.push_back({out + tri(I, c0), blk + (int64_t)i * T, (size_t)nc * 4});
        }
    }
    parallel_copy(spans);
}

}  // namespace

extern "C" {

bs_status bs_crmsd_all_vs_all_streamed(bs_context* ctx, const double* xyz, size_t n_conf,
                                       size_t n_atoms, bs_method method, size_t tile,
                                       float* out_upper)
{
    return guard([&] {
        BS_REQUIRE(n_atoms >= 1 && n_atoms <= (1u << 20), "n_atoms must be in [1, 2^20]");
        BS_REQUIRE(n_conf <= (1u << 21), "n_conf too large (%zu)", n_conf);
        BS_REQUIRE(method >= BS_FP64_PLAIN && method <= BS_BF16X3_GRAM, "unknown method %d",
                   (int)method);
        BS_REQUIRE(tile >= 64 && tile % 64 == 0 && tile <= 16384,
                   "tile must be a multiple of 64 in [64, 16384] (got %zu)", tile);
        if (n_conf < 2) return;
        BS_REQUIRE(xyz && out_upper, "null pointer argument");
        ctx = resolve_context(ctx);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream, cs = ctx->copy_stream;
        const int n = (int)n_conf, L = (int)n_atoms;
        const size_t in_bytes = n_conf * n_atoms * 24;
        double* d = static_cast<double*>(ctx->ws_xyz.ensure(in_bytes, "ensemble coordinates"));
        BS_HIP(hipMemcpyAsync(d, xyz, in_bytes, hipMemcpyHostToDevice, s));
        if (!is_gram(method)) {
            // the per-pair kernels write the whole triangle in one launch
            const size_t out_bytes = n_conf * (n_conf - 1) / 2 * 4;
            void* o = ctx->ws_out.ensure(out_bytes, "rmsd matrix");
            avsa_device(ctx, d, n, L, method, o, false);
            BS_HIP(hipMemcpyAsync(out_upper, o, out_bytes, hipMemcpyDeviceToHost, s));
            BS_HIP(hipStreamSynchronize(s));
            return;
        }
        const int T = std::min((int)tile, round_up(n, 64));
        const int nb = (n + T - 1) / T;
        std::vector<std::pair<int, int>> blocks;
        for (int bi = 0; bi < nb; ++bi)
            for (int bj = bi; bj < nb; ++bj) blocks.push_back({bi, bj});
        const int K = (int)blocks.size();
        const size_t blk_bytes = (size_t)T * T * 4;
        float* dev[2];
        for (int k = 0; k < 2; ++k) {
            dev[k] = static_cast<float*>(ctx->stream_blk[k].ensure(blk_bytes, "streamed block"));
            if (!ctx->ev_compute[k])
                BS_HIP(hipEventCreateWithFlags(&ctx->ev_compute[k], hipEventDisableTiming));
            if (!ctx->ev_copy[k]) BS_HIP(hipEventCreateWithFlags(&ctx->ev_copy[k], hipEventDisableTiming));
        }
        if (blk_bytes > ctx->h_blk_cap) {
            for (int k = 0; k < 2; ++k) {
                if (ctx->h_blk[k]) BS_HIP(hipHostFree(ctx->h_blk[k]));
                ctx->h_blk[k] = nullptr;
            }
            ctx->h_blk_cap = 0;
            for (int k = 0; k < 2; ++k)
                BS_HIP(hipHostMalloc(&ctx->h_blk[k], blk_bytes, hipHostMallocDefault));
            ctx->h_blk_cap = blk_bytes;
        }
        if ((size_t)K > ctx->h_blk_counts_cap) {
            if (ctx->h_blk_counts) BS_HIP(hipHostFree(ctx->h_blk_counts));
            ctx->h_blk_counts = nullptr;
            ctx->h_blk_counts_cap = 0;
            BS_HIP(hipHostMalloc((void**)&ctx->h_blk_counts, (size_t)K * 4, hipHostMallocDefault));
            ctx->h_blk_counts_cap = (size_t)K;
        }
        RunRecord* rec = ctx->begin_record();
        rec->base.pairs = (uint64_t)n * (n - 1) / 2;
        rec->base.main_launches = (uint64_t)K;
        rec->base.tile_m = rec->base.tile_n = T;
        rec->mark({kBTotal0, kBPrep0}, s);
        prepare(ctx->prep, d, n, L, mode_of(method), nullptr, s);
        rec->mark({kBPrep1, kBMain0}, s);
        auto drain = [&](int k) {  // block k's copy has been issued: wait, scatter
            BS_HIP(hipEventSynchronize(ctx->ev_copy[k % 2]));
            scatter_block(n, T, blocks[k].first, blocks[k].second,
                          static_cast<const float*>(ctx->h_blk[k % 2]), out_upper);
        };
        for (int k = 0; k < K; ++k) {
            const int slot = k % 2;
            const int bi = blocks[k].first, bj = blocks[k].second;
            if (k >= 2) BS_HIP(hipStreamWaitEvent(s, ctx->ev_copy[slot], 0));  // device slot free
            GramRun run;
            run.triangle = bi == bj;
            run.d_out = dev[slot];
            run.out_f64 = false;
            run.out_ld = T;
            run.h_count = ctx->h_blk_counts + k;
            run_gram(view_of(ctx->prep, bi * T, T), view_of(ctx->prep, bj * T, T), run, ctx->gsc, s,
                     nullptr);
            BS_HIP(hipEventRecord(ctx->ev_compute[slot], s));
            if (k >= 2) drain(k - 2);  // pinned slot free
            const int nr = std::min(T, n - bi * T);
            const size_t bytes = run.triangle ? (size_t)nr * (nr - 1) / 2 * 4 : (size_t)nr * T * 4;
            BS_HIP(hipStreamWaitEvent(cs, ctx->ev_compute[slot], 0));
            BS_HIP(hipMemcpyAsync(ctx->h_blk[slot], dev[slot], bytes, hipMemcpyDeviceToHost, cs));
            BS_HIP(hipEventRecord(ctx->ev_copy[slot], cs));
        }
        rec->mark({kBMain1, kBFallback0, kBFallback1, kBTotal1}, s);
        for (int k = std::max(0, K - 2); k < K; ++k) drain(k);
        BS_HIP(hipStreamSynchronize(s));
        uint64_t flagged = 0;
        for (int k = 0; k < K; ++k) flagged += ctx->h_blk_counts[k];
        rec->base.flagged_pairs = flagged;
        rec->gram = false;  // flagged_pairs is final in base
    });
}

}  // extern "C"
