// batch.cpp -- many independent ensembles (uneven n and L) per call, and the
// Gram tile-shape autotuner.
//
// The batched pipeline is a fixed sequence of launches whatever the number of
// ensembles (kernels_batch.hip).  The host plans the batch: it gives each
// ensemble its slice of the batch-wide device arrays, picks its tile side
// (autotuner table), and deals its tiles to 8 per-XCD work lists by
// longest-processing-time, whole ensembles per list.  The plan goes to the
// device as one metadata block (ensemble table, prefix sums, work lists), and
// the launch sequence is captured once per batch shape in a HIP graph and
// replayed.  Everything per ensemble is read from device memory, so one graph
// serves every batch whose shape (and buffer addresses) match.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>

#include "engine.hpp"

namespace bs {

int default_tile(int n, int /*L*/) { return n <= 96 ? 32 : 64; }

static int choose_tile(const bs_context* ctx, bs_method m, int n, int L)
{
    for (const TuneCell& c : ctx->tune[m])
        if (L <= c.l_max && n <= c.n_max) return c.tile;
    return default_tile(n, L);
}

void BatchWorkspace::bind(Budget* b)
{
    for (DevBuf* d : {&xyz, &out, &cent, &rot, &ref, &partial, &panels, &consts, &bits, &meta})
        d->bind(b);
}

void BatchWorkspace::release()
{
    if (meta_done) (void)hipEventSynchronize(meta_done);
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    exec_key.clear();
    meta_last.clear();
    meta_last_dev = nullptr;
    for (DevBuf* d : {&xyz, &out, &cent, &rot, &ref, &partial, &panels, &consts, &bits, &meta})
        d->reset();
    if (h_meta) (void)hipHostFree(h_meta);
    if (h_io) (void)hipHostFree(h_io);
    h_meta = h_io = nullptr;
    h_meta_cap = h_io_cap = 0;
    if (meta_done) (void)hipEventDestroy(meta_done);
    meta_done = nullptr;
}

namespace {

// Grow-only pinned host buffer (contents not preserved).
void* ensure_pinned(void*& p, size_t& cap, size_t bytes)
{
    if (bytes <= cap) return p;
    if (p) BS_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    const size_t want = bytes + bytes / 8;
    BS_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
    cap = want;
    return p;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The host-side plan of one batch.
constexpr int kLists = 8;  // one work list per XCD

struct Plan {
    static constexpr int kXcdLists = kLists;
    bs_method method = BS_BF16X3_GRAM;
    bool gram = true;
    GramMode mode = kGramBF16x3;
    std::vector<int> src;  // caller's index of each planned ensemble (n >= 2)
    std::vector<int> n, L, n_pad, n_kblk, tile, n_ref, slices;
    std::vector<int64_t> xyz_off, ref_off, part_off, pan_off, con_off, out_off;
    std::vector<int> conf_prefix, fit_prefix, slice_prefix, coord_prefix;
    std::vector<int64_t> word_prefix;
    std::vector<uint32_t> items;
    int list_start[kXcdLists + 1] = {};
    std::vector<uint64_t> t64, t32;
    int64_t xyz_doubles = 0, ref_doubles = 0, part_doubles = 0, pan_bytes = 0, con_doubles = 0,
            out_values = 0;
    uint64_t pairs = 0;
    std::vector<GramThresholds> th;
};

Plan make_plan(const bs_context* ctx, bs_method m, const size_t* n_conf, const size_t* n_atoms,
               size_t n_ens, int force_tile)
{
    Plan P;
    P.method = m;
    P.gram = is_gram(m);
    P.mode = mode_of(m);
    for (size_t e = 0; e < n_ens; ++e)
        if (n_conf[e] >= 2) P.src.push_back((int)e);
    const int E = (int)P.src.size();
    P.conf_prefix.assign(E + 1, 0);
    P.fit_prefix.assign(E + 1, 0);
    P.slice_prefix.assign(E + 1, 0);
    P.coord_prefix.assign(E + 1, 0);
    P.word_prefix.assign(E + 1, 0);
    int64_t conf_total = 0, pad_total = 0, fit_total = 0, slice_total = 0, coord_total = 0;
    std::vector<double> cost(E, 0.0);
    for (int e = 0; e < E; ++e) {
        const int n = (int)n_conf[P.src[e]], L = (int)n_atoms[P.src[e]];
        const int tile = force_tile ? force_tile : choose_tile(ctx, m, n, L);
        const int n_pad = round_up(n, tile), n_kblk = (L + kAtomBlock - 1) / kAtomBlock;
        const int n_ref = ref_sample_count(n);
        const int slices = ref_slices(n_ref);
        const int64_t pairs = (int64_t)n * (n - 1) / 2;
        P.n.push_back(n);
        P.L.push_back(L);
        P.n_pad.push_back(n_pad);
        P.n_kblk.push_back(n_kblk);
        P.tile.push_back(tile);
        P.n_ref.push_back(n_ref);
        P.slices.push_back(slices);
        P.th.push_back(gram_thresholds(P.mode, L, n_kblk));
        P.xyz_off.push_back(P.xyz_doubles);
        P.xyz_doubles += (int64_t)n * L * 3;
        P.ref_off.push_back(P.ref_doubles);
        P.ref_doubles += align_up((size_t)(3 * L + 6), 2);
        P.part_off.push_back(P.part_doubles);
        P.part_doubles += (int64_t)slices * 3 * L;
        P.pan_off.push_back(P.pan_bytes);
        P.pan_bytes += (int64_t)(n_pad / kConfBlock) * n_kblk * kPanelBlockU4 * 16;
        P.con_off.push_back(P.con_doubles);
        P.con_doubles += (int64_t)kNumConsts * n_pad;
        P.out_off.push_back(P.out_values);
        P.out_values += pairs;
        P.pairs += (uint64_t)pairs;
        conf_total += n;
        pad_total += n_pad;
        fit_total += fit_groups(n_pad);
        slice_total += slices;
        coord_total += 3 * L;
        BS_REQUIRE(conf_total < (1ll << 30) && pad_total < (1ll << 30) && coord_total < (1ll << 30),
                   "batch too large");
        P.conf_prefix[e + 1] = (int)conf_total;
        P.fit_prefix[e + 1] = (int)fit_total;
        P.slice_prefix[e + 1] = (int)slice_total;
        P.coord_prefix[e + 1] = (int)coord_total;
        P.word_prefix[e + 1] = P.word_prefix[e] + (int64_t)align_up((size_t)(pairs + 31) / 32, 4);
        // cost of the ensemble's Gram work in work-group tile times: a 64-tile
        // and a quad of 32-tiles each take about (n_kblk + 13) units (K loop +
        // epilogue)
        const int64_t T = n_pad / tile;
        const double tiles = (double)T * (T + 1) / 2;
        cost[e] = (tile == 64 ? tiles : tiles / 4.0) * (n_kblk + 13);
    }
    if (!P.gram || E == 0) return P;

    // Whole ensembles to 8 lists, largest first, each to the least-loaded list.
    std::vector<int> order(E);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<std::vector<int>> lists(Plan::kXcdLists);
    double load[Plan::kXcdLists] = {};
    for (int e : order) {
        const int x = (int)(std::min_element(load, load + Plan::kXcdLists) - load);
        lists[x].push_back(e);
        load[x] += cost[e];
    }
    for (int x = 0; x < Plan::kXcdLists; ++x) {
        P.list_start[x] = (int)P.items.size();
        std::vector<uint64_t> pend;  // 32-tiles waiting to fill a quad
        auto flush = [&](bool all) {
            while (pend.size() >= 4 || (all && !pend.empty())) {
                const size_t q = P.t32.size() / 4;
                for (int w = 0; w < 4; ++w) {
                    if (!pend.empty() && w < (int)pend.size()) {
                        P.t32.push_back(pend[w]);
                    } else {
                        P.t32.push_back(~0ull);
                    }
                }
                pend.erase(pend.begin(), pend.begin() + std::min<size_t>(4, pend.size()));
                BS_REQUIRE(q < 0x7FFFFFFFu, "batch too large");
                P.items.push_back(0x80000000u | (uint32_t)q);
            }
        };
        for (int e : lists[x]) {
            const int T = P.n_pad[e] / P.tile[e];
            BS_REQUIRE(T < 65536, "ensemble too large for the batch tile codes");
            for (int tr = 0; tr < T; ++tr)
                for (int tc = tr; tc < T; ++tc) {
                    const uint64_t code = ((uint64_t)e << 32) | ((uint64_t)tr << 16) | (uint64_t)tc;
                    if (P.tile[e] == 64) {
                        BS_REQUIRE(P.t64.size() < 0x7FFFFFFFu, "batch too large");
                        P.items.push_back((uint32_t)P.t64.size());
                        P.t64.push_back(code);
                    } else {
                        pend.push_back(code);
                        flush(false);
                    }
                }
        }
        flush(true);
    }
    P.list_start[Plan::kXcdLists] = (int)P.items.size();
    return P;
}

// Appends `bytes` of `src` to blob at a 256-byte boundary; returns the offset.
size_t put(std::vector<unsigned char>& blob, const void* src, size_t bytes)
{
    const size_t off = align_up(blob.size(), 256);
    blob.resize(off + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(blob.data() + off, src, bytes);
    return off;
}

// The batch's descriptors on the device: blob goes up (through the pinned
// staging buffer, on stream s) unless the device copy already holds it.
char* upload_meta(BatchWorkspace& W, std::vector<unsigned char>& blob, hipStream_t s)
{
    const size_t meta_bytes = align_up(blob.size(), 256);
    blob.resize(meta_bytes);
    const void* meta_before = W.meta.get();
    char* meta = static_cast<char*>(W.meta.ensure(meta_bytes, "batch metadata"));
    if (meta != meta_before || meta != W.meta_last_dev || blob != W.meta_last) {
        if (!W.meta_done) BS_HIP(hipEventCreateWithFlags(&W.meta_done, hipEventDisableTiming));
        BS_HIP(hipEventSynchronize(W.meta_done));  // the previous upload has left the pinned buffer
        void* h = ensure_pinned(W.h_meta, W.h_meta_cap, meta_bytes);
        std::memcpy(h, blob.data(), meta_bytes);
        BS_HIP(hipMemcpyAsync(meta, h, meta_bytes, hipMemcpyHostToDevice, s));
        BS_HIP(hipEventRecord(W.meta_done, s));
        W.meta_last.swap(blob);
        W.meta_last_dev = meta;
    }
    return meta;
}

// The per-pair kernel over the batch: the ensembles grouped by tile block
// (plain_block), one launch per group over all of its tiles.
void enqueue_plain(bs_context* ctx, const Plan& P, const std::vector<const double*>& d_xyz,
                   const std::vector<float*>& d_out, RunRecord* rec)
{
    hipStream_t s = ctx->stream;
    const bool fp64 = P.method == BS_FP64_PLAIN;
    const int E = (int)P.src.size();
    std::vector<PlainEns> grp[3];  // blocks 1, 2, 4
    std::vector<int> pre[3];
    for (int g = 0; g < 3; ++g) pre[g].push_back(0);
    for (int e = 0; e < E; ++e) {
        const int r = plain_block(P.n[e], fp64, false, ctx->n_cu);
        const int g = r == 1 ? 0 : r == 2 ? 1 : 2;
        const int nt = plain_tiles_per_side(P.n[e], r);
        const int64_t total = (int64_t)pre[g].back() + (int64_t)nt * (nt + 1) / 2;
        BS_REQUIRE(total < (1ll << 31), "batch too large for the per-pair kernel");
        grp[g].push_back({d_xyz[e], d_out[e], P.n[e], P.L[e], nt, 0});
        pre[g].push_back((int)total);
    }
    std::vector<unsigned char> blob;
    size_t o_ens[3], o_pre[3];
    for (int g = 0; g < 3; ++g) {
        o_ens[g] = put(blob, grp[g].data(), grp[g].size() * sizeof(PlainEns));
        o_pre[g] = put(blob, pre[g].data(), pre[g].size() * 4);
    }
    char* meta = upload_meta(ctx->batch, blob, s);
    rec->mark({kBTotal0, kBPrep0, kBPrep1, kBMain0}, s);
    uint64_t launches = 0;
    for (int g = 0; g < 3; ++g) {
        if (grp[g].empty()) continue;
        launch_plain_batch(reinterpret_cast<const PlainEns*>(meta + o_ens[g]),
                           reinterpret_cast<const int*>(meta + o_pre[g]), (int)grp[g].size(),
                           pre[g].back(), 1 << g, fp64, s);
        ++launches;
    }
    BS_HIP(hipGetLastError());
    rec->mark({kBMain1, kBFallback0, kBFallback1, kBTotal1}, s);
    rec->base.main_launches = launches;
}

// Enqueues the batch on the context's stream.  d_xyz[e] / d_out[e] are the
// device input and output of planned ensemble e.
void enqueue(bs_context* ctx, const Plan& P, const std::vector<const double*>& d_xyz,
             const std::vector<float*>& d_out)
{
    BatchWorkspace& W = ctx->batch;
    hipStream_t s = ctx->stream;
    RunRecord* rec = ctx->begin_record();
    const int E = (int)P.src.size();
    rec->base.pairs = P.pairs;
    rec->base.main_launches = 1;
    rec->base.tile_m = rec->base.tile_n = 0;
    if (E == 0) {
        rec->mark({kBTotal0, kBTotal1, kBPrep0, kBPrep1, kBMain0, kBMain1, kBFallback0, kBFallback1}, s);
        return;
    }
    if (!P.gram) {
        enqueue_plain(ctx, P, d_xyz, d_out, rec);
        return;
    }

    // device arrays of the batch
    const int64_t total_conf = P.conf_prefix[E];
    double* cent = static_cast<double*>(W.cent.ensure((size_t)total_conf * 3 * 8, "batch centroids"));
    double* part = static_cast<double*>(W.partial.ensure((size_t)std::max<int64_t>(P.part_doubles, 1) * 8, "batch partial sums"));
    double* rot = static_cast<double*>(W.rot.ensure((size_t)total_conf * 9 * 8, "batch rotations"));
    double* ref = static_cast<double*>(W.ref.ensure((size_t)P.ref_doubles * 8, "batch references"));
    char* pan = static_cast<char*>(W.panels.ensure((size_t)P.pan_bytes, "batch panels"));
    double* con = static_cast<double*>(W.consts.ensure((size_t)P.con_doubles * 8, "batch constants"));
    const int64_t total_words = P.word_prefix[E];
    uint32_t* bits = static_cast<uint32_t*>(W.bits.ensure(
        (size_t)(total_words + 16) * 4 + (size_t)kFlagListCap * 8, "batch flag bitmaps and queue"));

    std::vector<BatchEns> ens(E);
    for (int e = 0; e < E; ++e) {
        BatchEns& b = ens[e];
        const int64_t c0 = P.conf_prefix[e];
        b.xyz = d_xyz[e];
        b.cent = cent + c0 * 3;
        b.rot = rot + c0 * 9;
        b.ref = ref + P.ref_off[e];
        b.partial = part + P.part_off[e];
        b.panels = pan + P.pan_off[e];
        b.consts = con + P.con_off[e];
        b.bits = bits + P.word_prefix[e];
        b.out = d_out[e];
        b.n = P.n[e];
        b.L = P.L[e];
        b.n_pad = P.n_pad[e];
        b.n_kblk = P.n_kblk[e];
        b.n_ref = P.n_ref[e];
        b.n_slices = P.slices[e];
        b.tile = P.tile[e];
        b.lim = epilogue_limits(P.th[e], P.L[e]);
    }
    std::vector<unsigned char> blob;
    blob.reserve(4096);
    const size_t o_ens = put(blob, ens.data(), ens.size() * sizeof(BatchEns));
    const size_t o_pp = put(blob, P.fit_prefix.data(), P.fit_prefix.size() * 4);
    const size_t o_sp = put(blob, P.slice_prefix.data(), P.slice_prefix.size() * 4);
    const size_t o_kp = put(blob, P.coord_prefix.data(), P.coord_prefix.size() * 4);
    const size_t o_wp = put(blob, P.word_prefix.data(), P.word_prefix.size() * 8);
    const size_t o_ls = put(blob, P.list_start, sizeof(P.list_start));
    const size_t o_it = put(blob, P.items.data(), P.items.size() * 4);
    const size_t o_64 = put(blob, P.t64.data(), P.t64.size() * 8);
    const size_t o_32 = put(blob, P.t32.data(), P.t32.size() * 8);
    char* meta = upload_meta(W, blob, s);

    BatchArgs a{};
    a.ens = reinterpret_cast<const BatchEns*>(meta + o_ens);
    a.n_ens = E;
    a.mode = P.mode;
    a.fit_prefix = reinterpret_cast<const int*>(meta + o_pp);
    a.slice_prefix = reinterpret_cast<const int*>(meta + o_sp);
    a.coord_prefix = reinterpret_cast<const int*>(meta + o_kp);
    a.word_prefix = reinterpret_cast<const int64_t*>(meta + o_wp);
    a.total_fit = P.fit_prefix[E];
    a.total_slices = P.slice_prefix[E];
    a.total_coords = P.coord_prefix[E];
    a.total_words = total_words;
    a.items = reinterpret_cast<const uint32_t*>(meta + o_it);
    a.list_start = reinterpret_cast<const int*>(meta + o_ls);
    a.tiles64 = reinterpret_cast<const uint64_t*>(meta + o_64);
    a.tiles32 = reinterpret_cast<const uint64_t*>(meta + o_32);
    a.ens_bits0 = bits;
    a.counters = bits + total_words;
    a.flag_list = reinterpret_cast<int64_t*>(bits + total_words + 16);
    a.n_wg = std::max(1, std::min(2 * ctx->n_cu, (int)P.items.size()));

    // graph of this shape: the kernel arguments are all that it bakes in
    std::vector<unsigned char> key(sizeof(BatchArgs) + 2 * sizeof(void*));
    std::memcpy(key.data(), &a, sizeof(BatchArgs));
    std::memcpy(key.data() + sizeof(BatchArgs), &rot, sizeof(void*));
    std::memcpy(key.data() + sizeof(BatchArgs) + sizeof(void*), &bits, sizeof(void*));
    if (!W.exec || key != W.exec_key) {
        if (W.exec) BS_HIP(hipGraphExecDestroy(W.exec));
        W.exec = nullptr;
        W.exec_key.clear();
        hipGraph_t graph = nullptr;
        BS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            launch_batch_prep(a, s);
            launch_batch_gram(a, bits, total_words + 16, s);
            launch_batch_fallback(a, s);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        BS_HIP(hipStreamEndCapture(s, &graph));
        const hipError_t ie = hipGraphInstantiate(&W.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BS_HIP(ie);
        W.exec_key.swap(key);
        ++W.graph_builds;
    }
    // the stages run inside the graph: only the total is timed
    rec->mark({kBTotal0, kBPrep0, kBPrep1, kBMain0, kBMain1, kBFallback0, kBFallback1}, s);
    BS_HIP(hipGraphLaunch(W.exec, s));
    rec->mark({kBTotal1}, s);
    ++W.graph_launches;
    if (rec->h_count)
        BS_HIP(hipMemcpyAsync(rec->h_count, a.counters + kLists, 4, hipMemcpyDeviceToHost, s));
    rec->gram = true;
    rec->base.tile_m = rec->base.tile_n = P.tile.empty() ? 0 : P.tile[0];
}

void check_batch_args(bs_context* ctx, const size_t* n_conf, const size_t* n_atoms, size_t n_ens,
                      bs_method m)
{
    BS_REQUIRE(ctx != nullptr, "context is null");
    BS_REQUIRE(m >= BS_FP64_PLAIN && m <= BS_BF16X3_GRAM, "unknown method %d", (int)m);
    BS_REQUIRE(n_ens == 0 || (n_conf && n_atoms), "null pointer argument");
    BS_REQUIRE(n_ens < (1u << 20), "too many ensembles (%zu)", n_ens);
    for (size_t e = 0; e < n_ens; ++e) {
        BS_REQUIRE(n_atoms[e] >= 1 && n_atoms[e] <= (1u << 20),
                   "ensemble %zu: n_atoms must be in [1, 2^20] (got %zu)", e, n_atoms[e]);
        BS_REQUIRE(n_conf[e] <= (1u << 21), "ensemble %zu: n_conf too large (%zu)", e, n_conf[e]);
    }
}

}  // namespace

}  // namespace bs

using namespace bs;

extern "C" {

bs_status bs_crmsd_batch_device(bs_context* ctx, const double* const* d_xyz, const size_t* n_conf,
                                const size_t* n_atoms, size_t n_ens, bs_method method,
                                float* const* d_out_upper)
{
    return guard([&] {
        check_batch_args(ctx, n_conf, n_atoms, n_ens, method);
        BS_REQUIRE(n_ens == 0 || (d_xyz && d_out_upper), "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Plan P = make_plan(ctx, method, n_conf, n_atoms, n_ens, 0);
        std::vector<const double*> xs;
        std::vector<float*> os;
        for (int e : P.src) {
            BS_REQUIRE(d_xyz[e] && d_out_upper[e], "ensemble %d: null pointer", e);
            xs.push_back(d_xyz[e]);
            os.push_back(d_out_upper[e]);
        }
        enqueue(ctx, P, xs, os);
    });
}

bs_status bs_crmsd_batch(bs_context* ctx, const double* const* xyz, const size_t* n_conf,
                         const size_t* n_atoms, size_t n_ens, bs_method method,
                         float* const* out_upper)
{
    return guard([&] {
        ctx = resolve_context(ctx);
        check_batch_args(ctx, n_conf, n_atoms, n_ens, method);
        BS_REQUIRE(n_ens == 0 || (xyz && out_upper), "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const Plan P = make_plan(ctx, method, n_conf, n_atoms, n_ens, 0);
        const int E = (int)P.src.size();
        if (E == 0) return;
        BatchWorkspace& W = ctx->batch;
        hipStream_t s = ctx->stream;
        for (int e : P.src) BS_REQUIRE(xyz[e] && out_upper[e], "ensemble %d: null pointer", e);
        const size_t in_bytes = (size_t)P.xyz_doubles * 8, out_bytes = (size_t)P.out_values * 4;
        double* dx = static_cast<double*>(W.xyz.ensure(in_bytes, "batch coordinates"));
        float* dout = static_cast<float*>(W.out.ensure(out_bytes, "batch rmsd matrices"));
        // the staging buffer is free: every earlier use ended with a synchronisation
        char* h = static_cast<char*>(ensure_pinned(W.h_io, W.h_io_cap, std::max(in_bytes, out_bytes)));
        std::vector<Span> in;
        std::vector<const double*> xs;
        std::vector<float*> os;
        for (int e = 0; e < E; ++e) {
            in.push_back({h + P.xyz_off[e] * 8, xyz[P.src[e]], (size_t)P.n[e] * P.L[e] * 24});
            xs.push_back(dx + P.xyz_off[e]);
            os.push_back(dout + P.out_off[e]);
        }
        parallel_copy(in);
        BS_HIP(hipMemcpyAsync(dx, h, in_bytes, hipMemcpyHostToDevice, s));
        enqueue(ctx, P, xs, os);
        BS_HIP(hipMemcpyAsync(h, dout, out_bytes, hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
        std::vector<Span> out;
        for (int e = 0; e < E; ++e)
            out.push_back({out_upper[P.src[e]], h + P.out_off[e] * 4,
                           (size_t)P.n[e] * (P.n[e] - 1) / 2 * 4});
        parallel_copy(out);
    });
}

}  // extern "C"

// ----------------------------------------------------------------- autotuner
namespace {

const int kTuneL[] = {64, 128, 256, 512, 1 << 20};
const int kTuneLRep[] = {48, 100, 200, 400, 700};
const int kTuneN[] = {64, 128, 256, 512, 1 << 21};
const int kTuneNRep[] = {40, 100, 200, 400, 640};

std::string tune_json(bs_method m, const std::vector<TuneCell>& cells)
{
    std::ostringstream o;
    o << "{\n  \"arch\": \"gfx950\",\n  \"method\": " << (int)m << ",\n  \"cells\": [\n";
    for (size_t k = 0; k < cells.size(); ++k) {
        const TuneCell& c = cells[k];
        char line[256];
        std::snprintf(line, sizeof(line),
                      "    {\"l_max\": %d, \"n_max\": %d, \"tile\": %d, \"ms_tile32\": %.5f, "
                      "\"ms_tile64\": %.5f}%s\n",
                      c.l_max, c.n_max, c.tile, c.ms32, c.ms64, k + 1 < cells.size() ? "," : "");
        o << line;
    }
    o << "  ]\n}\n";
    return o.str();
}

// Integer or number after "key": in s from pos; false if absent before `end`.
bool field(const std::string& s, size_t pos, size_t end, const char* key, double* out)
{
    const std::string k = std::string("\"") + key + "\"";
    const size_t p = s.find(k, pos);
    if (p == std::string::npos || p >= end) return false;
    const size_t c = s.find(':', p + k.size());
    if (c == std::string::npos || c >= end) return false;
    char* stop = nullptr;
    *out = std::strtod(s.c_str() + c + 1, &stop);
    return stop != s.c_str() + c + 1;
}

void load_tune(bs_context* ctx, const char* path)
{
    std::ifstream f(path);
    BS_REQUIRE(f.good(), "cannot read autotune cache '%s'", path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    BS_REQUIRE(s.find("\"gfx950\"") != std::string::npos, "autotune cache '%s' is not for gfx950", path);
    double mv = -1;
    BS_REQUIRE(field(s, 0, s.size(), "method", &mv) && mv >= 0 && mv <= 3,
               "autotune cache '%s' has no valid method", path);
    const size_t cells = s.find("\"cells\"");
    BS_REQUIRE(cells != std::string::npos, "autotune cache '%s' has no cells", path);
    std::vector<TuneCell> out;
    size_t pos = cells;
    while (true) {
        const size_t a = s.find('{', pos);
        if (a == std::string::npos) break;
        const size_t b = s.find('}', a);
        BS_REQUIRE(b != std::string::npos, "autotune cache '%s' is malformed", path);
        TuneCell c;
        double v;
        BS_REQUIRE(field(s, a, b, "l_max", &v), "cell without l_max in '%s'", path);
        c.l_max = (int)v;
        BS_REQUIRE(field(s, a, b, "n_max", &v), "cell without n_max in '%s'", path);
        c.n_max = (int)v;
        BS_REQUIRE(field(s, a, b, "tile", &v) && (v == 32 || v == 64), "cell tile must be 32 or 64 in '%s'", path);
        c.tile = (int)v;
        if (field(s, a, b, "ms_tile32", &v)) c.ms32 = v;
        if (field(s, a, b, "ms_tile64", &v)) c.ms64 = v;
        out.push_back(c);
        pos = b + 1;
    }
    ctx->tune[(int)mv] = out;
}

}  // namespace

extern "C" {

bs_status bs_autotune_load(bs_context* ctx, const char* cache_path)
{
    return guard([&] {
        BS_REQUIRE(ctx != nullptr && cache_path != nullptr, "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        load_tune(ctx, cache_path);
    });
}

bs_status bs_autotune(bs_context* ctx, bs_method method, const char* cache_path)
{
    return guard([&] {
        BS_REQUIRE(ctx != nullptr, "context is null");
        BS_REQUIRE(is_gram(method), "the autotuner chooses Gram tile shapes; method %d has none",
                   (int)method);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream;
        Timer tm;
        tm.create();
        std::vector<TuneCell> cells;
        DevBuf base, outs;
        base.bind(&ctx->budget);
        outs.bind(&ctx->budget);
        const int pool = 700;
        for (int li = 0; li < 5; ++li) {
            const int L = kTuneLRep[li];
            // one pool of conformations per atom count; the batch's ensembles
            // are overlapping slices of it
            std::vector<double> h((size_t)pool * L * 3);
            BS_REQUIRE(bs_generate_ensemble(900 + li, pool, L, 0.5, 3.0, 0.0, 0.0, h.data()) == 0,
                       "generator failed");
            double* d = static_cast<double*>(base.ensure(h.size() * 8, "autotune pool"));
            BS_HIP(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, s));
            for (int ni = 0; ni < 5; ++ni) {
                const int n = kTuneNRep[ni];
                const int64_t per = (int64_t)n * (n - 1) / 2;
                const int E = (int)std::max<int64_t>(1, std::min<int64_t>(512, 3000000 / per));
                std::vector<size_t> nc(E, (size_t)n), na(E, (size_t)L);
                float* o = static_cast<float*>(outs.ensure((size_t)E * per * 4, "autotune output"));
                std::vector<const double*> xs(E);
                std::vector<float*> os(E);
                for (int e = 0; e < E; ++e) {
                    xs[e] = d + (size_t)((e * 37) % (pool - n + 1)) * L * 3;
                    os[e] = o + (size_t)e * per;
                }
                TuneCell c;
                c.l_max = kTuneL[li];
                c.n_max = kTuneN[ni];
                for (int tile : {32, 64}) {
                    const Plan P = make_plan(ctx, method, nc.data(), na.data(), (size_t)E, tile);
                    enqueue(ctx, P, xs, os);  // builds the graph
                    float best = 1e30f;
                    for (int r = 0; r < 3; ++r) {
                        tm.start(s);
                        enqueue(ctx, P, xs, os);
                        tm.stop(s);
                        best = std::min(best, tm.ms());
                    }
                    (tile == 32 ? c.ms32 : c.ms64) = best;
                }
                c.tile = c.ms32 <= c.ms64 ? 32 : 64;
                cells.push_back(c);
            }
        }
        BS_HIP(hipStreamSynchronize(s));
        tm.destroy();
        ctx->tune[method] = cells;
        if (cache_path && *cache_path) {
            std::ofstream f(cache_path);
            BS_REQUIRE(f.good(), "cannot write autotune cache '%s'", cache_path);
            f << tune_json(method, cells);
        }
    });
}

}  // extern "C"
