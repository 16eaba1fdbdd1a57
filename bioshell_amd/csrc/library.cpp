// library.cpp -- device-resident conformation libraries and query-vs-library
// cRMSD (full matrix or per-query top-k).
//
// A library is prepared once (centring, fit onto its Procrustes-mean
// reference, operand panels, constants) and kept in HBM.  Queries are centred
// and fitted onto the library's reference, packed the same way, and scored by
// the rectangular Gram kernel (rows = queries, columns = library) with the
// FP64 QCP epilogue; top-k then selects per query.
#include <algorithm>
#include <cstring>

#include "library.hpp"

using namespace bs;

namespace bs {

// Top-k results (n_q x k indices, then values) followed by the selection's candidates.
size_t topk_bytes(int n_q, int n_lib, int k)
{
    return (size_t)n_q * k * 8 + (k > 0 ? topk_scratch_bytes(n_q, n_lib, k) : 0);
}

// Tiles of a query batch: 64-row tiles of the queries x 64-column tiles of the library.
static size_t query_tiles(int n_q, int n_lib)
{
    return (size_t)((n_q + 63) / 64) * (size_t)((n_lib + 63) / 64);
}

size_t query_workspace_bytes(int n_q, int n_lib, int L, int k)
{
    return prepared_bytes(n_q, L) + (size_t)n_q * L * 3 * 8 + (size_t)n_q * n_lib * 4 +
           topk_bytes(n_q, n_lib, k) + GramScratch::order_bytes(query_tiles(n_q, n_lib)) + 4096;
}

void QueryWorkspace::bind(Budget* b)
{
    prep.bind(b);
    xyz.bind(b);
    scores.bind(b);
    topk.bind(b);
    gsc.bind(b);
}

void QueryWorkspace::reset()
{
    prep.reset();
    xyz.reset();
    scores.reset();
    topk.reset();
    gsc.reset();
}

void QueryWorkspace::reserve(int n_q, int n_lib, int L, int k)
{
    const int n_pad = round_up(n_q, 64), n_kblk = (L + kAtomBlock - 1) / kAtomBlock;
    prep.cent.ensure((size_t)n_q * 3 * 8, "centroids");
    prep.rot.ensure((size_t)n_q * 9 * 8, "rotations");
    prep.panels.ensure((size_t)n_pad / kConfBlock * n_kblk * kPanelBlockU4 * 16, "operand panels");
    prep.consts.ensure((size_t)kNumConsts * n_pad * 8, "constants");
    xyz.ensure((size_t)n_q * L * 3 * 8, "query coordinates");
    scores.ensure((size_t)n_q * n_lib * 4, "query scores");
    topk.ensure(topk_bytes(n_q, n_lib, k), "top-k");
    gsc.flags.ensure(flag_region_bytes((uint64_t)n_q * n_lib, kFlagListCap), "flag bitmap and queue");
    gsc.reserve_orders(query_tiles(n_q, n_lib));
}

// Enqueues the whole query path on stream s: H2D of the queries (from
// q_host, which must stay valid until the stream passes the copy), prep
// against the library's reference, Gram + epilogue + FP64 re-evaluation into
// ws.scores (n_q x n_lib), and, when k > 0, top-k into ws.topk.
void enqueue_query(bs_library* lib, QueryWorkspace& ws, const double* q_host, int n_q, int k,
                   hipStream_t s, RunRecord* rec)
{
    const int L = lib->P.L, n_lib = lib->P.n;
    double* dq = static_cast<double*>(ws.xyz.ensure((size_t)n_q * L * 3 * 8, "query coordinates"));
    float* scores =
        static_cast<float*>(ws.scores.ensure((size_t)n_q * n_lib * 4, "query scores"));
    BS_HIP(hipMemcpyAsync(dq, q_host, (size_t)n_q * L * 3 * 8, hipMemcpyHostToDevice, s));
    if (rec) rec->mark({kBTotal0, kBPrep0}, s);
    prepare(ws.prep, dq, n_q, L, lib->P.mode, lib->P.ref.as<double>(), s);
    GramRun run;
    run.triangle = false;
    run.d_out = scores;
    run.out_f64 = false;
    run.out_ld = n_lib;
    ws.gsc.max_wg = lib->ctx->gsc.max_wg;
    run_gram(view_of(ws.prep), view_of(lib->P), run, ws.gsc, s, rec);
    if (k > 0) {
        uint32_t* idx = static_cast<uint32_t*>(ws.topk.ensure(topk_bytes(n_q, n_lib, k), "top-k"));
        float* val = reinterpret_cast<float*>(idx + (size_t)n_q * k);
        launch_topk(scores, n_q, n_lib, k, idx, val, idx + (size_t)n_q * k * 2, s);
        BS_HIP(hipGetLastError());
    }
    if (rec) rec->mark({kBTotal1}, s);
    if (rec) {
        rec->base.pairs = (uint64_t)n_q * n_lib;
        rec->base.main_launches = 1;
        rec->base.tile_m = 64;
        rec->base.tile_n = 64;
    }
}

}  // namespace bs

namespace bs {
void library_retain(bs_library* lib) { lib->refs.fetch_add(1, std::memory_order_relaxed); }

void library_release(bs_library* lib)
{
    if (lib->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    bs_context* ctx = lib->ctx;
    (void)guard([&] {
        BS_HIP(hipSetDevice(ctx->device));
        lib->P.reset();
        lib->staging.reset();
    });
    delete lib;
    context_release(ctx);
}
}  // namespace bs

extern "C" {

bs_status bs_library_create(bs_context* ctx, const double* xyz, size_t n_conf, size_t n_atoms,
                            bs_library** out)
{
    return guard([&] {
        BS_REQUIRE(ctx != nullptr && out != nullptr, "null pointer argument");
        BS_REQUIRE(xyz != nullptr, "xyz is null");
        BS_REQUIRE(n_conf >= 1 && n_atoms >= 1, "a library needs >= 1 conformation of >= 1 atom");
        BS_REQUIRE(n_conf < (1u << 21) && n_atoms <= (1u << 20), "library too large");
        *out = nullptr;
        std::lock_guard<std::mutex> lk(ctx->mu);
        BS_HIP(hipSetDevice(ctx->device));
        auto lib = std::make_unique<bs_library>();
        lib->ctx = ctx;
        lib->P.bind(&ctx->budget);
        lib->staging.bind(&ctx->budget);
        hipStream_t s = ctx->stream;
        const size_t bytes = n_conf * n_atoms * 3 * 8;
        double* d = static_cast<double*>(lib->staging.ensure(bytes, "library coordinates"));
        BS_HIP(hipMemcpyAsync(d, xyz, bytes, hipMemcpyHostToDevice, s));
        // the coordinates stay resident beside the prepared set: the FP64
        // re-evaluation of flagged query pairs reads them
        prepare(lib->P, d, (int)n_conf, (int)n_atoms, kGramBF16x3, nullptr, s);
        BS_HIP(hipStreamSynchronize(s));
        lib->P.partial.reset();
        context_retain(ctx);
        *out = lib.release();
    });
}

void bs_library_destroy(bs_library* lib)
{
    if (lib) library_release(lib);
}

size_t bs_library_size(const bs_library* lib) { return lib ? (size_t)lib->P.n : 0; }
size_t bs_library_atoms(const bs_library* lib) { return lib ? (size_t)lib->P.L : 0; }

static void check_query(bs_context* ctx, bs_library* lib, const double* q, size_t n_q)
{
    BS_REQUIRE(ctx != nullptr && lib != nullptr, "null context or library");
    BS_REQUIRE(lib->ctx == ctx, "library belongs to another context");
    BS_REQUIRE(n_q >= 1 && n_q < (1u << 21), "n_q must be in [1, 2^21)");
    BS_REQUIRE(q != nullptr, "query coordinates are null");
}

bs_status bs_query_all(bs_context* ctx, bs_library* lib, const double* q_xyz, size_t n_q,
                       float* out)
{
    return guard([&] {
        check_query(ctx, lib, q_xyz, n_q);
        BS_REQUIRE(out != nullptr, "out is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        BS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ctx->qws.bind(&ctx->budget);
        enqueue_query(lib, ctx->qws, q_xyz, (int)n_q, 0, s, ctx->begin_record());
        BS_HIP(hipMemcpyAsync(out, ctx->qws.scores.get(), n_q * lib->P.n * 4,
                              hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
    });
}

bs_status bs_query_topk(bs_context* ctx, bs_library* lib, const double* q_xyz, size_t n_q,
                        size_t k, uint32_t* idx, float* rmsd)
{
    return guard([&] {
        check_query(ctx, lib, q_xyz, n_q);
        BS_REQUIRE(idx != nullptr && rmsd != nullptr, "null output");
        BS_REQUIRE(k >= 1 && k <= 1024 && k <= (size_t)lib->P.n,
                   "k must be in [1, min(1024, library size)]");
        std::lock_guard<std::mutex> lk(ctx->mu);
        BS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ctx->qws.bind(&ctx->budget);
        enqueue_query(lib, ctx->qws, q_xyz, (int)n_q, (int)k, s, ctx->begin_record());
        const uint32_t* d_idx = ctx->qws.topk.as<uint32_t>();
        const float* d_val = reinterpret_cast<const float*>(d_idx + n_q * k);
        BS_HIP(hipMemcpyAsync(idx, d_idx, n_q * k * 4, hipMemcpyDeviceToHost, s));
        BS_HIP(hipMemcpyAsync(rmsd, d_val, n_q * k * 4, hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
