// engine.hpp -- host orchestration shared by the ABI entry points.
#pragma once

#include <atomic>
#include <initializer_list>

#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "qcp.hpp"

namespace bs {

// `ctx`, or for a null `ctx` the process-wide default context (device 0, no
// HBM budget, created on first use and kept until the process ends).
bs_context* resolve_context(bs_context* ctx);

// Takes the current device's address of the no-allocation entry point's
// scratch array (once per device; called when a context is created on it).
void noalloc_init(int device);

// Conformations of one atom count, prepared on the device for the Gram path:
// the raw coordinates they came from (borrowed: they must outlive every run
// that reads this set, for the FP64 re-evaluation), centroids, rotations
// onto the common reference, the reference itself (+ R0), operand panels and
// constants.
struct Prepared {
    int n = 0, n_pad = 0, L = 0, n_kblk = 0;
    GramMode mode = kGramBF16x3;
    const double* xyz = nullptr;
    DevBuf cent, rot, ref, partial, panels, consts;
    void bind(Budget* b)
    {
        cent.bind(b); rot.bind(b); ref.bind(b); partial.bind(b); panels.bind(b); consts.bind(b);
    }
    void reset()
    {
        cent.reset(); rot.reset(); ref.reset(); partial.reset(); panels.reset(); consts.reset();
        xyz = nullptr;
    }
    size_t bytes() const
    {
        return cent.size() + rot.size() + ref.size() + partial.size() + panels.size() +
               consts.size();
    }
};

// Rows [first, first + count) of a prepared set, as the Gram launch reads
// them; first is a multiple of 64 (a whole number of 64-row tiles).
struct PrepView {
    const void* panels = nullptr;
    const double* consts = nullptr;
    const double* xyz = nullptr;
    const double* cent = nullptr;
    int ld = 0;         // constant stride (the set's n_pad)
    int n = 0;          // rows in the view
    int L = 0, n_kblk = 0;
    GramMode mode = kGramBF16x3;
};
PrepView view_of(const Prepared& P, int first = 0, int count = -1);

// Bytes prepare() will hold for n conformations of L atoms.
size_t prepared_bytes(int n, int L);

// The device buffers of a prepared set as plain pointers (those of a
// Prepared, or carved from any other memory of prepared_sizes' sizes).
struct PreparedPtrs {
    double* cent = nullptr;     // n x 3
    double* rot = nullptr;      // n x 9
    double* ref = nullptr;      // 3L + 6 (unused with a fixed reference)
    double* partial = nullptr;  // reference-mean slices x 3L (unused with a fixed reference)
    void* panels = nullptr;
    double* consts = nullptr;   // kNumConsts x n_pad
};
struct PreparedSizes {
    size_t cent, rot, ref, partial, panels, consts;
};
PreparedSizes prepared_sizes(int n, int L);

// Reference, fit, pack (prep_device.hpp).  If `fixed_ref` (device, 3L + 6
// doubles, the layout of Prepared::ref) is given, conformations are fitted
// onto it, read in place (Prepared::ref stays unused); otherwise the reference
// is one Procrustes-mean step from conformation 0 over the reference samples.
// d_xyz is borrowed (Prepared::xyz).
void prepare(Prepared& P, const double* d_xyz, int n, int L, GramMode mode,
             const double* fixed_ref, hipStream_t s);
// The same into caller-provided buffers; returns the view the Gram run reads.
PrepView prepare_into(const PreparedPtrs& p, const double* d_xyz, int n, int L, GramMode mode,
                      const double* fixed_ref, hipStream_t s);

// Flagging thresholds of the Gram epilogue's error model at atom count L.
struct GramThresholds {
    double kappa_dup, u_rep, acc_coef, tau, rep_g_max;
};
GramThresholds gram_thresholds(GramMode mode, int L, int n_kblk);
// The same as the epilogue takes them (scale 1/L, squared tolerances).
EpilogueLimits epilogue_limits(const GramThresholds& t, int L);

struct GramRun {
    uint32_t* h_count = nullptr;  // pinned: receives the flagged-pair count (async)
    bool triangle = true;
    void* d_out = nullptr;
    bool out_f64 = false;
    int64_t out_ld = 0;
    bool ends_total = false;      // the record's total ends with this run
    // row statistics from the epilogue (triangle, f32 output): tile results
    // (fused_stats_tiles(n) x 1 KiB) and the n keys of each statistic, which
    // the run leaves decoded to f32
    unsigned* stats_part = nullptr;
    unsigned* st_min = nullptr;
    unsigned* st_max = nullptr;
};

// Scratch a Gram run needs besides the prepared sets.
struct GramScratch {
    // Tile orders of the last few launch shapes (tile rows x tile cols x
    // triangle): a streamed all-vs-all alternates between a few shapes.
    struct TileOrder {
        uint64_t key = ~0ull;
        int n = 0;
        DevBuf buf;
    };
    static constexpr int kOrders = 6;
    DevBuf flags;
    // flags.generation() of the allocation whose region is known clean
    // (zeroed here, or left so by a launch sequence that was fully enqueued)
    uint64_t flags_clean = 0;
    TileOrder orders[kOrders];
    int next_order = 0;
    int max_wg = 0;               // persistent Gram work-groups (2 per CU)
    uint32_t* h_order = nullptr;  // pinned staging of a tile order being uploaded
    size_t h_order_cap = 0;
    hipEvent_t order_done = nullptr;
    void bind(Budget* b)
    {
        flags.bind(b);
        for (auto& o : orders) o.buf.bind(b);
    }
    void reset();
    // Device tile order for the shape; uploaded on first use (stream-ordered).
    const uint32_t* tile_order(int tiles_r, int tiles_c, bool triangle, hipStream_t s, int* n);
    // Sizes every cached order and the pinned staging for up to max_tiles
    // tiles, so that new shapes up to that size never allocate.
    void reserve_orders(size_t max_tiles);
    static size_t order_bytes(size_t max_tiles) { return kOrders * (max_tiles * 4 + 4); }
};

// Events and host-visible results of one call, read lazily.
// Stage boundaries of a call's record: 2k starts stage k, 2k + 1 ends it.
enum { kRecTotal = 0, kRecPrep = 1, kRecMain = 2, kRecFallback = 3 };
enum {
    kBTotal0 = 0, kBTotal1 = 1, kBPrep0 = 2, kBPrep1 = 3,
    kBMain0 = 4, kBMain1 = 5, kBFallback0 = 6, kBFallback1 = 7
};

// Timing of one call.  Each mark() records one event for the boundaries it
// names, so boundaries with no work between them share an event: an event
// record costs the stream a few microseconds, as much as a small kernel.
struct RunRecord {
    hipEvent_t ev[8] = {};        // in the order recorded
    int8_t slot[8] = {-1, -1, -1, -1, -1, -1, -1, -1};  // boundary -> index into ev
    int n_ev = 0;
    uint32_t* h_count = nullptr;  // flagged-pair count (pinned), stored by the resolve kernel
    bs_run_stats base{};          // pairs, launches, tile sizes
    bool gram = false;
    bool used = false;
    void create();
    void destroy();
    void reset();
    void mark(std::initializer_list<int> bounds, hipStream_t s);
    bs_run_stats read() const;  // synchronises on the last event
};

// Gram contraction + QCP epilogue + FP64 re-evaluation of flagged pairs over
// rows of A x cols of B (or the upper triangle of A when triangle).  Asynchronous:
// the flagged-pair count is stored to rec->h_count when rec is given.
// Words of the flag bitmap of n_out outputs (a whole number of 16-byte units).
inline uint64_t flag_words(uint64_t n_out) { return ((n_out + 31) / 32 + 3) / 4 * 4; }
// Bytes of the flag region: four control words (count, report, done, spare),
// list_cap queued entries, then the bitmap.  The list sits at a fixed offset
// and the bitmap grows at the end, so a region reused for a larger launch
// never finds list entries where its bitmap is.
inline size_t flag_region_bytes(uint64_t n_out, int list_cap)
{
    return 16 + (size_t)list_cap * 8 + (size_t)flag_words(n_out) * 4;
}
// The Gram launch sequence with every buffer given: `order` (n_tiles tile
// codes on the device), `flags` (flag_region_bytes(n_out, list_cap)); no
// allocation, no host copy except the optional counts of `run` / `rec`.
// Each launch leaves the region's control words zero and its bitmap clear;
// `clear_flags` zeroes the region first, for memory that may hold anything.
void run_gram_with(const PrepView& A, const PrepView& B, const GramRun& run, const uint32_t* order,
                   int n_tiles, uint32_t* flags, int list_cap, bool clear_flags, int max_wg, hipStream_t s,
                   RunRecord* rec);
void run_gram(const PrepView& A, const PrepView& B, const GramRun& run, GramScratch& sc,
              hipStream_t s, RunRecord* rec);

// Workspace of one in-flight query batch (query prep, score matrix, top-k).
struct QueryWorkspace {
    Prepared prep;
    DevBuf xyz, scores, topk;
    GramScratch gsc;
    void bind(Budget* b);
    void reset();
    // Allocates up front what a batch of n_q queries of L atoms against n_lib
    // conformations with top-k needs, so that such batches never reallocate
    // (a device free synchronises the whole device).
    void reserve(int n_q, int n_lib, int L, int k);
};

// Device bytes a query batch of n_q against n_lib conformations of L atoms needs.
size_t query_workspace_bytes(int n_q, int n_lib, int L, int k);

GramMode mode_of(bs_method m);
bool is_gram(bs_method m);

// All-vs-all on device-resident coordinates; fills ctx->last.  With d_rmin
// set (f32 output), also the row minima and maxima of the triangle
// (launch_row_stats' results): from the Gram epilogue where the method has one
// and the triangle has at most kFusedStatsMaxTiles tiles, else by
// launch_row_stats over the written triangle.
void avsa_device(bs_context* ctx, const double* d_xyz, int n, int L, bs_method m, void* d_out,
                 bool out_f64, float* d_rmin = nullptr, float* d_rmax = nullptr);

}  // namespace bs

// The context: one device, one stream, workspaces, stats, autotune choices.
struct bs_context {
    // the creator's reference plus one per library and server made from it:
    // the context goes when the last of them is destroyed, in any order
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    bs::Budget budget{0};
    bs::Prepared prep;         // the last ensemble's prepared set
    bs::GramScratch gsc;
    bs::DevBuf ws_xyz, ws_out, ws_misc;
    bs::DevBuf ws_stats;  // per-tile partial results of the row statistics
    hipStream_t copy_stream = nullptr;  // device-to-host copies overlapping the compute stream
    bs::DevBuf stream_blk[2];          // streamed all-vs-all: result blocks on the device
    void* h_blk[2] = {nullptr, nullptr};  // pinned result blocks
    size_t h_blk_cap = 0;
    uint32_t* h_blk_counts = nullptr;  // pinned flagged counts, one per block
    size_t h_blk_counts_cap = 0;
    hipEvent_t ev_compute[2] = {}, ev_copy[2] = {};
    bs::QueryWorkspace qws;    // synchronous query calls
    std::vector<bs::RunRecord> ring;  // the last calls' records
    size_t ring_next = 0, ring_used = 0;
    uint32_t* h_counts = nullptr;      // pinned, one per record
    bs::RunRecord* begin_record();     // next ring slot, recycled
    std::vector<bs::TuneCell> tune[4];  // autotuner: Gram tile side per (L, n) cell, per method
    bs::BatchWorkspace batch;          // batched pipeline (bs_crmsd_batch*)
    int n_cu = 0;
    std::mutex mu;
};

namespace bs {
void context_retain(bs_context* ctx);
// Drops a reference; the last one frees the context's streams and memory.
void context_release(bs_context* ctx);
}  // namespace bs
