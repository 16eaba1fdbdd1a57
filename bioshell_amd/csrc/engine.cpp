// engine.cpp -- the all-vs-all pipeline: preparation, Gram launch, FP64
// re-evaluation of flagged pairs, statistics.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace bs {

GramMode mode_of(bs_method m) { return m == BS_FP32_GRAM ? kGramFP32 : kGramBF16x3; }
bool is_gram(bs_method m) { return m == BS_FP32_GRAM || m == BS_BF16X3_GRAM; }

static constexpr int kTileRows = 64;  // pairs per work-group tile side

PreparedSizes prepared_sizes(int n, int L)
{
    const size_t n_pad = (size_t)round_up(n, kTileRows);
    const size_t n_kblk = (size_t)(L + kAtomBlock - 1) / kAtomBlock;
    const size_t slices = (size_t)ref_slices(ref_sample_count(n));
    PreparedSizes z;
    z.cent = (size_t)n * 3 * 8;
    z.rot = (size_t)n * 9 * 8;
    z.ref = (size_t)(3 * L + 6) * 8;
    z.partial = slices * 3 * L * 8;
    z.panels = n_pad / kConfBlock * n_kblk * kPanelBlockU4 * 16;
    z.consts = (size_t)kNumConsts * n_pad * 8;
    return z;
}

size_t prepared_bytes(int n, int L)
{
    const PreparedSizes z = prepared_sizes(n, L);
    return z.cent + z.rot + z.ref + z.partial + z.panels + z.consts;
}

static PrepView prep_view(const PreparedPtrs& p, const double* d_xyz, int n, int L, GramMode mode)
{
    PrepView v;
    v.panels = p.panels;
    v.consts = p.consts;
    v.xyz = d_xyz;
    v.cent = p.cent;
    v.ld = round_up(n, kTileRows);
    v.n = n;
    v.L = L;
    v.n_kblk = (L + kAtomBlock - 1) / kAtomBlock;
    v.mode = mode;
    return v;
}

PrepView prepare_into(const PreparedPtrs& p, const double* d_xyz, int n, int L, GramMode mode,
                      const double* fixed_ref, hipStream_t s)
{
    const int n_pad = round_up(n, kTileRows), n_kblk = (L + kAtomBlock - 1) / kAtomBlock;
    const double* ref = fixed_ref;  // a fixed reference is read in place
    if (!fixed_ref) {
        launch_reference(d_xyz, n, L, p.partial, p.ref, s);
        ref = p.ref;
    }
    launch_fit_pack(d_xyz, ref, n, n_pad, L, n_kblk, mode, p.rot, p.cent, p.panels, p.consts, n_pad, s);
    BS_HIP(hipGetLastError());
    return prep_view(p, d_xyz, n, L, mode);
}

void prepare(Prepared& P, const double* d_xyz, int n, int L, GramMode mode,
             const double* fixed_ref, hipStream_t s)
{
    P.n = n;
    P.L = L;
    P.n_pad = round_up(n, kTileRows);
    P.n_kblk = (L + kAtomBlock - 1) / kAtomBlock;
    P.mode = mode;
    P.xyz = d_xyz;
    const PreparedSizes z = prepared_sizes(n, L);
    PreparedPtrs p;
    p.cent = static_cast<double*>(P.cent.ensure(z.cent, "centroids"));
    p.rot = static_cast<double*>(P.rot.ensure(z.rot, "rotations"));
    p.panels = P.panels.ensure(z.panels, "operand panels");
    p.consts = static_cast<double*>(P.consts.ensure(z.consts, "constants"));
    if (!fixed_ref) {
        p.ref = static_cast<double*>(P.ref.ensure(z.ref, "reference"));
        p.partial = static_cast<double*>(P.partial.ensure(z.partial, "reference partial sums"));
    }
    (void)prepare_into(p, d_xyz, n, L, mode, fixed_ref, s);
}

GramThresholds gram_thresholds(GramMode mode, int L, int n_kblk)
{
    GramThresholds t;
    // Error model of the residual Gram: bf16 hi/lo represents e to 2^-16
    // relative (the dropped lo*lo term is bounded by 2^-16 sqrt(g_i g_j));
    // FP32 accumulation adds ~2^-24 sqrt(#accumulations) sqrt(g_i g_j).
    const double n_acc = mode == kGramBF16x3 ? 3.0 * n_kblk : 8.0 * n_kblk;
    t.kappa_dup = 8.0;
    t.u_rep = mode == kGramBF16x3 ? std::ldexp(1.0, -16) : std::ldexp(1.0, -24);
    // Error of H - mu in units of sqrt(g_i g_j): FP32 accumulation
    // (~2^-24 per accumulation, random-walk) plus, for the split, the
    // dropped lo*lo products (|sum lo_i lo_j| <= 2^-16 sqrt(g_i g_j), a sum
    // of 3L terms of random sign: 3 sigma = 3/sqrt(3L) of that bound).
    t.acc_coef = 2.0 * std::ldexp(1.0, -24) * std::sqrt(n_acc);
    if (mode == kGramBF16x3) t.acc_coef += t.u_rep * std::min(1.0, 3.0 / std::sqrt(3.0 * L));
    t.tau = 2e-5;
    // Rounding e to the split representation moves each coordinate by up
    // to u_rep |e|; its effect on RMSD has a standard deviation of about
    // u_rep sqrt(2 (g_i + g_j)) / (3 L).  Pairs whose 6-sigma value could
    // exceed 5e-5 A (tiny L, large residuals) are re-evaluated in FP64.
    const double r = 5e-5 * 3.0 * L / (6.0 * std::sqrt(2.0) * t.u_rep);
    t.rep_g_max = r * r;
    return t;
}

EpilogueLimits epilogue_limits(const GramThresholds& t, int L)
{
    EpilogueLimits e;
    const double Ld = (double)L;
    e.inv_L = 1.0 / Ld;
    e.dup_coef = t.kappa_dup * t.u_rep;
    e.acc_coef = t.acc_coef;
    e.tau2L = t.tau * t.tau * Ld;
    e.solve_tol2L = 0.0625 * e.tau2L;
    e.rep_h_max = 0.5 * t.rep_g_max;  // h = g / 2
    return e;
}

void RunRecord::create()
{
    for (auto& e : ev) BS_HIP(hipEventCreate(&e));
}

void RunRecord::destroy()
{
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
}

void RunRecord::reset()
{
    for (auto& x : slot) x = -1;
    n_ev = 0;
}

void RunRecord::mark(std::initializer_list<int> bounds, hipStream_t s)
{
    BS_REQUIRE(n_ev < 8, "too many events in one run record");
    BS_HIP(hipEventRecord(ev[n_ev], s));
    for (int b : bounds) slot[b] = (int8_t)n_ev;
    ++n_ev;
}

bs_run_stats RunRecord::read() const
{
    bs_run_stats st = base;
    if (!used || n_ev == 0) return st;
    BS_HIP(hipEventSynchronize(ev[n_ev - 1]));
    auto span = [&](int k) {
        const int a = slot[2 * k], b = slot[2 * k + 1];
        if (a < 0 || b < 0 || a == b) return 0.0;
        float v = 0.0f;
        BS_HIP(hipEventElapsedTime(&v, ev[a], ev[b]));
        return (double)v;
    };
    st.total_ms = span(kRecTotal);
    st.prep_ms = span(kRecPrep);
    st.main_ms = span(kRecMain);
    st.fallback_ms = span(kRecFallback);
    st.flagged_pairs = gram && h_count ? *h_count : base.flagged_pairs;
    return st;
}

PrepView view_of(const Prepared& P, int first, int count)
{
    BS_REQUIRE(first >= 0 && first % kTileRows == 0 && first <= P.n, "bad prepared-set view");
    PrepView v;
    v.n = count < 0 ? P.n - first : std::min(count, P.n - first);
    v.panels = static_cast<const char*>(P.panels.get()) +
               (size_t)(first / kConfBlock) * P.n_kblk * kPanelBlockU4 * 16;
    v.consts = P.consts.as<double>() + first;
    v.xyz = P.xyz + (size_t)first * P.L * 3;
    v.cent = P.cent.as<double>() + (size_t)first * 3;
    v.ld = P.n_pad;
    v.L = P.L;
    v.n_kblk = P.n_kblk;
    v.mode = P.mode;
    return v;
}

void run_gram_with(const PrepView& A, const PrepView& B, const GramRun& run, const uint32_t* order,
                   int n_tiles, uint32_t* flags, int list_cap, bool clear_flags, int max_wg, hipStream_t s,
                   RunRecord* rec)
{
    // Flagged outputs go to a list of list_cap positions (counted in the
    // control words at the region's start); any beyond it set bits of a
    // bitmap indexed by output position (one bit per pair: 6.25 MB for 50M
    // pairs), so there is no capacity limit and no host round trip.  One
    // kernel then re-evaluates them in FP64 and leaves the region clean.
    const uint64_t n_out =
        run.triangle ? (uint64_t)A.n * (A.n - 1) / 2 : (uint64_t)A.n * (uint64_t)run.out_ld;
    const uint64_t n_words = flag_words(n_out);
    uint32_t* ctl = flags;  // count, report, done, spare
    int64_t* list = reinterpret_cast<int64_t*>(flags + 4);
    uint32_t* bits = reinterpret_cast<uint32_t*>(list + list_cap);
    if (clear_flags) launch_zero_words(flags, (int64_t)(flag_region_bytes(n_out, list_cap) / 4), s);

    GramArgs g{};
    g.panels_a = A.panels;
    g.panels_b = B.panels;
    g.consts_a = A.consts;
    g.consts_b = B.consts;
    g.lda = A.ld;
    g.ldb = B.ld;
    g.n_rows = A.n;
    g.n_cols = B.n;
    g.L = A.L;
    g.n_kblk = A.n_kblk;
    g.tiles = order;
    g.n_tiles = n_tiles;
    g.triangle = run.triangle ? 1 : 0;
    g.out_ld = run.out_ld;
    g.out = run.d_out;
    g.out_f64 = run.out_f64 ? 1 : 0;
    g.flag_bits = bits;
    g.flag_list = list;
    g.flag_count = ctl;
    g.flag_cap = list_cap;
    g.lim = epilogue_limits(gram_thresholds(A.mode, A.L, A.n_kblk), A.L);
    g.max_wg = max_wg > 0 ? max_wg : 512;
    const bool stats = run.stats_part != nullptr;
    BS_REQUIRE(!stats || (run.triangle && !run.out_f64 && run.st_min && run.st_max),
               "row statistics need the f32 triangle");
    g.stats_part = run.stats_part;
    g.stats_T = (A.n + kTileRows - 1) / kTileRows;
    (void)n_words;

    if (rec) rec->mark({kBPrep1, kBMain0}, s);
    launch_gram(g, A.mode, s);
    if (rec) rec->mark({kBMain1, kBFallback0}, s);
    BS_HIP(hipGetLastError());
    // the statistics of the pairs the epilogue kept, as keys; the resolve
    // kernel then adds the re-evaluated pairs, and the keys become f32
    if (stats) launch_row_stats_combine(run.stats_part, A.n, run.st_min, run.st_max, s);
    launch_flag_resolve(A.xyz, A.cent, B.xyz, B.cent, A.L, ctl, list, list_cap, bits, (int64_t)n_out,
                        g.triangle, A.n, run.out_ld, run.d_out, g.out_f64, rec ? rec->h_count : nullptr,
                        run.h_count, run.st_min, run.st_max, s);
    if (stats) launch_row_stats_decode(run.st_min, run.st_max, A.n, s);
    BS_HIP(hipGetLastError());
    if (rec) {
        if (run.ends_total)
            rec->mark({kBFallback1, kBTotal1}, s);
        else
            rec->mark({kBFallback1}, s);
        rec->gram = true;
    }
}

void run_gram(const PrepView& A, const PrepView& B, const GramRun& run, GramScratch& sc,
              hipStream_t s, RunRecord* rec)
{
    BS_REQUIRE(A.L == B.L && A.mode == B.mode, "prepared sets disagree on atom count or mode");
    const int tiles_r = (A.n + kTileRows - 1) / kTileRows, tiles_c = (B.n + kTileRows - 1) / kTileRows;
    BS_REQUIRE(tiles_r < 65536 && tiles_c < 65536, "too many conformations for one launch");
    int n_tiles = 0;
    const uint32_t* order = sc.tile_order(tiles_r, tiles_c, run.triangle, s, &n_tiles);
    const uint64_t n_out =
        run.triangle ? (uint64_t)A.n * (A.n - 1) / 2 : (uint64_t)A.n * (uint64_t)run.out_ld;
    // every launch sequence leaves the region clean: only an allocation it
    // has not run in (made here or by a workspace's reserve) needs zeroing
    uint32_t* flags = static_cast<uint32_t*>(
        sc.flags.ensure(flag_region_bytes(n_out, kFlagListCap), "flag bitmap and queue"));
    if (sc.flags_clean != sc.flags.generation())
        launch_zero_words(flags, (int64_t)(sc.flags.size() / 16 * 4), s);
    sc.flags_clean = 0;  // a sequence that throws part-way leaves the region unknown
    run_gram_with(A, B, run, order, n_tiles, flags, kFlagListCap, false, sc.max_wg, s, rec);
    sc.flags_clean = sc.flags.generation();
}

void GramScratch::reset()
{
    if (order_done) (void)hipEventSynchronize(order_done);
    flags.reset();
    flags_clean = 0;
    for (auto& o : orders) {
        o.buf.reset();
        o.key = ~0ull;
        o.n = 0;
    }
    if (h_order) (void)hipHostFree(h_order);
    h_order = nullptr;
    h_order_cap = 0;
    if (order_done) (void)hipEventDestroy(order_done);
    order_done = nullptr;
}

void GramScratch::reserve_orders(size_t max_tiles)
{
    for (auto& o : orders) o.buf.ensure(max_tiles * 4 + 4, "tile order");
    if (max_tiles > h_order_cap) {
        if (!order_done) BS_HIP(hipEventCreateWithFlags(&order_done, hipEventDisableTiming));
        BS_HIP(hipEventSynchronize(order_done));
        if (h_order) BS_HIP(hipHostFree(h_order));
        h_order = nullptr;
        h_order_cap = 0;
        BS_HIP(hipHostMalloc((void**)&h_order, max_tiles * 4, hipHostMallocDefault));
        h_order_cap = max_tiles;
    }
}

const uint32_t* GramScratch::tile_order(int tiles_r, int tiles_c, bool triangle, hipStream_t s,
                                        int* n)
{
    const uint64_t key = ((uint64_t)tiles_r << 33) | ((uint64_t)tiles_c << 1) | (triangle ? 1 : 0);
    for (auto& o : orders)
        if (o.key == key) {
            *n = o.n;
            return o.buf.as<uint32_t>();
        }
    // The pinned staging is reused: wait until its last upload has been read.
    if (!order_done) BS_HIP(hipEventCreateWithFlags(&order_done, hipEventDisableTiming));
    BS_HIP(hipEventSynchronize(order_done));
    const size_t cap = (size_t)tiles_r * tiles_c;
    if (cap > h_order_cap) {
        if (h_order) BS_HIP(hipHostFree(h_order));
        h_order = nullptr;
        h_order_cap = 0;
        BS_HIP(hipHostMalloc((void**)&h_order, cap * 4, hipHostMallocDefault));
        h_order_cap = cap;
    }
    TileOrder& o = orders[next_order];
    next_order = (next_order + 1) % kOrders;
    o.key = ~0ull;  // the upload is ordered on s after every earlier launch that read the slot
    const int nt = build_tile_order(tiles_r, tiles_c, triangle, 8, h_order);
    uint32_t* d = static_cast<uint32_t*>(o.buf.ensure((size_t)nt * 4 + 4, "tile order"));
    BS_HIP(hipMemcpyAsync(d, h_order, (size_t)nt * 4, hipMemcpyHostToDevice, s));
    BS_HIP(hipEventRecord(order_done, s));
    o.key = key;
    o.n = nt;
    *n = nt;
    return d;
}

void avsa_device(bs_context* ctx, const double* d_xyz, int n, int L, bs_method m, void* d_out,
                 bool out_f64, float* d_rmin, float* d_rmax)
{
    hipStream_t s = ctx->stream;
    BS_REQUIRE(!d_rmin || (d_rmax && !out_f64 && n >= 2), "row statistics need the f32 triangle");
    // BIOSHELL_FUSED_STATS_MAX_TILES lowers the limit (tests of the other path)
    size_t max_tiles = kFusedStatsMaxTiles;
    if (const char* e = std::getenv("BIOSHELL_FUSED_STATS_MAX_TILES"))
        max_tiles = std::min<size_t>(max_tiles, (size_t)std::max(0, std::atoi(e)));
    const bool fused = d_rmin && is_gram(m) && fused_stats_tiles(n) <= max_tiles;
    RunRecord* rec = ctx->begin_record();
    rec->base.pairs = n >= 2 ? (uint64_t)n * (n - 1) / 2 : 0;
    if (n >= 2 && !is_gram(m)) {
        rec->mark({kBTotal0, kBPrep0, kBPrep1, kBMain0}, s);
        launch_plain_avsa(d_xyz, n, L, m == BS_FP64_PLAIN, d_out, out_f64, ctx->n_cu, s);
        BS_HIP(hipGetLastError());
        rec->mark({kBMain1, kBFallback0, kBFallback1, kBTotal1}, s);
        rec->base.main_launches = 1;
    } else if (n >= 2) {
        rec->mark({kBTotal0, kBPrep0}, s);
        prepare(ctx->prep, d_xyz, n, L, mode_of(m), nullptr, s);
        GramRun run;
        run.triangle = true;
        run.d_out = d_out;
        run.out_f64 = out_f64;
        run.ends_total = true;
        if (fused) {
            run.stats_part = static_cast<unsigned*>(
                ctx->ws_stats.ensure(fused_stats_tiles(n) * 256 * 4, "row statistics tiles"));
            run.st_min = reinterpret_cast<unsigned*>(d_rmin);
            run.st_max = reinterpret_cast<unsigned*>(d_rmax);
        }
        run_gram(view_of(ctx->prep), view_of(ctx->prep), run, ctx->gsc, s, rec);
        rec->base.main_launches = 1;
        rec->base.tile_m = kTileRows;
        rec->base.tile_n = kTileRows;
    } else {
        rec->mark({kBTotal0, kBTotal1, kBPrep0, kBPrep1, kBMain0, kBMain1, kBFallback0, kBFallback1}, s);
    }
    if (d_rmin && !fused) {
        void* scratch = ctx->ws_stats.ensure(row_stats_scratch_bytes(n), "row statistics tiles");
        launch_row_stats(static_cast<const float*>(d_out), n, d_rmin, d_rmax, scratch, s);
        BS_HIP(hipGetLastError());
    }
}

}  // namespace bs

bs::RunRecord* bs_context::begin_record()
{
    constexpr size_t kRing = 64;
    if (ring.empty()) {
        ring.resize(kRing);
        for (auto& r : ring) r.create();
        BS_HIP(hipHostMalloc((void**)&h_counts, kRing * sizeof(uint32_t), hipHostMallocDefault));
        for (size_t k = 0; k < kRing; ++k) ring[k].h_count = h_counts + k;
    }
    bs::RunRecord* r = &ring[ring_next];
    ring_next = (ring_next + 1) % ring.size();
    ring_used = std::min(ring_used + 1, ring.size());
    if (r->used && r->n_ev > 0) BS_HIP(hipEventSynchronize(r->ev[r->n_ev - 1]));  // never reuse in flight
    r->reset();
    r->base = bs_run_stats{};
    r->gram = false;
    r->used = true;
    return r;
}
