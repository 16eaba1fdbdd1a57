// kernels.hpp -- host-side launch wrappers for the device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bs {

// ------------------------------------------------------------ plain path
// Per-pair kernel over the upper triangle of one ensemble: centre, 3x3
// covariance, textbook QCP, sqrt -- in T = double (fp64 != 0) or float.
// xyz: device, n x L x 3 f64 AoS.  out: packed triangle, f64 or f32.
// The tile side is 16 R conformations, R = 1, 2 or 4 (fp32) and 1 or 2
// (fp64), chosen by plain_block from the pair slots the tiles cover and, for
// a lone launch (alone), from how many rounds of 2 work-groups per CU they
// take on n_cu CUs.
void launch_plain_avsa(const double* xyz, int n, int L, bool fp64, void* out, bool out_f64,
                       int n_cu, hipStream_t s);
int plain_block(int n, bool fp64, bool alone, int n_cu);
inline int plain_tiles_per_side(int n, int r) { return (n + 16 * r - 1) / (16 * r); }
// One ensemble of a batched plain launch (f32 output); nt tiles per side.
struct PlainEns {
    const double* xyz;
    void* out;
    int n, L, nt, pad;
};
// Every ensemble of ens[0 .. n_ens) with block r in one launch: prefix (n_ens + 1
// entries, device) counts their tiles, tiles = prefix[n_ens].
void launch_plain_batch(const PlainEns* ens, const int* prefix, int n_ens, int tiles, int r, bool fp64,
                        hipStream_t s);

// ------------------------------------------------------------ Gram path
enum GramMode : int { kGramBF16x3 = 0, kGramFP32 = 1 };

// Operand panels: conformations in blocks of 32, atoms in blocks of 16.
// One (conf block, atom block) is 6 KiB = 384 x 16 B:
//   BF16x3: 6 planes (x_hi, x_lo, y_hi, y_lo, z_hi, z_lo) x 64 lanes x 8 bf16,
//           lane l = conformation (l % 32), atoms 8*(l/32) .. +7 of the block;
//   FP32:   3 planes (x, y, z) x 64 lanes x 8 f32,
//           lane l = conformation (l % 32), atoms (l/32) + 2s, s = 0..7.
constexpr int kConfBlock = 32;
constexpr int kAtomBlock = 16;
constexpr int kPanelBlockU4 = 384;  // 16-byte units per (conf block, atom block)

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Per-conformation fits (prep_device.hpp): kFitConfs conformations per
// work-group of 256 threads, kFitLanes threads each.
constexpr int kFitConfs = 16;
constexpr int kFitLanes = 256 / kFitConfs;
__host__ __device__ inline int fit_groups(int count) { return (count + kFitConfs - 1) / kFitConfs; }
// The reference mean's partial sums: one per work-group of sample fits.
__host__ __device__ inline int ref_slices(int S) { return fit_groups(S); }

// The reference r of the shifted frame is the Procrustes mean of S(n)
// conformations spread evenly over the ensemble: sample s is conformation
// floor(s n / S), so S = n takes every one.  Any r makes the frame exact; r
// only sets the residual sizes, hence the split's error and the flags.  A
// mean over S samples departs from the full mean by about sigma / sqrt(S)
// per coordinate, which adds 1/S of a typical residual variance, while the
// first two preparation passes read S conformations instead of n.  S grows
// with n (n/8 above 512) so that near-identical conformations of a large
// ensemble keep residuals far below the noise of the rest.  A floor of 64
// rather than 128 left the largest error of a 512-ensemble batch unchanged
// and cut its step by 6 %.
constexpr int kRefSampleMin = 64;
__host__ __device__ inline int ref_sample_count(int n)
{
    const int s = n / 8 > kRefSampleMin ? n / 8 : kRefSampleMin;
    return n < s ? n : s;
}
__host__ __device__ inline int ref_sample_index(int s, int n, int S)
{
    return (int)((int64_t)s * n / S);
}

// Reference of one ensemble (prep_device.hpp): for the S = ref_sample_count(n)
// samples, the slices' partial sums of the fitted samples (partial,
// ref_slices(S) x 3L), then r (ref: 3L values) and R0 (6).
void launch_reference(const double* xyz, int n, int L, double* partial, double* ref, hipStream_t s);
// Fit onto ref and pack: rotations (rot), centroids (cent), operand panels of
// e_i = rot_i (x_i - c_i) - ref and the constants, for n_pad rows.
void launch_fit_pack(const double* xyz, const double* ref, int n, int n_pad, int L, int n_kblk,
                     GramMode mode, double* rot, double* cent, void* panels, double* consts, int ld,
                     hipStream_t s);

// The epilogue's flagging limits and RMSD scale at one atom count L, formed
// on the host so that the kernels hold them in scalar registers rather than
// recompute them per pair.
struct EpilogueLimits {
    double inv_L;        // 1 / L
    double dup_coef;     // flag H < dup_coef sqrt(g_i g_j): near-duplicate residuals
    double acc_coef;     // estimated error of H - mu per sqrt(g_i g_j)
    double tau2L;        // tau^2 L, tau the tolerated RMSD error (A)
    double solve_tol2L;  // the solve's own budget, tau2L / 16
    double rep_h_max;    // h_i + h_j above which the split's rounding of e may show
};

struct GramArgs {
    const void* panels_a;   // row operand panels
    const void* panels_b;   // column operand panels
    const double* consts_a; // row constants (SoA, stride lda)
    const double* consts_b; // column constants (SoA, stride ldb)
    int lda, ldb;
    int n_rows, n_cols;     // valid rows / cols
    int L;                  // real atom count (for the RMSD denominator)
    int n_kblk;             // atom blocks of 16
    const uint32_t* tiles;  // (tile_row << 16 | tile_col) per logical tile
    int n_tiles;
    int triangle;           // 1: packed upper triangle of rows==cols, only i<j
    int64_t out_ld;         // dense output: row stride (triangle: ignored)
    void* out;
    int out_f64;
    uint32_t* flag_bits;    // bit o set: output o is re-evaluated in FP64
    // when flag_list is set, flagged outputs are queued there first
    // (*flag_count counts every flagged output; the bitmap takes those
    // beyond flag_cap)
    int64_t* flag_list;
    uint32_t* flag_count;
    int flag_cap;
    EpilogueLimits lim;     // flagging limits and the RMSD scale
    // launch shape
    int max_wg;             // resident work-groups (2 per CU)
    // Row statistics from the epilogue (triangle, f32 output): when set, tile
    // (tr, tc) writes the 256 keys of row_stats_tile_kernel's layout to
    // stats_part + 256 (tr stats_T - tr (tr - 1) / 2 + tc - tr), leaving out
    // the flagged pairs (the resolve kernel adds their final values).
    unsigned* stats_part;
    int stats_T;            // tiles per side of the triangle
};

// Builds the logical tile order (XCD-aware super-blocks) on the host.
// Returns the number of tiles written to `out` (capacity must be enough).
int build_tile_order(int tiles_r, int tiles_c, bool triangle, int super, uint32_t* out);

void launch_gram(const GramArgs& a, GramMode mode, hipStream_t s);
void launch_zero_u32(uint32_t* p, int n, hipStream_t s);

// FP64 re-evaluation of the outputs the Gram kernel flagged, from raw
// coordinates and centroids of rows (a) and columns (b): one wave per entry of
// `list` (min(count, cap) entries), then, only when count > cap, a scan of the
// bitmap `bits` (n_out bits) that clears every word it reads.  ctl holds
// count, report, done: the last work-group to finish copies count to report
// and zeroes count and done, so the region is ready for the next launch.  It
// also stores the count into h_report_a / h_report_b when given (pinned host
// words, written with a system-scope store: no separate device-to-host copy).
// With st_min / st_max set (triangle, f32 output), each re-evaluated value
// also goes into the row-statistics keys of both its elements (atomic
// min / max on order-preserving keys).
void launch_flag_resolve(const double* xyz_a, const double* cent_a, const double* xyz_b,
                         const double* cent_b, int L, uint32_t* ctl, const int64_t* list, int cap,
                         uint32_t* bits, int64_t n_out, int triangle, int n, int64_t out_ld, void* out,
                         int out_f64, uint32_t* h_report_a, uint32_t* h_report_b, unsigned* st_min,
                         unsigned* st_max, hipStream_t s);
// Queues the set bits of bits[0 .. n_words) as bit indices (word * 32 + bit)
// into list (cap entries), clearing the queued ones; *count += set bits.
void launch_flag_compact(uint32_t* bits, int64_t n_words, int64_t* list, int cap, uint32_t* count,
                         hipStream_t s);
constexpr int kFlagListCap = 1 << 20;  // queued FP64 re-evaluations per call (8 MiB)
void launch_zero_words(uint32_t* p, int64_t n_words, hipStream_t s);

// Per-query top-k over a dense n_q x n_lib score matrix (ascending, ties by index).
// Per-query k smallest of an n_q x n_lib score matrix (k <= min(1024, n_lib)),
// sorted by (value, index): idx and val get n_q x k each.  scratch holds
// topk_scratch_bytes(n_q, n_lib, k).
size_t topk_scratch_bytes(int n_q, int n_lib, int k);
// Whether the two-stage selection handles k of n_lib (k <= 1024, and the
// chunks' candidates fit the merge): launch_topk fails otherwise.  Monotone in
// k for a given n_lib.
bool topk_supported(int n_lib, int k);
void launch_topk(const float* scores, int n_q, int n_lib, int k, uint32_t* idx, float* val, void* scratch,
                 hipStream_t s);

// ------------------------------------------------------------ batches
// One ensemble of a batch, as the batched kernels see it (device memory).
// Per-conformation arrays (centred, rot) are slices of batch-wide arrays.
struct BatchEns {
    const double* xyz;      // n x L x 3 f64 input
    double* cent;           // n x 3 centroids
    double* rot;            // n x 9 rotations
    double* ref;            // 3L + 6
    double* partial;        // n_slices x 3L
    void* panels;           // n_pad / 32 x n_kblk blocks
    double* consts;         // kNumConsts x n_pad
    uint32_t* bits;         // flag bitmap of the n(n-1)/2 outputs
    float* out;             // packed triangle, f32
    int n, L, n_pad, n_kblk;
    int n_ref;              // conformations sampled for the reference mean
    int n_slices, tile;     // reference-mean slices; tile side 32 or 64
    EpilogueLimits lim;     // error-model limits at this L
};

// A batch: ensemble table, the prefix sums that map a batch-wide index
// (reference sample, padded conformation, mean slice, reference coordinate,
// bitmap word) to its ensemble, and the Gram work lists.
//
// Work items are dealt to 8 per-XCD lists (whole ensembles per list, costs
// balanced); a work-group on XCD x pulls items from list x through an atomic
// counter and then helps the other lists.  An item is either one 64 x 64
// tile (bit 31 clear: index into tiles64) or four 32 x 32 tiles, one per
// wave (bit 31 set: index q, tiles32[4q .. 4q+3]).  Tile codes are
// (ensemble << 32) | (tile row << 16) | tile col; ~0 marks an empty slot.
struct BatchArgs {
    const BatchEns* ens;
    int n_ens;
    int mode;
    const int* fit_prefix;   // fit + pack work-groups (fit_groups of n_pad)
    const int* slice_prefix;
    const int* coord_prefix;
    const int64_t* word_prefix;
    int total_fit, total_slices, total_coords;
    int64_t total_words;
    const uint32_t* items;
    const int* list_start;  // 9 entries: list x is items[list_start[x] .. list_start[x+1])
    const uint64_t* tiles64;
    const uint64_t* tiles32;
    uint32_t* ens_bits0;    // the batch-wide bitmap (ensemble 0's bitmap first)
    uint32_t* counters;     // 8 list counters, then the flagged-pair count
    int64_t* flag_list;     // kFlagListCap queued re-evaluations (batch-wide bit indices)
    int n_wg;               // persistent work-groups of the Gram kernel
};

// The batched pipeline's launches, in order (each over all ensembles).
void launch_batch_prep(const BatchArgs& b, hipStream_t s);
void launch_batch_gram(const BatchArgs& b, uint32_t* zero_base, int64_t zero_words, hipStream_t s);
void launch_batch_fallback(const BatchArgs& b, hipStream_t s);

// ------------------------------------------------------------ clustering
// Row minimum and maximum over j != i of a packed triangle, as scans from
// f32::MAX and -f32::MAX with `d < min` / `d > max` give them (NaN never
// counts); scratch holds row_stats_scratch_bytes(n).
size_t row_stats_scratch_bytes(int n);
void launch_row_stats(const float* up, int n, float* rmin, float* rmax, void* scratch, hipStream_t s);
// Order-preserving unsigned keys of f32 bits: every non-NaN float, negative
// ones included, keeps its order.  All-ones is the key of a NaN only, so it
// serves as "no value" for a minimum.
__host__ __device__ inline unsigned fkey(unsigned bits) { return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u); }
__host__ __device__ inline unsigned fkey_inverse(unsigned key) { return (key & 0x80000000u) ? (key ^ 0x80000000u) : ~key; }
constexpr unsigned kKeyFltMax = 0xFF7FFFFFu;     // fkey(bits of f32::MAX)
constexpr unsigned kKeyNegFltMax = 0x00800000u;  // fkey(bits of -f32::MAX)

// Row statistics taken in the Gram epilogue: the tile results of the whole
// triangle at once (1 KiB per tile), for at most kFusedStatsMaxTiles tiles
// (n up to about 23k; larger triangles take launch_row_stats afterwards).
constexpr int kFusedStatsMaxTiles = 1 << 16;
inline size_t fused_stats_tiles(int n)
{
    const size_t T = (size_t)(n + 63) / 64;
    return T * (T + 1) / 2;
}
// Combines the epilogue's tile results into keys (kmin / kmax, n each, n >= 2),
// which the resolve kernel then updates; decode turns the keys into f32.
void launch_row_stats_combine(const unsigned* part, int n, unsigned* kmin, unsigned* kmax, hipStream_t s);
void launch_row_stats_decode(unsigned* kmin, unsigned* kmax, int n, hipStream_t s);
// For each of m members, the largest distance to the other members.
void launch_members_row_max(const float* up, int n, const unsigned* members, int m, float* out,
                            hipStream_t s);
// Agglomerative clustering: the dense matrix D (n rows of ld >= n floats, ld a
// multiple of 4), sizes and a cleared merge state from the packed triangle;
// then n - 1 merge steps (n_part work-groups scan, one decides; state holds
// cluster_state_bytes(), part n_part candidates of cluster_candidate_bytes()).
size_t cluster_state_bytes();
size_t cluster_candidate_bytes();
void launch_cluster_init(const float* up, int n, int ld, float* D, unsigned* sizes, void* state, hipStream_t s);
// One merge; the merge index lives in `state`, so every step's launches are
// identical (a captured run of steps replays as is; steps past the last merge
// do nothing).
void launch_cluster_step(float* D, int n, int ld, void* state, void* part, int n_part, unsigned* sizes,
                         int rule, unsigned* mi, unsigned* mj, float* md, hipStream_t s);

// Single-pair superposition (rotation + translation) in FP64.
void launch_superpose(const double* a, const double* b, int L, double* out13, hipStream_t s);
// Every conformation of xyz (n x L x 3) onto ref (L x 3): rot n x 9, trans
// n x 3, rmsd n and, when out is not null, the superposed n x L x 3.  scratch
// holds superpose_scratch_bytes(n, L); any of rot / trans / rmsd may be null.
size_t superpose_scratch_bytes(int n, int L);
void launch_superpose_ensemble(const double* xyz, int n, int L, const double* ref, double* scratch,
                               double* rot, double* trans, double* rmsd, double* out, hipStream_t s);

}  // namespace bs
