/*
 * bioshell_gpu.h -- C ABI of the MI355X all-vs-all cRMSD engine.
 *
 * The reference (dgront/BioShell v4.0.0-dev) exports no C-callable symbols:
 * bioshell-pdb and bioshell-clustering are built as rlib+cdylib
 * (bioshell-pdb/Cargo.toml:11, bioshell-clustering/Cargo.toml:12) but carry no
 * extern "C" items, and the crmsd function the path needs is only announced
 * (bioshell-pdb/Cargo.toml:27, "nalgebra ... for helical axis and crmsd").
 * Each entry point below therefore names the Rust item it stands in for:
 * the intended bioshell_pdb::calc function, the driver template it follows,
 * or the consumer it feeds.
 *
 * Conventions
 *  - Coordinates are f64 x,y,z triples per atom (bioshell-core/src/vec3.rs:43-51),
 *    packed AoS; an ensemble is n_conf x n_atoms x 3, row-major, the layout of
 *    Structure::model_coordinates (bioshell-pdb/src/structure.rs:101).
 *  - All-vs-all results are the packed upper triangle, row-major over i, i<j:
 *    index(i,j) = i*n - i*(i+1)/2 + (j-i-1); n(n-1)/2 values.
 *  - Every call returns a bs_status; 0 is success, negative is an error whose
 *    text is returned by bs_last_error() (per thread).  Nothing unwinds or
 *    aborts across this boundary: preconditions the reference asserts on
 *    (substructures.rs:19) come back as BS_ERR_INVALID_ARGUMENT.
 *  - Host buffers are borrowed for the duration of a synchronous call.
 *    Device memory belongs to the context / library / server that allocated it.
 *  - Calls on one bs_context are serialised by the context (each takes its
 *    lock); distinct contexts run independently.
 *  - Default context.  The synchronous entry points that take host buffers
 *    (bs_crmsd_pair_f64, bs_superpose_f64, bs_superpose_ensemble,
 *    bs_crmsd_all_vs_all, bs_crmsd_all_vs_all_streamed, bs_crmsd_batch,
 *    bs_rmsd_row_stats, bs_rmsd_members_row_max, bs_hierarchical_clustering)
 *    accept ctx == NULL: they then run on the process-wide default context
 *    (device 0, no HBM budget), created by the first such call and kept until
 *    the process ends.  This is what the context-free Rust functions
 *    crmsd(a, b) and crmsd_all_vs_all(models, method) bind to.  If it cannot
 *    be created (no GPU), every call that needs it returns that error.
 */
#ifndef BIOSHELL_GPU_H
#define BIOSHELL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bs_status;

#define BS_OK 0
#define BS_ERR_INVALID_ARGUMENT (-1)
#define BS_ERR_NO_DEVICE (-2)
#define BS_ERR_HIP (-3)
#define BS_ERR_OUT_OF_BUDGET (-4)
#define BS_ERR_UNSUPPORTED (-5)
#define BS_ERR_NOT_READY (-6)
#define BS_ERR_CANCELLED (-7)
#define BS_ERR_INTERNAL (-8)
#define BS_ERR_PARSE (-9) /* malformed structure text (bioshell_pdb::PDBError's parsing variants) */
#define BS_ERR_IO (-10)   /* a file could not be read or written */

/* How the pairwise cRMSD is evaluated.
 *  BS_FP64_PLAIN   per-pair kernel, FP64: centre, 3x3 covariance, QCP, sqrt.
 *  BS_FP32_PLAIN   the same kernel in FP32.
 *  BS_FP32_GRAM    covariance as a Gram contraction on FP32 MFMA, FP64 QCP epilogue.
 *  BS_BF16X3_GRAM  covariance as a Gram contraction on BF16 MFMA with hi/lo
 *                  split operands (hi*hi + hi*lo + lo*hi, FP32 accumulation),
 *                  FP64 QCP epilogue and FP64 re-evaluation of flagged pairs. */
typedef enum bs_method {
    BS_FP64_PLAIN = 0,
    BS_FP32_PLAIN = 1,
    BS_FP32_GRAM = 2,
    BS_BF16X3_GRAM = 3
} bs_method;

typedef struct bs_context bs_context;
typedef struct bs_library bs_library;
typedef struct bs_server bs_server;

/* Timings and counters of the last all-vs-all / batch / query call on a
 * context.  Times are device times from HIP events on the context's stream
 * (adjacent stage boundaries share one event; a batch times only its total). */
typedef struct bs_run_stats {
    double total_ms;        /* first prep kernel start .. last kernel end      */
    double prep_ms;         /* centring, reference fit, packing, flag clear    */
    double main_ms;         /* the pair kernel(s): plain or Gram+QCP            */
    double fallback_ms;     /* FP64 re-evaluation of flagged pairs, count copy */
    uint64_t pairs;         /* pairs evaluated                                 */
    uint64_t flagged_pairs; /* pairs re-evaluated in FP64                       */
    uint64_t main_launches; /* launches of the main pair kernel                 */
    int32_t tile_m;         /* Gram tile rows used (0 for plain methods)        */
    int32_t tile_n;         /* Gram tile cols used                              */
} bs_run_stats;

/* ---------------------------------------------------------------- basics */

/* Text of the last error raised on this thread ("" if none). */
const char* bs_last_error(void);

/* Library version string. */
const char* bs_version(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int bs_device_count(void);

/* Creates a context on `device`.  hbm_budget_bytes caps the device memory the
 * context (and libraries / servers made from it) may allocate; 0 = no cap. */
bs_status bs_context_create(int device, size_t hbm_budget_bytes, bs_context** out);
/* Gives up the creator's hold on the context.  Libraries and servers made
 * from it hold it too, so handles may be destroyed in any order: the context
 * goes when the last of them is destroyed.  No call may use the handle after
 * its own destroy. */
void bs_context_destroy(bs_context* ctx);
bs_status bs_context_synchronize(bs_context* ctx);
/* The hipStream_t (as void*) on which the context launches its kernels. */
void* bs_context_stream(bs_context* ctx);
bs_status bs_context_last_stats(const bs_context* ctx, bs_run_stats* out);
/* Stats of up to max_calls most recent calls (oldest first; the context keeps
 * the last 64); *n_out receives how many were written.  Synchronises on them. */
bs_status bs_context_history(const bs_context* ctx, bs_run_stats* out, size_t max_calls,
                             size_t* n_out);
/* Device bytes currently allocated by the context (workspaces included). */
size_t bs_context_bytes_in_use(const bs_context* ctx);
/* The process-wide default context that ctx == NULL selects (created here if
 * no call has needed it yet).  Never destroy it; it lives until the process
 * ends.  For its stream, stats and history, and for the device entry points. */
bs_status bs_default_context(bs_context** out);

/* ------------------------------------------------------- single pair */

/* cRMSD of two conformations of n_atoms each, FP64 on the GPU.
 * Stands in for the intended `bioshell_pdb::calc::crmsd(a: &[T], b: &[T]) -> f64`
 * (bioshell-pdb/Cargo.toml:27), generic over HasCartesians like
 * SubstructureAxis::from_3d_points (bioshell-pdb/src/calc/substructures.rs:17).
 * ctx may be NULL (the default context): the Rust crmsd(a, b) takes none. */
bs_status bs_crmsd_pair_f64(bs_context* ctx, const double* a_xyz, const double* b_xyz,
                            size_t n_atoms, double* out);

/* Optimal superposition of b onto a: a ~= R b + t, R row-major 3x3, plus the
 * RMSD of the fit.  Output is a bioshell_pdb::calc::Rototranslation
 * (bioshell-pdb/src/calc/transformation.rs:22-37, applied as at 171-174). */
bs_status bs_superpose_f64(bs_context* ctx, const double* a_xyz, const double* b_xyz,
                           size_t n_atoms, double* rot9, double* trans3, double* rmsd);

/* Superposition of every conformation of an ensemble (n_conf x n_atoms x 3)
 * onto one reference -- ref_xyz (n_atoms x 3), or conformation ref_index when
 * ref_xyz is NULL: for conformation m, R_m (row-major 3x3, rot9[9m..]) and
 * t_m (trans3[3m..]) with ref ~= R_m x + t_m, the fit's RMSD (rmsd[m]) and the
 * superposed coordinates R_m x + t_m (superposed, n_conf x n_atoms x 3).
 * Each output may be NULL.  FP64 throughout.  This is the Rototranslation of
 * bioshell-pdb/src/calc/transformation.rs:22-37 (x' = R x + t, apply_mut at
 * 171-174) fitted per model, the companion of writing superposed models
 * (bs_ensemble_write_pdb). */
bs_status bs_superpose_ensemble(bs_context* ctx, const double* xyz, size_t n_conf, size_t n_atoms,
                                const double* ref_xyz, size_t ref_index, double* rot9,
                                double* trans3, double* rmsd, double* superposed);
/* The same on device buffers, asynchronous on the context's stream; d_ref may
 * be NULL (conformation ref_index is the reference); d_superposed must not
 * overlap d_xyz. */
bs_status bs_superpose_ensemble_device(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                       size_t n_atoms, const double* d_ref, size_t ref_index,
                                       double* d_rot9, double* d_trans3, double* d_rmsd,
                                       double* d_superposed);

/* ------------------------------------------------------- all-vs-all */

/* All-vs-all cRMSD of one ensemble; host in, host out.  out_upper holds
 * n_conf(n_conf-1)/2 values, f64 if out_f64 != 0 else f32.
 * Stands in for the absent structural all-vs-all driver, whose template is
 * align_all_pairs (bioshell-seq/src/alignment/alignment_protocols.rs:83-115,
 * triangle at 97); its result is what hierarchical_clustering reads through
 * its Fn(usize,usize)->f32 closure (bioshell-clustering/src/hierarchical/
 * hierarchical.rs:22-26, densified in clustering_matrix.rs:11-21). */
bs_status bs_crmsd_all_vs_all(bs_context* ctx, const double* xyz, size_t n_conf,
                              size_t n_atoms, bs_method method, void* out_upper, int out_f64);

/* The same on device-resident buffers (d_xyz and d_out_upper are device
 * pointers, e.g. from hipMalloc).  The call is asynchronous on the context's
 * stream (see bs_context_stream / bs_context_synchronize); d_xyz must stay
 * valid and d_out_upper unread until the stream reaches the call's end.  Used
 * when the ensemble already lives in HBM. */
bs_status bs_crmsd_all_vs_all_device(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                     size_t n_atoms, bs_method method, void* d_out_upper,
                                     int out_f64);

/* bs_crmsd_all_vs_all_device with f32 output, and in the same call the row
 * minimum and maximum over j != i of the matrix it writes: for each of the
 * n_conf elements, bit for bit what bs_rmsd_row_stats_device returns for
 * d_out_upper afterwards (retrieve_outliers and medoid_by_min_max,
 * hierarchical.rs:198-217 and 105-131, read these).  The Gram methods take
 * them from the kernel's epilogue, so the triangle is not read back; the
 * plain methods, and triangles of more than 65,536 64 x 64 tiles (n_conf
 * above about 23,000), run the row-statistics kernels over d_out_upper.
 * n_conf >= 2.  Device buffers, asynchronous on the context's stream. */
bs_status bs_crmsd_all_vs_all_row_stats_device(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                               size_t n_atoms, bs_method method, float* d_out_upper,
                                               float* d_row_min, float* d_row_max);

/* The same as bs_crmsd_all_vs_all_device, without allocating anything.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its whole workspace -- prepared operand panels and
 * constants, centroids, rotations, the reference, the tile order, the flag
 * bitmap and queue -- lives in one fixed 64 MiB array of the library's device
 * memory, which every GPU the library runs on holds until the process ends;
 * inputs that need more return BS_ERR_OUT_OF_BUDGET with the size (10,000
 * conformations of 300 atoms need about 47 MB).  Results equal
 * bs_crmsd_all_vs_all_device's bit for bit.  Every call on one GPU shares that
 * array: calls that overlap on a GPU race and must be ordered by the caller.
 * No run statistics are recorded.  Asynchronous on the context's stream. */
bs_status bs_crmsd_all_vs_all_device_noalloc(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                             size_t n_atoms, bs_method method, void* d_out_upper,
                                             int out_f64);

/* Streams the pair space in tile x tile blocks (e.g. 4096) and copies each
 * finished block of the packed triangle to the host while the next one is
 * computed; out_upper is a host f32 buffer of n(n-1)/2 values. */
bs_status bs_crmsd_all_vs_all_streamed(bs_context* ctx, const double* xyz, size_t n_conf,
                                       size_t n_atoms, bs_method method, size_t tile,
                                       float* out_upper);

/* ------------------------------------------------------- batching */

/* Many independent ensembles (uneven n_conf and n_atoms) in one persistent,
 * graph-captured launch.  out_upper[e] receives ensemble e's packed triangle
 * (f32).  Tile shapes come from the autotuner's cache when present. */
bs_status bs_crmsd_batch(bs_context* ctx, const double* const* xyz, const size_t* n_conf,
                         const size_t* n_atoms, size_t n_ens, bs_method method,
                         float* const* out_upper);

/* The same on device-resident buffers: d_xyz[e] (n_conf[e] x n_atoms[e] x 3
 * f64) and d_out_upper[e] (n_conf[e](n_conf[e]-1)/2 f32) are device
 * pointers.  Asynchronous on the context's stream; the pointer arrays
 * themselves are read before the call returns.  Ensembles with fewer than 2
 * conformations are skipped.  The launch sequence (all preparation steps, one
 * persistent Gram launch over every ensemble's tiles, one FP64 re-evaluation
 * launch) is a HIP graph, built on the first call of a batch shape and
 * replayed after that. */
bs_status bs_crmsd_batch_device(bs_context* ctx, const double* const* d_xyz,
                                const size_t* n_conf, const size_t* n_atoms, size_t n_ens,
                                bs_method method, float* const* d_out_upper);

/* Runs the tile-shape autotuner for `method` (a Gram method) over
 * representative batches -- atom counts up to 64, 128, 256, 512 and beyond,
 * times conformation counts up to 64, 128, 256, 512 and beyond -- timing
 * 32 x 32 and 64 x 64 tiles for each cell, stores the faster shape per cell
 * in `cache_path` (JSON; may be null) and loads the table into the context. */
bs_status bs_autotune(bs_context* ctx, bs_method method, const char* cache_path);
/* Loads a previously written autotune cache into the context. */
bs_status bs_autotune_load(bs_context* ctx, const char* cache_path);

/* ------------------------------------------------------- clustering */

/* Linkage (merging rule) of agglomerative clustering, the closures of
 * bioshell-clustering/src/hierarchical/strategies/mod.rs: single_link,
 * complete_link, average_link, median_link, centroid_link, wards_method. */
typedef enum bs_linkage {
    BS_SINGLE_LINK = 0,
    BS_COMPLETE_LINK = 1,
    BS_AVERAGE_LINK = 2,
    BS_MEDIAN_LINK = 3,
    BS_CENTROID_LINK = 4,
    BS_WARDS_METHOD = 5
} bs_linkage;

/* Row minimum and maximum over j != i of an n x n distance matrix given as
 * its packed f32 upper triangle (distances >= 0).  The row minimum against a
 * cutoff is retrieve_outliers (bioshell-clustering/src/hierarchical/
 * hierarchical.rs:198-217); the row maximum is what medoid_by_min_max
 * minimises (hierarchical.rs:105-131).  n = 1 gives f32::MAX / -f32::MAX. */
bs_status bs_rmsd_row_stats(bs_context* ctx, const float* upper, size_t n, float* row_min,
                            float* row_max);
/* The same on device buffers, asynchronous on the context's stream (n >= 2). */
bs_status bs_rmsd_row_stats_device(bs_context* ctx, const float* d_upper, size_t n,
                                   float* d_row_min, float* d_row_max);
/* For each of the m members (indices into the matrix), the largest distance
 * to the other members: the inner loop of medoid_by_min_max
 * (hierarchical.rs:105-131) for one cluster. */
bs_status bs_rmsd_members_row_max(bs_context* ctx, const float* upper, size_t n,
                                  const uint32_t* members, size_t m, float* row_max);

/* Agglomerative clustering of the n elements of a packed f32 distance
 * triangle, as hierarchical_clustering (hierarchical.rs:22-80) with the
 * matrix of clustering_matrix.rs, on the device: merge k joins the clusters
 * at current matrix indices merge_i[k] < merge_j[k] at distance merge_d[k]
 * (n - 1 merges).  The merged cluster takes index i and the last active
 * cluster moves to index j, so the tree is rebuilt by replaying the merges
 * (the Python mirror does).  Ties, the sizes[j] argument of the merging rule
 * and the sizes kept in place on a move are those of the reference.  The
 * dense matrix needs n^2 * 4 bytes of device memory (n <= 65536). */
bs_status bs_hierarchical_clustering(bs_context* ctx, const float* upper, size_t n,
                                     bs_linkage linkage, uint32_t* merge_i, uint32_t* merge_j,
                                     float* merge_d);
/* The same with the triangle already in HBM (e.g. from
 * bs_crmsd_all_vs_all_device with f32 output); synchronous. */
bs_status bs_hierarchical_clustering_device(bs_context* ctx, const float* d_upper, size_t n,
                                            bs_linkage linkage, uint32_t* merge_i,
                                            uint32_t* merge_j, float* merge_d);

/* Writes a distance matrix (packed f32 upper triangle of n elements) as the
 * tab-separated "label_i<TAB>label_j<TAB>value" lines that BioShell's clust
 * tool reads (bioshell-clustering/src/bin/clust.rs:102-107; the layout of its
 * fixtures tests/test_files/cities.tsv and d5.tsv): one line per pair i < j in
 * row order, or, with both_triangles != 0, per ordered pair i != j (for clust
 * runs without --symmetric).  labels holds n NUL-terminated labels, or NULL
 * for "0".."n-1".  decimals < 0 writes each value in the shortest positional
 * form that reads back to the same f32 (Rust's `{}` for f32); 0..30 writes that
 * many decimals of the exact f32 value.  n_threads <= 0 picks up to 16. */
bs_status bs_write_distance_tsv(const float* upper, size_t n, const char* const* labels,
                                int decimals, int both_triangles, int n_threads,
                                const char* path);

/* ------------------------------------------------------- ensembles from files */

/* A conformational ensemble read from a multi-model structure file: the atom
 * records of the first model and the coordinates of the selected atoms in
 * every model (n_models x n_atoms x 3, f64, the layout every cRMSD entry point
 * takes).  It is what the reference gives for Deposit::from_file
 * (bioshell-pdb/src/deposit.rs:81-85) followed, for every model, by
 * Structure::set_model (structure.rs:277-285) and an atom predicate over
 * Structure::atoms(): models are numbered in order of appearance, atoms come
 * from the first model, and a model whose atom count differs from the first
 * model's is an error (PDBError::WrongAtomsNumberInModel). */
typedef struct bs_ensemble bs_ensemble;

typedef enum bs_format {
    BS_FORMAT_AUTO = 0, /* is_cif_file, then is_pdb_file (load_cif.rs:155-167, load_pdb.rs:225-237) */
    BS_FORMAT_CIF = 1,  /* Deposit::from_cif_reader + PdbAtom::from_cif_data (pdb_atom.rs:130-203) */
    BS_FORMAT_PDB = 2   /* Deposit::from_pdb_reader (load_pdb.rs:43-115) */
} bs_format;

/* Atom predicates of bioshell-pdb/src/pdb_atom_filters.rs. */
typedef enum bs_selection {
    BS_SELECT_ALL = 0,
    BS_SELECT_CA = 1,          /* IsCA (pdb_atom_filters.rs:388-394) */
    BS_SELECT_CB = 2,          /* IsCB (pdb_atom_filters.rs:415-422) */
    BS_SELECT_NOT_HYDROGEN = 3 /* !is_hydrogen (pdb_atom.rs:352-365) */
} bs_selection;

typedef struct bs_ensemble_info {
    size_t n_models;          /* Structure::count_models (structure.rs:255) */
    size_t n_atoms;           /* selected atoms per model */
    size_t n_structure_atoms; /* Structure::count_atoms of the first model */
    size_t n_chains;          /* Structure::count_chains (structure.rs:250-252) */
    size_t n_residues;        /* Structure::count_residues (structure.rs:236-239) */
    int format;               /* BS_FORMAT_CIF or BS_FORMAT_PDB */
} bs_ensemble_info;

/* One atom record (bioshell_pdb::PdbAtom); strings are UTF-8, NUL-terminated
 * and cut at the field size. */
typedef struct bs_atom_record {
    int32_t serial;
    int32_t res_seq;
    double occupancy;
    double temp_factor;
    char name[16];
    char alt_loc[8];
    char res_name[16];
    char chain_id[16];
    char i_code[8];
    char element[16];
    int32_t has_element;
    int32_t hetero;
} bs_atom_record;

/* Reads an ensemble from mmCIF or PDB text (len bytes, not NUL-terminated).
 * chain_id (or NULL) adds the ByChain predicate (pdb_atom_filters.rs:191-199).
 * n_threads <= 0 picks up to 16.  Parse errors return BS_ERR_PARSE. */
bs_status bs_ensemble_parse(const char* text, size_t len, int format, int selection,
                            const char* chain_id, int n_threads, bs_ensemble** out);
/* The same from a file; a name ending in .gz is inflated
 * (bioshell-core/src/io/utils.rs:312-330). */
bs_status bs_ensemble_load(const char* path, int format, int selection, const char* chain_id,
                           int n_threads, bs_ensemble** out);
void bs_ensemble_destroy(bs_ensemble* ens);
bs_status bs_ensemble_get_info(const bs_ensemble* ens, bs_ensemble_info* info);
/* Copies the n_models x n_atoms x 3 coordinates out. */
bs_status bs_ensemble_copy_coords(const bs_ensemble* ens, double* xyz);
bs_status bs_ensemble_get_atom(const bs_ensemble* ens, size_t k, bs_atom_record* out);
/* Copies the coordinates to device memory (d_xyz, n_models x n_atoms x 3 f64)
 * on the context's stream and waits for the copy. */
bs_status bs_ensemble_upload(bs_context* ctx, const bs_ensemble* ens, double* d_xyz);
/* Writes models in PDB format, each with write_pdb's records
 * (bioshell-pdb/src/structure.rs:828-850, lines as impl Display for PdbAtom,
 * pdb_atom.rs:256-278); more than one model is framed by MODEL / ENDMDL
 * records, which Deposit::from_pdb_reader reads back as models.  xyz (or NULL
 * for the ensemble's own coordinates) holds n_models x n_atoms x 3. */
bs_status bs_ensemble_write_pdb(const bs_ensemble* ens, const double* xyz, size_t n_models,
                                const char* path);

/* ------------------------------------------------------- libraries */

/* A device-resident library of n_conf conformations of n_atoms each. */
bs_status bs_library_create(bs_context* ctx, const double* xyz, size_t n_conf,
                            size_t n_atoms, bs_library** out);
/* As bs_context_destroy: a server serving the library keeps it until the
 * server is destroyed. */
void bs_library_destroy(bs_library* lib);
size_t bs_library_size(const bs_library* lib);
size_t bs_library_atoms(const bs_library* lib);

/* For each of n_q query conformations, the k library members of lowest cRMSD
 * (ascending; ties by lower index): idx[q*k + r], rmsd[q*k + r]. */
bs_status bs_query_topk(bs_context* ctx, bs_library* lib, const double* q_xyz, size_t n_q,
                        size_t k, uint32_t* idx, float* rmsd);

/* Full query x library cRMSD matrix (f32, row-major n_q x n_lib). */
bs_status bs_query_all(bs_context* ctx, bs_library* lib, const double* q_xyz, size_t n_q,
                       float* out);

/* ------------------------------------------------------- serving */

/* A continuously batching request loop over device-resident libraries.
 * Requests are admitted while their device memory fits in hbm_budget_bytes and
 * a slot is free; queued queries of one library are merged into one launch. */
bs_status bs_server_create(bs_context* ctx, bs_library* const* libs, size_t n_libs,
                           size_t hbm_budget_bytes, int max_inflight, bs_server** out);
void bs_server_destroy(bs_server* srv);
/* Submits n_q queries against library lib_index.  The queries are copied
 * before the call returns; results go to idx/rmsd (n_q*k each), which must
 * stay valid until the ticket completes or is released.  k beyond the top-k
 * selection's range for that library returns BS_ERR_UNSUPPORTED here, so a
 * request never fails later because of the requests it was merged with.
 * The server keeps each ticket's record (status, latency) until
 * bs_server_release; bs_server_destroy finishes every queued request. */
bs_status bs_server_submit(bs_server* srv, size_t lib_index, const double* q_xyz, size_t n_q,
                           size_t k, uint32_t* idx, float* rmsd, uint64_t* ticket);
/* 0 = done, BS_ERR_NOT_READY = still running, other = the request's error. */
bs_status bs_server_poll(bs_server* srv, uint64_t ticket);
/* Blocks until the ticket completes; returns its status. */
bs_status bs_server_wait(bs_server* srv, uint64_t ticket);
/* Cancels a ticket that has not been admitted yet. */
bs_status bs_server_cancel(bs_server* srv, uint64_t ticket);
/* Latency of a completed ticket (submit .. result visible on host), in ms. */
bs_status bs_server_latency_ms(bs_server* srv, uint64_t ticket, double* ms);
/* Drops a ticket's record.  A finished ticket goes at once.  A queued one is
 * withdrawn.  A running one finishes without writing idx/rmsd (they may be
 * freed as soon as this returns) and its record goes when it ends.  The
 * ticket id is invalid afterwards. */
bs_status bs_server_release(bs_server* srv, uint64_t ticket);
/* Ticket records the server holds (submitted and not yet released). */
size_t bs_server_ticket_count(bs_server* srv);

/* ------------------------------------------------------- synthetic input */

/* Deterministic synthetic ensemble: a random-walk CA chain placed by NeRF
 * (bond 3.8 A, bond angle U(85,150) deg, dihedral U(-180,180] deg; the
 * PI - planar convention of bioshell-builder/src/nerf.rs:6-30), then per
 * conformation iid N(0, s^2) noise with s ~ U(sigma_lo, sigma_hi) (or
 * s = near_sigma with probability near_fraction), a uniform random rotation
 * and a translation U(-50,50)^3.  out_xyz: n_conf x n_atoms x 3.
 * Conformation i depends only on (seed, i), so any subset can be regenerated. */
bs_status bs_generate_ensemble(uint64_t seed, size_t n_conf, size_t n_atoms, double sigma_lo,
                               double sigma_hi, double near_fraction, double near_sigma,
                               double* out_xyz);

#ifdef __cplusplus
}
#endif

#endif /* BIOSHELL_GPU_H */
