/*
 * fehip.h -- coarse-grained C-ABI of the MI355X estimation core (libfe_gpu_hip.so).
 *
 * The fine-grained drop-in ABI (fe_gpu_runtime.h) mirrors the reference's per-kernel calls.
 * These entry points expose the whole estimation core as a few calls so the work can be laid
 * out for CDNA4: a device-resident panel (FE ids re-ordered by the largest FE dim, segment and
 * run structures built once), a within-transform solved in FE-table space, fused normal
 * equations, cluster-robust sandwich and 2SLS stages.  They replace, as one call, the
 * reference's fe_gpu_estimate (src/core/fe_pipeline.f90:516-1896) with its within transform
 * (src/gpu/fe_gpu_demean.f90:63-311), and are what the serving and batching front ends use.
 *
 * All pointers named d_* are device pointers; everything else is host memory.  Status
 * conventions are those of fe_gpu_runtime.h (0 = success, message via
 * fe_gpu_runtime_get_last_error).
 */
#ifndef FEHIP_H
#define FEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEHIP_MAX_FE 8

typedef struct fehip_panel fehip_panel;

typedef struct {
    int64_t n_obs;
    int32_t n_fe;
    int32_t primary_dim;               /* 0-based FE dim the rows are ordered by */
    int32_t sorted_input;              /* 1 if the caller's rows were already in that order */
    int32_t max_segment;               /* longest primary segment (rows) */
    int32_t group_sizes[FEHIP_MAX_FE]; /* G_d = max id (src/core/fe_grouping.f90:11-37) */
    int64_t n_runs[FEHIP_MAX_FE];      /* runs of equal id inside primary segments (0 for the primary) */
    int64_t device_bytes;              /* device memory held by the panel */
} fehip_panel_info;

/* Build a panel from n_fe device arrays of 1-based int32 FE ids (d_fe_ids is a host array of
 * device pointers).  Ids < 1 are rejected (the reference aborts on them).  The ids are copied;
 * the caller's arrays may be freed afterwards. */
int fehip_panel_create(const int* const* d_fe_ids, int n_fe, int64_t n_obs, fehip_panel** out);
int fehip_panel_destroy(fehip_panel* panel);
int fehip_panel_get_info(const fehip_panel* panel, fehip_panel_info* info);

/* Keep an FP32 copy of n_cols FP64 device columns (caller row order, leading dim ld >= n_obs) with the
 * panel.  A request with opts.demean.precision = 1 (FP32 storage) whose columns lie in a registered block
 * reads the copy in its right-hand-side and final passes, half the bytes of the FP64 columns; its results
 * then carry the columns' FP32 rounding (about 1e-7 relative).  The caller must not change registered
 * columns while the panel lives, or must register them again after a change.  No counterpart in the
 * reference, whose columns are FP64 throughout (src/io/fe_data_io.f90:243-299 widens FP32 files). */
int fehip_panel_register_columns(fehip_panel* panel, const double* d_cols, int n_cols, int64_t ld);

/* Within-transform options. */
typedef struct {
    double tol;           /* stop when max |group-mean update| < tol (the reference's criterion) */
    int32_t max_iter;     /* iteration cap */
    int32_t method;       /* 0: PCG in FE space (the only method of the coarse path; the reference's
                             relaxed AP in row space runs through the fine-grained ABI) */
    uint32_t fe_mask;     /* bit d set => FE dim d is absorbed; 0 means all dims */
    int32_t precision;    /* 0: FP64; 1 (fehip_estimate): within-transformed columns stored in FP32.
                             The FP64 cross products come from the transform's final pass when the
                             request's columns share a column tile (beta is then the FP64 solution);
                             otherwise beta is refined once in FP64 on the residual.  2SLS requests
                             stay FP64 */
    int32_t check_every;  /* convergence read-back period in iterations (0 -> 4) */
    int32_t reserved;
} fehip_demean_opts;

typedef struct {
    int32_t iterations;   /* max over columns */
    int32_t converged;    /* all columns */
    double max_update;    /* max over columns of the final criterion */
    double seconds;       /* device time of the transform */
    int64_t bytes_moved;  /* algorithmic bytes (ids, weights, tables, columns) */
} fehip_demean_stats;

/* In-place M_F on n_cols device columns (column-major, leading dim ld, caller row order).
 * d_weights (optional, n_obs uint8 in caller order) restricts the transform to rows with
 * weight 1; other rows are left untouched. */
int fehip_demean(fehip_panel* panel, double* d_cols, int n_cols, int64_t ld, const uint8_t* d_weights,
                 const fehip_demean_opts* opts, fehip_demean_stats* stats);

/* Estimation options (formula-level choices stay with the host, as in the reference). */
typedef struct {
    fehip_demean_opts demean;
    int32_t n_cluster;            /* cluster FE dims (1-based, <= 3 used, order kept) */
    int32_t cluster_dims[4];
    int32_t n_iv_reg;             /* endogenous regressors (1-based indices into W) */
    const int32_t* iv_reg;
    int32_t n_iv_inst;            /* selected instrument columns (1-based into Z); 0 => all */
    const int32_t* iv_inst;
} fehip_opts;

typedef struct {
    int32_t converged, fe_iterations, solver_info, rank;
    int32_t dof_model, dof_fes, is_iv, n_instruments;
    int64_t dof_resid, n_obs;
    int32_t n_cluster_dims, cluster_dims[3];
    int64_t n_clusters[7];
    int64_t min_single_cluster;
    int32_t psd_fix_applied;
    int32_t fe_refinements;  /* FP64 true-residual products of a solve with the FP32 operator (0 in FP64) */
    double rss, tss_total, tss_within, r2, r2_adj, r2_within, f_stat, max_update;
    double t_demean, t_regression, t_se;  /* seconds */
} fehip_result;

/* Full estimation on device columns y (n), W (n x K), Z (n x L), column-major with leading dim
 * ld, caller row order; optional row weights (uint8).  beta/se (K) and vcov (K x K, may be NULL)
 * are host outputs.  Follows the GPU-path conventions of the reference (relaxed-AP fixed point,
 * zero-variance column filter, Cholesky, IV normal equations, CGM with the nested small-sample
 * scalar and PSD clip). */
int fehip_estimate(fehip_panel* panel, const double* d_y, const double* d_W, int K, const double* d_Z, int L,
                   int64_t ld, const uint8_t* d_weights, const fehip_opts* opts, fehip_result* result, double* beta,
                   double* se, double* vcov);

/* One specification of a batch: the arguments of fehip_estimate in one struct. */
typedef struct {
    const double* d_y;
    const double* d_W;
    const double* d_Z;
    int32_t K, L;
    int64_t ld;
    const uint8_t* d_weights;  /* optional per-request row subsample (uint8, caller order) */
    fehip_opts opts;
    fehip_result* result;      /* host outputs */
    double* beta;
    double* se;
    double* vcov;              /* may be NULL */
} fehip_request;

/* Estimate n_req specifications on one panel.  Requests that absorb the same FE dims are solved
 * together: their columns share every within-transform sweep (up to 32 requests per sweep, each
 * with its own row subsample, Schur preconditioner and CG scalars), then each request takes its
 * own normal-equation and sandwich step.  Each request's outputs match fehip_estimate on that
 * request alone (same stopping rule per column; summation order inside the sweeps may differ in
 * the last bits).  Batching is the multi-specification path that has no counterpart in the
 * reference, which estimates one specification per process run (src/src_main.f90). */
int fehip_estimate_batch(fehip_panel* panel, const fehip_request* reqs, int n_req);

/* Request server over a device-resident panel (no counterpart in the reference, which runs one
 * specification per process).  Worker threads on the panel's device take queued requests in arrival
 * order; each solves as many as fit max_batch and the HBM budget left by the batches in flight as one
 * batch (fehip_estimate_batch semantics), on its own stream, so requests that arrive while a batch is
 * on the device start as soon as memory allows and overlap with it.  The request's device inputs and
 * host outputs must stay valid until fehip_server_wait returns for its ticket; the option arrays
 * (iv_reg, iv_inst) are copied at submit. */
typedef struct fehip_server fehip_server;

typedef struct {
    int64_t hbm_budget;  /* device bytes the batches in flight may hold; 0 -> 80% of free memory */
    int32_t max_batch;   /* requests per batch; 0 -> 32 */
    int32_t workers;     /* batches in flight at once (one host thread and stream each); 0 -> 2 */
    int32_t window_us;   /* a partial batch waits up to this long after its oldest request's submit for
                            more requests; 0 -> 2000, < 0 -> no wait */
    int32_t reserved;
} fehip_server_opts;

typedef struct {
    int64_t submitted, completed, failed, batches, queued;
    int64_t budget_bytes;
    double mean_batch;   /* requests per batch */
    double busy_s;       /* worker time spent solving */
    double span_s;       /* first batch start to last batch end */
    double p50_s, p99_s, max_s;  /* submit-to-completion latency of completed requests */
} fehip_server_stats;

int fehip_server_create(fehip_panel* panel, const fehip_server_opts* opts, fehip_server** out);
int fehip_server_submit(fehip_server* server, const fehip_request* req, int64_t* ticket);
/* Blocks until the ticket's request is done; returns its status (message via get_last_error). */
int fehip_server_wait(fehip_server* server, int64_t ticket, double* latency_s);
int fehip_server_get_stats(fehip_server* server, fehip_server_stats* stats);
int fehip_server_reset_stats(fehip_server* server);
/* Finishes every queued request, then stops the worker. */
int fehip_server_destroy(fehip_server* server);

/* The estimated fixed effects of a fitted regression: replaces reconstruct_and_save_fixed_effects
 * (src/core/fe_pipeline.f90:1619-1715), whose alternating projection over r = y - W beta converges to
 * per-dim values that each have row-weighted mean zero; here the FE-space solve of r gives the same
 * fixed point.  For every dim in opts->fe_mask (0 = all), in ascending dim order, G_d entries are
 * written at the running offset: d_values (level id - 1 indexes the dim's block; 0 for levels without
 * rows) and d_counts (rows per level; the reference's output files list the levels with rows).  beta is
 * host memory (K doubles); opts->tol and opts->max_iter bound the solve. */
int fehip_fe_values(fehip_panel* panel, const double* d_y, const double* d_W, int K, int64_t ld, const double* beta,
                    const fehip_demean_opts* opts, double* d_values, int32_t* d_counts, fehip_demean_stats* stats);

/* ---- formula design expansion on the device (replaces build_formula_design_matrix,
 * src/src_main.f90:595-736, which builds the expanded design on the host) ----
 * fehip_design_levels: the levels of a categorical column as collect_levels_for_column (:1389-1432)
 * finds them: the first 40 distinct values within 1e-6 of an integer, in row order, sorted ascending;
 * levels receives up to 40 values (host memory), truncated = 1 when the list filled up (the reference
 * warns then).  fehip_design_expand: output column j = product over its factors of (categorical:
 * nint(x) == level ? 1 : 0; continuous: x), x read from raw column col of d_raw (n_raw columns,
 * leading dim ld_raw); the caller lists the columns in the reference's order (a categorical term: its
 * levels but the base, ascending; an interaction: the factors' blocks multiplied out, the first factor
 * outermost). */
#define FEHIP_DESIGN_MAX_FACTORS 4
typedef struct {
    int32_t n_factors;
    int32_t col[FEHIP_DESIGN_MAX_FACTORS];
    int32_t level[FEHIP_DESIGN_MAX_FACTORS];
    int32_t categorical[FEHIP_DESIGN_MAX_FACTORS];
} fehip_design_column;

int fehip_design_levels(const double* d_col, int64_t n, int32_t* levels, int32_t* n_levels, int32_t* truncated);
int fehip_design_expand(const double* d_raw, int64_t ld_raw, int n_raw, int64_t n, const fehip_design_column* cols,
                        int n_cols, double* d_out, int64_t ld_out);

/* ---- `.bin` dataset files (the reference's format) straight into device memory ----
 * Replaces load_dataset_from_file (src/io/fe_data_io.f90:23-66, header rules :78-202, size formula
 * :342-358) followed by the upload of src/gpu/fe_gpu_data.f90:41-120.  The header layout (with or
 * without L) and the value precision are decided from the file size exactly as the reference does;
 * metadata_bytes counts the trailing META block (names; not read here). */
typedef struct {
    int64_t n_obs;
    int32_t n_regressors;       /* K */
    int32_t n_instruments;      /* L (0 in the old layout) */
    int32_t n_fe;
    int32_t has_cluster;
    int32_t has_weights;
    int32_t precision;          /* 0 = FP64 values, 1 = FP32 values */
    int32_t new_format;         /* 1: 28-byte header with L; 0: 24-byte header without it */
    int32_t has_precision_flag; /* an int32 flag word follows the header */
    int64_t data_offset;        /* byte offset of y */
    int64_t metadata_bytes;     /* bytes after the data blocks */
} fehip_bin_header;

int fehip_bin_read_header(const char* path, fehip_bin_header* out);
/* Streams the blocks into device buffers: d_y (n), d_W (n x K, col-major), d_Z (n x L), d_fe_ids
 * (n_fe x n int32, dim-major), d_cluster (n int32), d_weights (n).  Values land in FP64 (an FP32 file
 * is widened on the device, as the reference widens it on the host).  A null destination skips its
 * block.  The file is read in 64 MiB chunks into two pinned buffers, each chunk copied while the
 * next one is read. */
int fehip_bin_load(const char* path, const fehip_bin_header* h, double* d_y, double* d_W, double* d_Z,
                   int32_t* d_fe_ids, int32_t* d_cluster, double* d_weights);

/* Deterministic synthetic worker-firm-time panel generated on the device (the 150M-row benchmark
 * panel does not fit a host round trip).  Rows are ordered by worker then period.  Writes
 * d_fe (3 x n int32, dim-major: worker, firm, period), d_y (n) and d_W (n x 8, col-major). */
int fehip_synth_worker_firm(int64_t n_workers, int32_t n_periods, int32_t n_firms, double switch_p, uint64_t seed,
                            int32_t* d_fe, double* d_y, double* d_W);

/* Deterministic wide 2SLS panel generated on the device: n_ind x n_periods rows (ordered by individual
 * then period), firms moving with probability switch_p, time = period.  Writes d_fe (3 x n int32,
 * dim-major: individual, firm, period), d_y (n), d_W (n x 252: xend, 7 cat8 levels, x1..x4, 29 ind
 * levels, 7 msp levels, 203 ind x msp cells, x1*x2; columns 1-8 are endogenous) and d_Z (n x 21: z1,
 * z2, 19 zcat20 levels), column-major. */
int fehip_synth_iv_wide(int64_t n_ind, int32_t n_periods, int32_t n_firms, double switch_p, uint64_t seed,
                        int32_t* d_fe, double* d_y, double* d_W, double* d_Z);

/* Deterministic row subsample on the device: d_mask[i] = 1 with probability p (a counter-based hash of
 * seed and i), else 0; n uint8 values (a request's row mask of the serving workload). */
int fehip_synth_mask(int64_t n, double p, uint64_t seed, uint8_t* d_mask);

/* Opt-in per-kernel timing with HIP events on the stream each kernel is launched on.  Kernel ids:
 * 0 apply_iter (the CG operator product, the dominant kernel), 1 apply_data (right-hand side),
 * 2 finalize (M_F columns), 3 reduce_runs, 4 gram, 5 segsum (cluster score sums), 6 apply_iter32 (the
 * FP32 operator product of a mixed-precision solve), 7 apply_resid (its FP64 true-residual products),
 * 8 couplings (the one-time products that build the run-level operator's small-dim couplings).
 * bytes/flops accumulate the algorithmic work of every timed launch. */
int fehip_kstats_enable(int on);
int fehip_kstats_reset(void);
int fehip_kstats_get(int id, char* name, int name_len, int64_t* launches, double* total_ms, double* bytes,
                     double* flops);

/* Plain device-to-device streaming copy of bytes (a multiple of 16) by a kernel: the reference
 * workload for calibrating hardware byte counters (reads and writes exactly bytes each). */
int fehip_device_copy(const void* d_src, void* d_dst, int64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* FEHIP_H */
