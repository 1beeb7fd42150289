/*
 * CPU oracle for the fixed-effects estimation core.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against; it is never linked into or called by the product
 * library. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.
 *
 * Every function restates an algorithm of the reference (tgstavares/fe_gpu) in
 * plain C; each cites the reference file:line it follows.  Integer outputs
 * (group sizes, intersection ids, dof counts, drop decisions) are exact; the
 * floating-point outputs are the reference's formulas evaluated in FP64.
 *
 * Array conventions used throughout:
 *   y      : n doubles
 *   W      : n x K doubles, column-major, leading dimension n
 *   Z      : n x L doubles, column-major, leading dimension n
 *   fe     : n_fe x n int32, dim-major (dim d occupies fe[d*n .. d*n+n-1]),
 *            ids are 1-based; ids < 1 are skipped like the reference kernels do.
 */
#ifndef FE_ORACLE_H
#define FE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Estimation options.  All integers are int64 so that the ctypes mirror is trivial. */
typedef struct {
    int64_t conv;      /* 0: GPU-path conventions (fe_gpu_estimate), 1: CPU fallback (fe_cpu_estimate) */
    int64_t ap_kind;   /* 0: relaxed symmetric AP (fe_gpu_demean.f90), 1: plain forward AP (fe_pipeline.f90:225-277) */
    int64_t max_iter;  /* AP iteration cap */
    double tol;        /* AP tolerance on max |group-mean update| */
} fo_options;

typedef struct {
    int64_t converged;
    int64_t fe_iterations;
    int64_t solver_info;
    int64_t rank;
    int64_t dof_model;
    int64_t dof_fes;
    int64_t dof_resid;
    int64_t is_iv;
    int64_t n_instruments;
    int64_t n_cluster_dims;      /* cluster dims actually used (0 => homoskedastic) */
    int64_t cluster_dims[3];     /* 1-based FE dims, in use order */
    int64_t n_clusters[7];       /* clusters per subset mask 1..7 (0 if unused) */
    int64_t min_single_cluster;
    int64_t psd_fix_applied;
    double rss;
    double tss_total;
    double tss_within;
    double r2;
    double r2_adj;
    double r2_within;
    double f_stat;
    double max_update;
} fo_result;

/* G_d = max id of dim d; returns 1 (reference aborts) if any id < 1.  fe_grouping.f90:11-37 */
int fo_group_sizes(const int32_t* fe, int64_t n_fe, int64_t n, int32_t* G);

/* Cluster ids with the reference's signature (fe_ids interleaved (ld_fe, n_obs), 1-based subset
 * dims).  Restates cluster_id_builder.c:271-314: sort path => 1 + dense rank of the mixed-radix key
 * sum_j (id_j - 1) * prod_{l<j} size_l; hash path (span > 2^62 or a zero size) => ids in order of
 * first appearance.  Status codes follow the reference. */
int fo_build_cluster_ids(const int32_t* fe_ids, int ld_fe, long long n_obs, const int32_t* subset_dims,
                         const int32_t* subset_sizes, int subset_len, int32_t* out_ids, int32_t* out_n_clusters);

/* Same numbering, ids given as one contiguous array per dim (dim-major layout). */
int fo_cluster_ids_dimmajor(const int32_t* fe, int64_t n, const int32_t* subset_dims, const int32_t* subset_sizes,
                            int subset_len, int32_t* out_ids, int32_t* out_n_clusters);

/* Plain forward AP on y and W (CPU fallback).  fe_pipeline.f90:136-150, 225-277.
 * Returns iterations with Fortran DO semantics (max_iter+1 when not converged). */
int64_t fo_demean_forward(double* y, double* W, int64_t n, int64_t K, const int32_t* fe, int64_t n_fe,
                          const int32_t* G, double tol, int64_t max_iter, int64_t* converged, double* max_update);

/* Relaxed symmetric AP on y, W and Z (GPU schedule).  fe_gpu_demean.f90:63-311. */
int64_t fo_demean_relaxed(double* y, double* W, double* Z, int64_t n, int64_t K, int64_t L, const int32_t* fe,
                          int64_t n_fe, const int32_t* G, double tol, int64_t max_iter, int64_t* converged,
                          double* max_update);

/* Upper-Cholesky solve and inverse of an n x n SPD matrix (col-major).  fe_solver.f90:41-74.
 * On return beta solves Q beta = b, Qinv = Q^{-1} (full, symmetric).  Returns LAPACK-style info. */
int fo_chol_solve_and_invert(const double* Q, const double* b, int64_t n, double* beta, double* Qinv);

/* Symmetric eigen-decomposition (cyclic Jacobi) used for the PSD fix; stands in for dsyev.
 * A (n x n col-major, symmetric) -> eigenvalues w (ascending) and eigenvectors V (columns). */
int fo_sym_eig(const double* A, int64_t n, double* w, double* V);

/* Full estimation.  y, W, Z are demeaned in place.  beta/se have K entries, vcov K*K (may be NULL).
 * conv=0 restates fe_gpu_estimate (fe_pipeline.f90:516-1896), conv=1 restates fe_cpu_estimate
 * (fe_pipeline.f90:71-513). */
int fo_estimate(int64_t n, int64_t K, int64_t L, int64_t n_fe, double* y, double* W, double* Z, const int32_t* fe,
                const int32_t* G, const fo_options* opt, const int32_t* cluster_req, int64_t n_cluster_req,
                const int32_t* iv_reg, int64_t n_iv_reg, const int32_t* iv_inst, int64_t n_iv_inst, double* beta,
                double* se, double* vcov, fo_result* res);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int fo_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
