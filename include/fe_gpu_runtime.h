/*
 * fe_gpu_runtime.h -- the drop-in C-ABI of the fixed-effects estimation core,
 * implemented by libfe_gpu_hip.so with hand-written HIP kernels for gfx950 (MI355X).
 *
 * The signatures, argument meaning, layouts and status conventions are those the
 * reference's Fortran host binds through ISO_C_BINDING (src/gpu/fe_gpu_runtime.h:23-131
 * of tgstavares/fe_gpu; bind(C) interfaces in src/gpu/fe_gpu_runtime.f90:59-154,
 * src/gpu/fe_gpu_linalg.f90:18-90, src/gpu/fe_gpu_demean.f90:14-59), so the unchanged
 * host links against this library instead of the CUDA backend.
 *
 * Conventions (all entry points):
 *   - return 0 on success; nonzero is a HIP error code or 1, and the message is kept in a
 *     thread-local buffer read by fe_gpu_runtime_get_last_error().  Unlike the reference,
 *     the linalg entry points also set the message on failure.
 *   - device pointers are passed by value; the caller allocates and frees every buffer
 *     through fe_gpu_runtime_malloc/free.
 *   - matrices are column-major with an explicit leading dimension; FE ids and cluster ids
 *     are int32 and 1-based (ids < 1 are skipped); FE group tables are group-major
 *     (entry [g * n_cols + k]).
 *   - every entry point is ordered on the default stream; those returning a host value
 *     (absmax, dot, build_multi_cluster_ids) and cluster_scores synchronise.
 */
#ifndef FE_GPU_RUNTIME_H
#define FE_GPU_RUNTIME_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef FE_GPU_NAME_LEN
#define FE_GPU_NAME_LEN 256
#endif

/* Device description filled by fe_gpu_runtime_init.  Field order is the one of the C header
 * (src/gpu/fe_gpu_runtime.h:14-21), the ABI of record; the Fortran mirror
 * (src/gpu/fe_gpu_runtime.f90:11-19) orders the ints first, which only garbles its log line. */
typedef struct {
    int device_id;
    size_t total_global_mem;
    int multiprocessor_count;
    int major;
    int minor;
    char name[FE_GPU_NAME_LEN];
} fe_gpu_device_info;

/* ---- runtime (src/gpu/fe_gpu_runtime.h:23-33; CUDA impl fe_gpu_runtime.cu:449-571) ---- */

/* 0 when a HIP device is present (the reference's "available" value). */
int fe_gpu_runtime_is_available(void);
/* Select device_id (out-of-range -> 0) and describe it. */
int fe_gpu_runtime_init(int device_id, fe_gpu_device_info* info);
/* Release library-held device memory and reset the device. */
int fe_gpu_runtime_shutdown(void);
/* bytes == 0 yields *ptr = NULL and success. */
int fe_gpu_runtime_malloc(void** ptr, size_t bytes);
/* Also drops any device structures the library cached for this pointer. */
int fe_gpu_runtime_free(void* ptr);
int fe_gpu_runtime_memcpy_htod(void* dst, const void* src, size_t bytes);
int fe_gpu_runtime_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int fe_gpu_runtime_memcpy_dtod(void* dst, const void* src, size_t bytes);
int fe_gpu_runtime_memset(void* ptr, int value, size_t bytes);
/* Copies the last message into buffer (NUL-terminated); 1 if buffer is NULL or length 0. */
int fe_gpu_runtime_get_last_error(char* buffer, size_t length);
int fe_gpu_runtime_clear_error(void);

/* ---- dense algebra (src/gpu/fe_gpu_runtime.h:35-80; CUDA impl fe_gpu_linalg.cu:81-231) ---- */

/* No handle is needed (own kernels); kept for the host's init/shutdown calls. */
int fe_gpu_linalg_init(void);
int fe_gpu_linalg_shutdown(void);
/* Q(upper) = alpha * W^T W + beta * Q;  W n_rows x n_cols (ld ldW), Q n_cols x n_cols (ld n_cols). */
int fe_gpu_syrk(int n_rows, int n_cols, double alpha, const double* W, int ldW, double beta, double* Q);
/* b = alpha * W^T y + beta * b. */
int fe_gpu_gemv(int n_rows, int n_cols, double alpha, const double* W, int ldW, const double* y, double beta,
                double* b);
/* residual = y - W beta (nothing written when n_cols == 0, as in the reference). */
int fe_gpu_residual(int n_rows, int n_cols, const double* W, int ldW, const double* beta, const double* y,
                    double* residual);
/* *result (host) = x . y */
int fe_gpu_dot(int n_rows, const double* x, const double* y, double* result);
/* C = alpha * op(A) op(B) + beta * C, op = 'N' or 'T' (column-major BLAS semantics). */
int fe_gpu_gemm(char transA, char transB, int m, int n, int k, double alpha, const double* A, int ldA,
                const double* B, int ldB, double beta, double* C, int ldC);
/* scores[c + j * n_clusters] += residual_i * W[i + j * ldW] for rows with 1 <= id_i <= n_clusters;
 * the caller zeroes scores first. */
int fe_gpu_cluster_scores(const double* residual, const double* W, const int* cluster_ids, int n_rows, int n_cols,
                          int ldW, int n_clusters, double* scores);
/* meat(upper) = scores^T scores. */
int fe_gpu_cluster_meat(int n_clusters, int n_cols, const double* scores, int ldScores, double* meat);
/* fe_gpu_syrk without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and result as fe_gpu_syrk; its split-K partials live in a
 * fixed 32 MiB device array held by each GPU the library runs on until the process ends, so
 * n_cols * n_cols * 8 must not exceed 32 MiB (n_cols <= 2048; larger is an error).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other host threads or streams) race. */
int fe_gpu_syrk_noalloc(int n_rows, int n_cols, double alpha, const double* W, int ldW, double beta, double* Q);

/* ---- FE projection pieces (src/gpu/fe_gpu_runtime.h:82-114; CUDA impl fe_gpu_runtime.cu:573-641) ---- */

/* Adds per-group sums of y, W, Z and counts into caller-zeroed group-major tables. */
int fe_gpu_fe_accumulate(const double* y, const double* W, const double* Z, const int* fe_ids, size_t n_obs,
                         int n_groups, int n_reg, int n_inst, size_t leading_dim, double* group_sum_y,
                         double* group_sum_W, double* group_sum_Z, int* group_counts);
/* In place: sums -> means (0 for empty groups). */
int fe_gpu_fe_compute_means(double* group_sum_y, double* group_sum_W, double* group_sum_Z, const int* group_counts,
                            int n_groups, int n_reg, int n_inst);
/* x_i -= mean[id_i] for y, W, Z (relaxation is accepted and unused, as in the reference). */
int fe_gpu_fe_subtract(double* y, double* W, double* Z, const int* fe_ids, size_t n_obs, int n_reg, int n_inst,
                       size_t leading_dim, const double* group_mean_y, const double* group_mean_W,
                       const double* group_mean_Z, double relaxation);
/* mixed = prev + relaxation * (mean - prev); prev = mean = mixed. */
int fe_gpu_mix_means(double* mean, double* prev, size_t n, double relaxation);
/* y += alpha * x (exported for completeness; the host never binds it). */
int fe_gpu_axpy(int n, double alpha, const double* x, double* y);

/* ---- misc (src/gpu/fe_gpu_runtime.h:116-131; CUDA impl fe_gpu_runtime.cu:643-886) ---- */

/* *out (host) = max |data[i]|; n <= 0 is an error with *out = 0. */
int fe_gpu_absmax(const double* data, long long n, double* out);
/* dst[:, dest_offset + j] = src[:, indices[j] - 1] for n_rows rows; indices is a HOST array. */
int fe_gpu_copy_columns(const double* src, int ld_src, const int* indices, int n_indices, int n_rows, double* dst,
                        int ld_dst, int dest_offset);
/* Intersection ids: key_i = sum_d (id_d,i - 1) * strides[d] over ids > 0; out_ids = 1 + dense rank
 * of key_i among the distinct keys (the CPU builder's numbering, src/core/cluster_id_builder.c:146-194).
 * fe_ptrs is a HOST array of device pointers, strides a HOST array, out_n_clusters a HOST int. */
int fe_gpu_build_multi_cluster_ids(const void* const* fe_ptrs, const unsigned long long* strides, int n_dims,
                                   long long n_obs, int* out_ids, int* out_n_clusters);

#ifdef __cplusplus
}
#endif

#endif /* FE_GPU_RUNTIME_H */
