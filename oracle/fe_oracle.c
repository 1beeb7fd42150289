/*
 * CPU oracle for the fixed-effects estimation core (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in plain C, the algorithms of the reference tgstavares/fe_gpu that the
 * HIP product path replaces.  Each function cites the reference file:line it follows.
 * The product library never links or calls this file; tests compare against it.
 */
#include "fe_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int fo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Group sizes.  fe_grouping.f90:11-37: G_d = max id of the dim, abort on id < 1. */
int fo_group_sizes(const int32_t* fe, int64_t n_fe, int64_t n, int32_t* G) {
    for (int64_t d = 0; d < n_fe; ++d) {
        const int32_t* ids = fe + d * n;
        int32_t mx = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (ids[i] < 1) return 1;
            if (ids[i] > mx) mx = ids[i];
        }
        G[d] = mx;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Intersection ids.  cluster_id_builder.c:271-314 (entry), :259-269 (sort vs hash decision),
 * :146-194 (ids = 1 + dense rank of the key; depends only on key order, so any correct sort
 * reproduces it), :202-257 (hash fallback: ids in order of first appearance). */

typedef struct {
    const int32_t* base;
    int64_t stride; /* element stride between consecutive observations */
    int valid;
} fo_dim_view;

/* LSD radix sort of (key, index) pairs with 16-bit digits; passes whose digit is constant
 * across all keys are skipped. */
static int fo_sort_pairs(uint64_t* keys, int64_t* idx, int64_t n) {
    uint64_t* tk = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    int64_t* ti = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    size_t* hist = (size_t*)malloc(65536 * sizeof(size_t));
    if (!tk || !ti || !hist) {
        free(tk);
        free(ti);
        free(hist);
        return 6;
    }
    uint64_t* sk = keys;
    int64_t* si = idx;
    uint64_t* dk = tk;
    int64_t* di = ti;
    for (int pass = 0; pass < 4; ++pass) {
        const unsigned shift = (unsigned)(16 * pass);
        memset(hist, 0, 65536 * sizeof(size_t));
        for (int64_t i = 0; i < n; ++i) hist[(sk[i] >> shift) & 0xFFFFu]++;
        int trivial = 0;
        for (size_t b = 0; b < 65536; ++b) {
            if (hist[b] == (size_t)n) trivial = 1;
        }
        if (trivial) continue;
        size_t run = 0;
        for (size_t b = 0; b < 65536; ++b) {
            size_t c = hist[b];
            hist[b] = run;
            run += c;
        }
        for (int64_t i = 0; i < n; ++i) {
            size_t pos = hist[(sk[i] >> shift) & 0xFFFFu]++;
            dk[pos] = sk[i];
            di[pos] = si[i];
        }
        uint64_t* t1 = sk;
        sk = dk;
        dk = t1;
        int64_t* t2 = si;
        si = di;
        di = t2;
    }
    if (sk != keys) {
        memcpy(keys, sk, (size_t)n * sizeof(uint64_t));
        memcpy(idx, si, (size_t)n * sizeof(int64_t));
    }
    free(tk);
    free(ti);
    free(hist);
    return 0;
}

static uint64_t fo_mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static int fo_cluster_ids_views(const fo_dim_view* dims, const int32_t* sizes, int len, int64_t n, int32_t* out,
                                int32_t* n_out) {
    /* :259-269 use_sort_encoding */
    uint64_t span = 1;
    int sort_ok = 1;
    for (int j = 0; j < len; ++j) {
        uint64_t s = (uint64_t)sizes[j];
        if (s == 0 || s > UINT64_MAX / span) {
            sort_ok = 0;
            break;
        }
        span *= s;
        if (span > (1ULL << 62)) {
            sort_ok = 0;
            break;
        }
    }
    if (sort_ok) {
        uint64_t mult[4];
        uint64_t stride = 1;
        for (int j = 0; j < len; ++j) {
            mult[j] = stride;
            stride *= (uint64_t)sizes[j];
        }
        uint64_t* keys = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
        int64_t* idx = (int64_t*)malloc((size_t)n * sizeof(int64_t));
        if (!keys || !idx) {
            free(keys);
            free(idx);
            return 6;
        }
        for (int64_t i = 0; i < n; ++i) {
            uint64_t k = 0;
            for (int j = 0; j < len; ++j) {
                if (!dims[j].valid) continue;
                int32_t v = dims[j].base[i * dims[j].stride];
                k += mult[j] * (uint64_t)(v - 1);
            }
            keys[i] = k;
            idx[i] = i;
        }
        int st = fo_sort_pairs(keys, idx, n);
        if (st != 0) {
            free(keys);
            free(idx);
            return st;
        }
        int32_t id = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (i == 0 || keys[i] != keys[i - 1]) {
                if (id == INT32_MAX) {
                    free(keys);
                    free(idx);
                    return 5;
                }
                ++id;
            }
            out[idx[i]] = id;
        }
        *n_out = id;
        free(keys);
        free(idx);
        return 0;
    }
    /* hash path: first-appearance numbering */
    for (int j = 0; j < len; ++j) {
        if (!dims[j].valid) return 4;
    }
    size_t cap = 16;
    while (cap < (size_t)n * 2 + 1) cap <<= 1;
    int64_t* slot_row = (int64_t*)malloc(cap * sizeof(int64_t));
    int32_t* slot_id = (int32_t*)malloc(cap * sizeof(int32_t));
    if (!slot_row || !slot_id) {
        free(slot_row);
        free(slot_id);
        return 6;
    }
    for (size_t s = 0; s < cap; ++s) slot_row[s] = -1;
    int32_t next = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t h = 0x12345678ULL;
        for (int j = 0; j < len; ++j) h = fo_mix64(h ^ (uint64_t)(uint32_t)dims[j].base[i * dims[j].stride]);
        size_t pos = (size_t)(h & (cap - 1));
        for (;;) {
            int64_t r = slot_row[pos];
            if (r < 0) {
                if (next == INT32_MAX) {
                    free(slot_row);
                    free(slot_id);
                    return 5;
                }
                slot_row[pos] = i;
                slot_id[pos] = ++next;
                out[i] = next;
                break;
            }
            int same = 1;
            for (int j = 0; j < len; ++j) {
                if (dims[j].base[r * dims[j].stride] != dims[j].base[i * dims[j].stride]) {
                    same = 0;
                    break;
                }
            }
            if (same) {
                out[i] = slot_id[pos];
                break;
            }
            pos = (pos + 1) & (cap - 1);
        }
    }
    *n_out = next;
    free(slot_row);
    free(slot_id);
    return 0;
}

int fo_build_cluster_ids(const int32_t* fe_ids, int ld_fe, long long n_obs, const int32_t* subset_dims,
                         const int32_t* subset_sizes, int subset_len, int32_t* out_ids, int32_t* out_n_clusters) {
    if (!fe_ids || !subset_dims || !subset_sizes || !out_ids || !out_n_clusters) return 1;
    if (ld_fe <= 0 || n_obs <= 0) return 2;
    if (subset_len <= 0 || subset_len > 4) return 3;
    fo_dim_view v[4];
    for (int j = 0; j < subset_len; ++j) {
        int d = subset_dims[j];
        v[j].valid = (d >= 1 && d <= ld_fe);
        v[j].base = fe_ids + (v[j].valid ? (d - 1) : 0);
        v[j].stride = ld_fe;
    }
    return fo_cluster_ids_views(v, subset_sizes, subset_len, (int64_t)n_obs, out_ids, out_n_clusters);
}

int fo_cluster_ids_dimmajor(const int32_t* fe, int64_t n, const int32_t* subset_dims, const int32_t* subset_sizes,
                            int subset_len, int32_t* out_ids, int32_t* out_n_clusters) {
    if (!fe || !subset_dims || !subset_sizes || !out_ids || !out_n_clusters) return 1;
    if (n <= 0) return 2;
    if (subset_len <= 0 || subset_len > 4) return 3;
    fo_dim_view v[4];
    for (int j = 0; j < subset_len; ++j) {
        v[j].valid = subset_dims[j] >= 1;
        v[j].base = fe + (int64_t)(subset_dims[j] - 1) * n;
        v[j].stride = 1;
    }
    return fo_cluster_ids_views(v, subset_sizes, subset_len, n, out_ids, out_n_clusters);
}

/* ------------------------------------------------------------------------- */
/* Plain forward AP (CPU fallback).  fe_pipeline.f90:136-150 (loop), :225-277 (one dim). */
static void fo_forward_one_dim(double* y, double* W, int64_t n, int64_t K, const int32_t* ids, int32_t G,
                               double* max_change) {
    if (G <= 0) return;
    const int64_t Gl = G;
    const int64_t ncol = 1 + K;
    double* mean = (double*)calloc((size_t)(Gl * ncol), sizeof(double)); /* [col][g] */
    int64_t* cnt = (int64_t*)calloc((size_t)Gl, sizeof(int64_t));
#ifdef _OPENMP
    int nt = omp_get_max_threads();
#else
    int nt = 1;
#endif
    if (nt > 1 && n >= 65536) {
        double* part = (double*)calloc((size_t)nt * (size_t)(Gl * ncol), sizeof(double));
        int64_t* pc = (int64_t*)calloc((size_t)nt * (size_t)Gl, sizeof(int64_t));
#pragma omp parallel num_threads(nt)
        {
#ifdef _OPENMP
            int t = omp_get_thread_num();
#else
            int t = 0;
#endif
            double* mp = part + (size_t)t * (size_t)(Gl * ncol);
            int64_t* cp = pc + (size_t)t * (size_t)Gl;
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                int32_t g = ids[i];
                if (g < 1 || g > G) continue;
                --g;
                mp[g] += y[i];
                cp[g] += 1;
                for (int64_t j = 0; j < K; ++j) mp[(1 + j) * Gl + g] += W[j * n + i];
            }
        }
        for (int t = 0; t < nt; ++t) {
            const double* mp = part + (size_t)t * (size_t)(Gl * ncol);
            const int64_t* cp = pc + (size_t)t * (size_t)Gl;
            for (int64_t e = 0; e < Gl * ncol; ++e) mean[e] += mp[e];
            for (int64_t g = 0; g < Gl; ++g) cnt[g] += cp[g];
        }
        free(part);
        free(pc);
    } else {
        for (int64_t i = 0; i < n; ++i) {
            int32_t g = ids[i];
            if (g < 1 || g > G) continue;
            --g;
            mean[g] += y[i];
            cnt[g] += 1;
            for (int64_t j = 0; j < K; ++j) mean[(1 + j) * Gl + g] += W[j * n + i];
        }
    }
    for (int64_t g = 0; g < Gl; ++g) {
        for (int64_t c = 0; c < ncol; ++c) {
            if (cnt[g] > 0)
                mean[c * Gl + g] = mean[c * Gl + g] / (double)cnt[g];
            else
                mean[c * Gl + g] = 0.0;
        }
    }
    double mc = *max_change;
#pragma omp parallel for schedule(static) reduction(max : mc) if (n >= 65536)
    for (int64_t i = 0; i < n; ++i) {
        int32_t g = ids[i];
        if (g < 1 || g > G) continue;
        --g;
        double m = mean[g];
        y[i] -= m;
        if (fabs(m) > mc) mc = fabs(m);
        for (int64_t j = 0; j < K; ++j) {
            double mw = mean[(1 + j) * Gl + g];
            W[j * n + i] -= mw;
            if (fabs(mw) > mc) mc = fabs(mw);
        }
    }
    *max_change = mc;
    free(mean);
    free(cnt);
}

int64_t fo_demean_forward(double* y, double* W, int64_t n, int64_t K, const int32_t* fe, int64_t n_fe,
                          const int32_t* G, double tol, int64_t max_iter, int64_t* converged, double* max_update) {
    double mu = 0.0;
    if (n_fe <= 0) {
        *converged = 1;
        *max_update = 0.0;
        return 0;
    }
    int64_t iter;
    for (iter = 1; iter <= max_iter; ++iter) {
        mu = 0.0;
        for (int64_t d = 0; d < n_fe; ++d) fo_forward_one_dim(y, W, n, K, fe + d * n, G[d], &mu);
        if (mu < tol) break;
    }
    *converged = (mu < tol);
    *max_update = mu;
    return iter;
}

/* ------------------------------------------------------------------------- */
/* Relaxed symmetric AP on y, W, Z.  fe_gpu_demean.f90:63-311.
 * Per projection: zero tables (:150-155), accumulate sums+counts (:157-162, kernel
 * fe_gpu_runtime.cu:72-111), means by reciprocal (:183-211), relaxation mix (:361-380),
 * subtract (:213-253), change = max|mixed mean| (:196-211), relaxation update (:213-218). */
typedef struct {
    int64_t n, K, L;
    double *y, *W, *Z;
} fo_cols;

/* The tables are kept column by column here ([col][g]; y, then W, then Z): each column's group sums run
 * over the rows in row order on their own, so the columns are independent work for the threads and the
 * results do not depend on the thread count.  The table layout is internal to this restatement (the
 * reference's group-major layout changes neither a sum's order nor any value). */
static double fo_relaxed_project(const fo_cols* c, const int32_t* ids, int32_t G, double r, double* my, double* mW,
                                 double* mZ, int32_t* cnt, double* py, double* pW, double* pZ, double* max_update) {
    const int64_t n = c->n, K = c->K, L = c->L;
    const int64_t ncol = 1 + K + L;
    memset(cnt, 0, (size_t)G * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) {
        int32_t g = ids[i] - 1;
        if (g < 0) continue;
        cnt[g] += 1;
    }
    double cd = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(max : cd) if (n * ncol >= (1 << 18))
    for (int64_t col = 0; col < ncol; ++col) {
        double* x = col == 0 ? c->y : (col <= K ? c->W + (col - 1) * n : c->Z + (col - 1 - K) * n);
        double* m = col == 0 ? my : (col <= K ? mW + (col - 1) * (int64_t)G : mZ + (col - 1 - K) * (int64_t)G);
        double* pv = col == 0 ? py : (col <= K ? pW + (col - 1) * (int64_t)G : pZ + (col - 1 - K) * (int64_t)G);
        memset(m, 0, (size_t)G * sizeof(double));
        for (int64_t i = 0; i < n; ++i) {
            int32_t g = ids[i] - 1;
            if (g < 0) continue;
            m[g] += x[i];
        }
        for (int32_t g = 0; g < G; ++g) {
            double inv = cnt[g] > 0 ? 1.0 / (double)cnt[g] : 0.0;
            m[g] *= inv;
            /* mix: mixed = prev + r (m - prev); prev = mixed; mean = mixed */
            double v = pv[g] + r * (m[g] - pv[g]);
            pv[g] = v;
            m[g] = v;
            if (fabs(v) > cd) cd = fabs(v);
        }
        for (int64_t i = 0; i < n; ++i) {
            int32_t g = ids[i] - 1;
            if (g < 0) continue;
            x[i] -= m[g];
        }
    }
    /* change = max |mixed mean| over the y, W and Z tables */
    if (cd > *max_update) *max_update = cd;
    return cd;
}

int64_t fo_demean_relaxed(double* y, double* W, double* Z, int64_t n, int64_t K, int64_t L, const int32_t* fe,
                          int64_t n_fe, const int32_t* G, double tol, int64_t max_iter, int64_t* converged,
                          double* max_update) {
    *converged = 1;
    *max_update = 0.0;
    if (n_fe <= 0 || n <= 0) return 0;
    /* dims in ascending group count, by the reference's exchange sort (:104-113) */
    int64_t order[64];
    for (int64_t d = 0; d < n_fe; ++d) order[d] = d;
    for (int64_t a = 0; a < n_fe - 1; ++a)
        for (int64_t b = a + 1; b < n_fe; ++b)
            if (G[order[b]] < G[order[a]]) {
                int64_t t = order[a];
                order[a] = order[b];
                order[b] = t;
            }
    int32_t Gmax = 1;
    for (int64_t d = 0; d < n_fe; ++d)
        if (G[d] > Gmax) Gmax = G[d];
    double* my = (double*)malloc((size_t)Gmax * sizeof(double));
    double* mW = (double*)malloc((size_t)Gmax * (size_t)(K > 0 ? K : 1) * sizeof(double));
    double* mZ = (double*)malloc((size_t)Gmax * (size_t)(L > 0 ? L : 1) * sizeof(double));
    int32_t* cnt = (int32_t*)malloc((size_t)Gmax * sizeof(int32_t));
    double* relax = (double*)malloc((size_t)n_fe * sizeof(double));
    double* prev_change = (double*)malloc((size_t)n_fe * sizeof(double));
    double** py = (double**)calloc((size_t)n_fe, sizeof(double*));
    double** pW = (double**)calloc((size_t)n_fe, sizeof(double*));
    double** pZ = (double**)calloc((size_t)n_fe, sizeof(double*));
    for (int64_t d = 0; d < n_fe; ++d) {
        relax[d] = 1.5;
        prev_change[d] = DBL_MAX;
        py[d] = (double*)calloc((size_t)(G[d] > 0 ? G[d] : 1), sizeof(double));
        pW[d] = (double*)calloc((size_t)(G[d] > 0 ? G[d] : 1) * (size_t)(K > 0 ? K : 1), sizeof(double));
        pZ[d] = (double*)calloc((size_t)(G[d] > 0 ? G[d] : 1) * (size_t)(L > 0 ? L : 1), sizeof(double));
    }
    fo_cols c = {n, K, L, y, W, Z};
    int hit = 0;
    int64_t iterations = 0;
    double mu = 0.0;
    for (int64_t iter = 1; iter <= max_iter; ++iter) {
        mu = 0.0;
        for (int sweep = 0; sweep < 2; ++sweep) {
            for (int64_t pass = 0; pass < 2 * n_fe; ++pass) {
                int64_t s = pass < n_fe ? pass : (2 * n_fe - 1 - pass);
                int64_t d = order[s];
                double cd = fo_relaxed_project(&c, fe + d * n, G[d], relax[d], my, mW, mZ, cnt, py[d], pW[d], pZ[d],
                                               &mu);
                if (cd > prev_change[d] * 1.05) {
                    relax[d] = relax[d] * 0.5;
                    if (relax[d] < 0.4) relax[d] = 0.4;
                } else if (cd < prev_change[d] * 0.8) {
                    relax[d] = relax[d] * 1.10;
                    if (relax[d] > 2.0) relax[d] = 2.0;
                }
                prev_change[d] = cd;
            }
        }
        iterations = iter;
        if (mu < tol) {
            hit = 1;
            break;
        }
    }
    *converged = hit;
    *max_update = mu;
    for (int64_t d = 0; d < n_fe; ++d) {
        free(py[d]);
        free(pW[d]);
        free(pZ[d]);
    }
    free(py);
    free(pW);
    free(pZ);
    free(my);
    free(mW);
    free(mZ);
    free(cnt);
    free(relax);
    free(prev_change);
    return iterations;
}

/* ------------------------------------------------------------------------- */
/* Dense K x K algebra (col-major). */

/* Upper Cholesky A = U^T U, then solve and invert.  fe_solver.f90:41-74 (dpotrf/dpotrs/dpotri 'U'). */
int fo_chol_solve_and_invert(const double* Q, const double* b, int64_t n, double* beta, double* Qinv) {
    double* U = (double*)calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(double));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i <= j; ++i) U[j * n + i] = Q[j * n + i];
    for (int64_t j = 0; j < n; ++j) {
        double s = U[j * n + j];
        for (int64_t k = 0; k < j; ++k) s -= U[j * n + k] * U[j * n + k];
        if (!(s > 0.0)) {
            free(U);
            return (int)(j + 1);
        }
        double ujj = sqrt(s);
        U[j * n + j] = ujj;
        for (int64_t i = j + 1; i < n; ++i) {
            double t = U[i * n + j];
            for (int64_t k = 0; k < j; ++k) t -= U[j * n + k] * U[i * n + k];
            U[i * n + j] = t / ujj;
        }
    }
    /* solve U^T z = b then U x = z */
    for (int64_t i = 0; i < n; ++i) {
        double t = b[i];
        for (int64_t k = 0; k < i; ++k) t -= U[i * n + k] * beta[k];
        beta[i] = t / U[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        double t = beta[i];
        for (int64_t k = i + 1; k < n; ++k) t -= U[k * n + i] * beta[k];
        beta[i] = t / U[i * n + i];
    }
    /* inverse = inv(U) inv(U)^T */
    double* Ui = (double*)calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(double));
    for (int64_t j = 0; j < n; ++j) {
        Ui[j * n + j] = 1.0 / U[j * n + j];
        for (int64_t i = j - 1; i >= 0; --i) {
            double t = 0.0;
            for (int64_t k = i + 1; k <= j; ++k) t += U[k * n + i] * Ui[j * n + k];
            Ui[j * n + i] = -t / U[i * n + i];
        }
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i <= j; ++i) {
            double t = 0.0;
            for (int64_t k = j; k < n; ++k) t += Ui[k * n + i] * Ui[k * n + j];
            Qinv[j * n + i] = t;
            Qinv[i * n + j] = t;
        }
    free(U);
    free(Ui);
    return 0;
}

/* LU with partial pivoting, solve and inverse (CPU fallback: dgesv + dgetri, fe_pipeline.f90:485-512). */
static int fo_lu_solve_and_invert(const double* A, double* b, int64_t n, double* inv) {
    double* LU = (double*)malloc((size_t)(n * n > 0 ? n * n : 1) * sizeof(double));
    int64_t* piv = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    memcpy(LU, A, (size_t)(n * n) * sizeof(double));
    int info = 0;
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = k;
        double mx = fabs(LU[k * n + k]);
        for (int64_t i = k + 1; i < n; ++i)
            if (fabs(LU[k * n + i]) > mx) {
                mx = fabs(LU[k * n + i]);
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int64_t j = 0; j < n; ++j) {
                double t = LU[j * n + k];
                LU[j * n + k] = LU[j * n + p];
                LU[j * n + p] = t;
            }
        if (LU[k * n + k] == 0.0) {
            if (!info) info = (int)(k + 1);
            continue;
        }
        for (int64_t i = k + 1; i < n; ++i) {
            double l = LU[k * n + i] / LU[k * n + k];
            LU[k * n + i] = l;
            for (int64_t j = k + 1; j < n; ++j) LU[j * n + i] -= l * LU[j * n + k];
        }
    }
    if (info) {
        free(LU);
        free(piv);
        return info;
    }
    double* col = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int64_t c = -1; c < n; ++c) {
        double* x = (c < 0) ? b : col;
        if (c >= 0)
            for (int64_t i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
        for (int64_t k = 0; k < n; ++k)
            if (piv[k] != k) {
                double t = x[k];
                x[k] = x[piv[k]];
                x[piv[k]] = t;
            }
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = 0; k < i; ++k) x[i] -= LU[k * n + i] * x[k];
        for (int64_t i = n - 1; i >= 0; --i) {
            for (int64_t k = i + 1; k < n; ++k) x[i] -= LU[k * n + i] * x[k];
            x[i] /= LU[i * n + i];
        }
        if (c >= 0)
            for (int64_t i = 0; i < n; ++i) inv[c * n + i] = col[i];
    }
    free(col);
    free(LU);
    free(piv);
    return 0;
}

/* Cyclic Jacobi eigen-decomposition of a symmetric matrix. */
int fo_sym_eig(const double* A, int64_t n, double* w, double* V) {
    double* a = (double*)malloc((size_t)(n * n > 0 ? n * n : 1) * sizeof(double));
    memcpy(a, A, (size_t)(n * n) * sizeof(double));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) V[j * n + i] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) {
                tot += a[j * n + i] * a[j * n + i];
                if (i != j) off += a[j * n + i] * a[j * n + i];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int64_t p = 0; p < n; ++p)
            for (int64_t q = p + 1; q < n; ++q) {
                double apq = a[q * n + p];
                if (apq == 0.0) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int64_t k = 0; k < n; ++k) {
                    double akp = a[p * n + k], akq = a[q * n + k];
                    a[p * n + k] = cs * akp - sn * akq;
                    a[q * n + k] = sn * akp + cs * akq;
                }
                for (int64_t k = 0; k < n; ++k) {
                    double apk = a[k * n + p], aqk = a[k * n + q];
                    a[k * n + p] = cs * apk - sn * aqk;
                    a[k * n + q] = sn * apk + cs * aqk;
                }
                for (int64_t k = 0; k < n; ++k) {
                    double vkp = V[p * n + k], vkq = V[q * n + k];
                    V[p * n + k] = cs * vkp - sn * vkq;
                    V[q * n + k] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int64_t i = 0; i < n; ++i) w[i] = a[i * n + i];
    /* ascending order */
    for (int64_t i = 0; i < n; ++i) {
        int64_t m = i;
        for (int64_t j = i + 1; j < n; ++j)
            if (w[j] < w[m]) m = j;
        if (m != i) {
            double t = w[i];
            w[i] = w[m];
            w[m] = t;
            for (int64_t k = 0; k < n; ++k) {
                double tv = V[i * n + k];
                V[i * n + k] = V[m * n + k];
                V[m * n + k] = tv;
            }
        }
    }
    free(a);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Helpers for the estimator. */

static double fo_dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* compute_tss with intercept.  fe_pipeline.f90 compute_tss. */
static double fo_tss(const double* y, int64_t n) {
    if (n <= 0) return 0.0;
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) m += y[i];
    m /= (double)n;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (y[i] - m) * (y[i] - m);
    return s;
}

/* finalize_regression_stats.  fe_pipeline.f90 (end of module). */
static void fo_finalize_stats(fo_result* r, int64_t n_obs, int has_fes) {
    const double EPS = 1e-12;
    double n = (double)(n_obs > 1 ? n_obs : 1);
    double dfm = (double)(r->dof_model > 0 ? r->dof_model : 0);
    double dfr = (double)(r->dof_resid > 1 ? r->dof_resid : 1);
    double dev = r->rss;
    double dev0 = r->tss_total > EPS ? r->tss_total : EPS;
    if (dev0 > EPS) {
        r->r2 = 1.0 - dev / dev0;
        r->r2 = fmax(0.0, fmin(1.0, r->r2));
    } else {
        r->r2 = 0.0;
    }
    if (r->tss_within > EPS) {
        r->r2_within = 1.0 - dev / r->tss_within;
        r->r2_within = fmax(0.0, fmin(1.0, r->r2_within));
    } else {
        r->r2_within = r->r2;
    }
    (void)has_fes;
    double fe_flag = 1.0; /* has_intercept is always true */
    double k_total = (double)(r->dof_model + r->dof_fes + 1);
    double denom = fmax(n - k_total, 1.0);
    if (n > 1.0 && dev0 > EPS) {
        r->r2_adj = 1.0 - ((dev * (n - fe_flag)) / (dev0 * denom));
        r->r2_adj = fmax(0.0, fmin(1.0, r->r2_adj));
    } else {
        r->r2_adj = r->r2;
    }
    double ssr = fmax(0.0, dev0 - dev);
    if (dfm > 0.0 && dfr > 0.0 && dev > EPS)
        r->f_stat = (ssr / dfm) / (dev / dfr);
    else
        r->f_stat = 0.0;
}

/* scores = X' r per cluster, meat = S'S (K x K), ids 1..nc (others skipped).
 * fe_gpu_linalg.cu:29-77 semantics, fe_gpu_linalg.cu:200-206 (meat). */
static void fo_meat(const int32_t* ids, int64_t nc, const double* r, const double* X, int64_t n, int64_t K,
                    double* meat) {
    double* S = (double*)calloc((size_t)(nc > 0 ? nc : 1) * (size_t)(K > 0 ? K : 1), sizeof(double));
    /* column j of the scores on its own thread, rows in order: values independent of the thread count */
#pragma omp parallel for schedule(dynamic) if (n * K >= (1 << 18))
    for (int64_t j = 0; j < K; ++j)
        for (int64_t i = 0; i < n; ++i) {
            int64_t c = (int64_t)ids[i] - 1;
            if (c < 0 || c >= nc) continue;
            S[j * nc + c] += r[i] * X[j * n + i];
        }
#pragma omp parallel for schedule(dynamic) if (nc * K >= (1 << 16))
    for (int64_t a = 0; a < K; ++a)
        for (int64_t b = a; b < K; ++b) {
            double s = 0.0;
            for (int64_t c = 0; c < nc; ++c) s += S[a * nc + c] * S[b * nc + c];
            meat[b * K + a] = s;
            meat[a * K + b] = s;
        }
    free(S);
}

/* ------------------------------------------------------------------------- */
/* GPU-path estimator.  fe_pipeline.f90:516-1896. */
static int fo_estimate_gpu(int64_t n, int64_t K, int64_t L, int64_t n_fe, double* y, double* W, double* Z,
                           const int32_t* fe, const int32_t* G, const fo_options* opt, const int32_t* cluster_req,
                           int64_t n_cluster_req, const int32_t* iv_reg, int64_t n_iv_reg, const int32_t* iv_inst,
                           int64_t n_iv_inst, double* beta, double* se, double* vcov, fo_result* res) {
    /* cluster dims: valid, unique, first 3, request order (initialize_cluster_dims) */
    int64_t ncd = 0;
    for (int64_t i = 0; i < n_cluster_req; ++i) {
        int32_t d = cluster_req[i];
        if (d < 1 || d > n_fe) continue;
        int dup = 0;
        for (int64_t j = 0; j < ncd; ++j)
            if (res->cluster_dims[j] == d) dup = 1;
        if (dup) continue;
        if (ncd == 3) break;
        res->cluster_dims[ncd++] = d;
    }
    res->n_cluster_dims = ncd;
    res->tss_total = fo_tss(y, n);
    int64_t dof_fes = 0;
    for (int64_t d = 0; d < n_fe; ++d) dof_fes += (G[d] - 1 > 0 ? G[d] - 1 : 0);
    res->dof_fes = dof_fes;
    int64_t dof_nested = 0;
    for (int64_t j = 0; j < ncd; ++j) dof_nested += (G[res->cluster_dims[j] - 1] - 1 > 0 ? G[res->cluster_dims[j] - 1] - 1 : 0);

    int64_t conv = 0;
    double mu = 0.0;
    if (opt->ap_kind == 1)
        res->fe_iterations = fo_demean_forward(y, W, n, K, fe, n_fe, G, opt->tol, opt->max_iter, &conv, &mu);
    else
        res->fe_iterations = fo_demean_relaxed(y, W, Z, n, K, L, fe, n_fe, G, opt->tol, opt->max_iter, &conv, &mu);
    if (opt->ap_kind == 1 && L > 0) {
        /* the forward scheme has no Z argument: sweep the instruments on their own (M_F acts column-wise) */
        int64_t c2;
        double m2;
        double* dummy = (double*)calloc((size_t)n, sizeof(double));
        fo_demean_forward(dummy, Z, n, L, fe, n_fe, G, opt->tol, opt->max_iter, &c2, &m2);
        free(dummy);
    }
    res->converged = conv;
    res->max_update = mu;
    for (int64_t k = 0; k < K; ++k) {
        beta[k] = 0.0;
        se[k] = 0.0;
    }
    if (vcov)
        for (int64_t e = 0; e < K * K; ++e) vcov[e] = 0.0;
    if (K <= 0) return 0;
    res->tss_within = fo_dot(y, y, n);

    /* cross products, upper then symmetrised (symmetrize_upper) */
    double* WW = (double*)malloc((size_t)(K * K) * sizeof(double));
    double* Wy = (double*)malloc((size_t)K * sizeof(double));
    /* threads take whole entries (each dot runs serially), so the values do not depend on their count */
#pragma omp parallel for schedule(dynamic) if (n * K >= (1 << 18))
    for (int64_t a = 0; a < K; ++a) {
        Wy[a] = fo_dot(W + a * n, y, n);
        for (int64_t b = a; b < K; ++b) {
            double s = fo_dot(W + a * n, W + b * n, n);
            WW[b * K + a] = s;
            WW[a * K + b] = s;
        }
    }
    double* vdiag = (double*)malloc((size_t)K * sizeof(double));
    for (int64_t k = 0; k < K; ++k) vdiag[k] = WW[k * K + k];

    /* endogenous / exogenous / instruments (:664-753) */
    int* is_endog = (int*)calloc((size_t)K, sizeof(int));
    for (int64_t i = 0; i < n_iv_reg; ++i)
        if (iv_reg[i] >= 1 && iv_reg[i] <= K) is_endog[iv_reg[i] - 1] = 1;
    int64_t n_endog = 0;
    for (int64_t k = 0; k < K; ++k) n_endog += is_endog[k];
    int64_t n_exog = K - n_endog;
    int64_t* inst = (int64_t*)malloc((size_t)(L > 0 ? L : 1) * sizeof(int64_t) + (size_t)n_iv_inst * sizeof(int64_t));
    int64_t n_sel = 0;
    if (L > 0) {
        if (n_iv_inst > 0) {
            for (int64_t i = 0; i < n_iv_inst; ++i)
                if (iv_inst[i] >= 1 && iv_inst[i] <= L) inst[n_sel++] = iv_inst[i];
            if (n_sel == 0) {
                for (int64_t i = 0; i < L; ++i) inst[i] = i + 1;
                n_sel = L;
            }
        } else {
            for (int64_t i = 0; i < L; ++i) inst[i] = i + 1;
            n_sel = L;
        }
    }
    int64_t n_aug = n_exog + n_sel;
    int use_iv = n_endog > 0;
    if (use_iv && n_sel < n_endog) use_iv = 0;
    res->is_iv = use_iv;
    res->n_instruments = use_iv ? n_aug : 0;

    double* Qm = (double*)malloc((size_t)(K * K) * sizeof(double));
    double* bm = (double*)malloc((size_t)K * sizeof(double));
    double* Xhat = NULL;
    if (use_iv) {
        /* build_iv_normal_equations (:947-1032) */
        double* Za = (double*)malloc((size_t)n * (size_t)n_aug * sizeof(double));
        int64_t col = 0;
        for (int64_t k = 0; k < K; ++k)
            if (!is_endog[k]) memcpy(Za + (col++) * n, W + k * n, (size_t)n * sizeof(double));
        for (int64_t i = 0; i < n_sel; ++i) memcpy(Za + (col++) * n, Z + (inst[i] - 1) * n, (size_t)n * sizeof(double));
        double* Qzz = (double*)malloc((size_t)(n_aug * n_aug) * sizeof(double));
        double* Qzy = (double*)malloc((size_t)n_aug * sizeof(double));
        double* Qzx = (double*)malloc((size_t)(n_aug * K) * sizeof(double));
#pragma omp parallel for schedule(dynamic) if (n * n_aug >= (1 << 18))
        for (int64_t a = 0; a < n_aug; ++a) {
            Qzy[a] = fo_dot(Za + a * n, y, n);
            for (int64_t b = a; b < n_aug; ++b) {
                double s = fo_dot(Za + a * n, Za + b * n, n);
                Qzz[b * n_aug + a] = s;
                Qzz[a * n_aug + b] = s;
            }
            for (int64_t k = 0; k < K; ++k) Qzx[k * n_aug + a] = fo_dot(Za + a * n, W + k * n, n);
        }
        double* rhs = (double*)malloc((size_t)n_aug * sizeof(double));
        double* Qzinv = (double*)malloc((size_t)(n_aug * n_aug) * sizeof(double));
        int info = fo_chol_solve_and_invert(Qzz, Qzy, n_aug, rhs, Qzinv);
        if (info != 0) {
            res->solver_info = info;
            free(Za);
            free(Qzz);
            free(Qzy);
            free(Qzx);
            free(rhs);
            free(Qzinv);
            free(WW);
            free(Wy);
            free(vdiag);
            free(is_endog);
            free(inst);
            free(Qm);
            free(bm);
            return 0;
        }
        double* proj = (double*)malloc((size_t)(n_aug * K) * sizeof(double));
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < K; ++k)
            for (int64_t a = 0; a < n_aug; ++a) {
                double s = 0.0;
                for (int64_t c = 0; c < n_aug; ++c) s += Qzinv[c * n_aug + a] * Qzx[k * n_aug + c];
                proj[k * n_aug + a] = s;
            }
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < K; ++p)
            for (int64_t q = 0; q < K; ++q) {
                double s = 0.0;
                for (int64_t a = 0; a < n_aug; ++a) s += Qzx[p * n_aug + a] * proj[q * n_aug + a];
                Qm[q * K + p] = s;
            }
        for (int64_t p = 0; p < K; ++p)
            for (int64_t q = p + 1; q < K; ++q) {
                double s = 0.5 * (Qm[q * K + p] + Qm[p * K + q]);
                Qm[q * K + p] = s;
                Qm[p * K + q] = s;
            }
        for (int64_t p = 0; p < K; ++p) {
            double s = 0.0;
            for (int64_t a = 0; a < n_aug; ++a) s += Qzx[p * n_aug + a] * rhs[a];
            bm[p] = s;
        }
        Xhat = (double*)malloc((size_t)n * (size_t)K * sizeof(double));
#pragma omp parallel for schedule(dynamic)
        for (int64_t k = 0; k < K; ++k)
            for (int64_t i = 0; i < n; ++i) {
                double s = 0.0;
                for (int64_t a = 0; a < n_aug; ++a) s += Za[a * n + i] * proj[k * n_aug + a];
                Xhat[k * n + i] = s;
            }
        free(Za);
        free(Qzz);
        free(Qzy);
        free(Qzx);
        free(rhs);
        free(Qzinv);
        free(proj);
    } else {
        memcpy(Qm, WW, (size_t)(K * K) * sizeof(double));
        memcpy(bm, Wy, (size_t)K * sizeof(double));
    }

    /* solve_with_column_filter (:1034-1107) */
    double dmax = 0.0;
    for (int64_t k = 0; k < K; ++k)
        if (fabs(vdiag[k]) > dmax) dmax = fabs(vdiag[k]);
    if (dmax < 1.0) dmax = 1.0;
    double dtol = 1e-10 * dmax;
    int64_t* keep = (int64_t*)malloc((size_t)K * sizeof(int64_t));
    int64_t kept = 0;
    for (int64_t k = 0; k < K; ++k)
        if (!(fabs(vdiag[k]) <= dtol)) keep[kept++] = k;
    res->dof_model = kept;
    int info = 0;
    double* Qinv = NULL;
    if (kept == 0) {
        info = -1;
    } else {
        double* Qs = (double*)malloc((size_t)(kept * kept) * sizeof(double));
        double* bs = (double*)malloc((size_t)kept * sizeof(double));
        double* bsol = (double*)malloc((size_t)kept * sizeof(double));
        Qinv = (double*)malloc((size_t)(kept * kept) * sizeof(double));
        for (int64_t a = 0; a < kept; ++a) {
            bs[a] = bm[keep[a]];
            for (int64_t b = 0; b < kept; ++b) Qs[b * kept + a] = Qm[keep[b] * K + keep[a]];
        }
        info = fo_chol_solve_and_invert(Qs, bs, kept, bsol, Qinv);
        if (info == 0) {
            for (int64_t a = 0; a < kept; ++a) beta[keep[a]] = bsol[a];
            res->rank = kept;
        }
        free(Qs);
        free(bs);
        free(bsol);
    }
    res->solver_info = info;
    if (info == 0) {
        /* compute_standard_errors (:1109-1420) */
        double* r = (double*)malloc((size_t)n * sizeof(double));
#pragma omp parallel for schedule(static) if (n * K >= (1 << 18))
        for (int64_t i = 0; i < n; ++i) {
            double s = y[i];
            for (int64_t k = 0; k < K; ++k) s -= W[k * n + i] * beta[k];
            r[i] = s;
        }
        double rss = fo_dot(r, r, n);
        int64_t pct = (res->dof_model > 0 ? res->dof_model : 0) + dof_fes + 1;
        int64_t df = n - pct;
        if (df < 1) df = 1;
        res->dof_resid = df;
        res->rss = rss;
        double sigma2 = rss / (double)df;
        int cl_ok = 1;
        double* cov = NULL;
        if (ncd > 0) {
            const double* X = use_iv ? Xhat : W;
            cov = (double*)calloc((size_t)(kept * kept), sizeof(double));
            double* meat = (double*)malloc((size_t)(K * K) * sizeof(double));
            double* tmp = (double*)malloc((size_t)(kept * kept) * sizeof(double));
            int32_t* ids = (int32_t*)malloc((size_t)n * sizeof(int32_t));
            int64_t min_single = INT32_MAX;
            for (int64_t mask = 1; mask < (1LL << ncd); ++mask) {
                int32_t sd[3], ss[3];
                int len = 0;
                for (int64_t j = 0; j < ncd; ++j)
                    if (mask & (1LL << j)) {
                        sd[len] = (int32_t)res->cluster_dims[j];
                        ss[len] = G[sd[len] - 1];
                        ++len;
                    }
                double w = (len % 2 == 1) ? 1.0 : -1.0;
                int64_t nc = 0;
                const int32_t* idp = NULL;
                if (len == 1) {
                    nc = G[sd[0] - 1];
                    if (nc <= 0) {
                        cl_ok = 0;
                        break;
                    }
                    if (nc < min_single) min_single = nc;
                    idp = fe + (int64_t)(sd[0] - 1) * n;
                } else {
                    int32_t ncl = 0;
                    int st = fo_cluster_ids_dimmajor(fe, n, sd, ss, len, ids, &ncl);
                    if (st != 0 || ncl <= 0) {
                        cl_ok = 0;
                        break;
                    }
                    nc = ncl;
                    idp = ids;
                }
                res->n_clusters[mask - 1] = nc;
                fo_meat(idp, nc, r, X, n, K, meat);
                /* cov += w * Qinv M_kept Qinv */
#pragma omp parallel for schedule(static) if (kept >= 32)
                for (int64_t a = 0; a < kept; ++a)
                    for (int64_t b = 0; b < kept; ++b) {
                        double s = 0.0;
                        for (int64_t c = 0; c < kept; ++c) s += meat[keep[c] * K + keep[a]] * Qinv[b * kept + c];
                        tmp[b * kept + a] = s;
                    }
#pragma omp parallel for schedule(static) if (kept >= 32)
                for (int64_t a = 0; a < kept; ++a)
                    for (int64_t b = 0; b < kept; ++b) {
                        double s = 0.0;
                        for (int64_t c = 0; c < kept; ++c) s += Qinv[c * kept + a] * tmp[b * kept + c];
                        cov[b * kept + a] += w * s;
                    }
            }
            free(meat);
            free(tmp);
            free(ids);
            if (cl_ok) {
                int64_t nested_adj = dof_nested > 0 ? 1 : 0;
                int64_t pce = (res->dof_model > 0 ? res->dof_model : 0) + nested_adj;
                int64_t den = n - pce;
                if (den < 1) den = 1;
                int64_t num = n - 1;
                if (num < 1) num = 1;
                double scalar = (double)num / (double)den;
                if (min_single > 1) scalar = scalar * (double)min_single / (double)(min_single - 1);
                for (int64_t e = 0; e < kept * kept; ++e) cov[e] *= scalar;
                /* symmetrize_upper then PSD fix via eigen-decomposition of the upper triangle */
                for (int64_t b = 0; b < kept; ++b)
                    for (int64_t a = b + 1; a < kept; ++a) cov[b * kept + a] = cov[a * kept + b];
                double* ev = (double*)malloc((size_t)kept * sizeof(double));
                double* V = (double*)malloc((size_t)(kept * kept) * sizeof(double));
                fo_sym_eig(cov, kept, ev, V);
                int fixed = 0;
                for (int64_t a = 0; a < kept; ++a) {
                    double lam = ev[a];
                    if (lam < 0.0) {
                        lam = 0.0;
                        fixed = 1;
                    }
                    double sl = lam > 0.0 ? sqrt(lam) : 0.0;
                    for (int64_t c = 0; c < kept; ++c) V[a * kept + c] *= sl;
                }
                for (int64_t a = 0; a < kept; ++a)
                    for (int64_t b = 0; b < kept; ++b) {
                        double s = 0.0;
                        for (int64_t c = 0; c < kept; ++c) s += V[c * kept + a] * V[c * kept + b];
                        cov[b * kept + a] = s;
                    }
                res->psd_fix_applied = fixed;
                free(ev);
                free(V);
                for (int64_t a = 0; a < kept; ++a) se[keep[a]] = sqrt(fmax(0.0, cov[a * kept + a]));
                res->min_single_cluster = min_single;
                res->dof_resid = (min_single - 1 > 1) ? min_single - 1 : 1;
                if (vcov)
                    for (int64_t a = 0; a < kept; ++a)
                        for (int64_t b = 0; b < kept; ++b) vcov[keep[b] * K + keep[a]] = cov[b * kept + a];
            } else {
                res->n_cluster_dims = 0;
                for (int64_t m = 0; m < 7; ++m) res->n_clusters[m] = 0;
            }
        }
        if (ncd == 0 || !cl_ok) {
            for (int64_t a = 0; a < kept; ++a) se[keep[a]] = sqrt(fmax(0.0, sigma2 * Qinv[a * kept + a]));
            if (vcov)
                for (int64_t a = 0; a < kept; ++a)
                    for (int64_t b = 0; b < kept; ++b) vcov[keep[b] * K + keep[a]] = sigma2 * Qinv[b * kept + a];
        }
        free(cov);
        free(r);
        fo_finalize_stats(res, n, n_fe > 0);
    }
    free(Qinv);
    free(keep);
    free(Xhat);
    free(WW);
    free(Wy);
    free(vdiag);
    free(is_endog);
    free(inst);
    free(Qm);
    free(bm);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* CPU-fallback estimator.  fe_pipeline.f90:71-513 (no IV, no filter, LU, sigma2 = rss/(N-K),
 * CGM scalar (N-1)/(N-K) G/(G-1), symmetrize by averaging, no PSD fix). */
static int fo_estimate_cpu(int64_t n, int64_t K, int64_t n_fe, double* y, double* W, const int32_t* fe,
                           const int32_t* G, const fo_options* opt, const int32_t* cluster_req, int64_t n_cluster_req,
                           double* beta, double* se, double* vcov, fo_result* res) {
    int64_t ncd = 0;
    for (int64_t i = 0; i < n_cluster_req; ++i) {
        int32_t d = cluster_req[i];
        if (d < 1 || d > n_fe) continue;
        int dup = 0;
        for (int64_t j = 0; j < ncd; ++j)
            if (res->cluster_dims[j] == d) dup = 1;
        if (dup) continue;
        if (ncd == 3) break; /* initialize_cpu_cluster_dims keeps the first 3 valid */
        res->cluster_dims[ncd++] = d;
    }
    res->n_cluster_dims = ncd;
    res->tss_total = fo_tss(y, n);
    for (int64_t k = 0; k < K; ++k) {
        beta[k] = 0.0;
        se[k] = 0.0;
    }
    if (vcov)
        for (int64_t e = 0; e < K * K; ++e) vcov[e] = 0.0;
    int64_t conv = 1;
    double mu = 0.0;
    if (n_fe > 0)
        res->fe_iterations = fo_demean_forward(y, W, n, K, fe, n_fe, G, opt->tol, opt->max_iter, &conv, &mu);
    else
        res->fe_iterations = 0;
    res->converged = conv;
    res->max_update = mu;
    res->tss_within = fo_dot(y, y, n);
    int64_t dof_fes = 0;
    for (int64_t d = 0; d < n_fe; ++d) dof_fes += (G[d] - 1 > 0 ? G[d] - 1 : 0);
    res->dof_fes = dof_fes;
    if (K > 0) {
        double* WW = (double*)malloc((size_t)(K * K) * sizeof(double));
        double* Wy = (double*)malloc((size_t)K * sizeof(double));
        double* Qi = (double*)malloc((size_t)(K * K) * sizeof(double));
#pragma omp parallel for schedule(dynamic) if (n >= 65536)
        for (int64_t a = 0; a < K; ++a) {
            Wy[a] = fo_dot(W + a * n, y, n);
            for (int64_t b = 0; b < K; ++b) WW[b * K + a] = fo_dot(W + a * n, W + b * n, n);
        }
        int info = fo_lu_solve_and_invert(WW, Wy, K, Qi);
        res->solver_info = info;
        if (info != 0) {
            free(WW);
            free(Wy);
            free(Qi);
            return 0;
        }
        for (int64_t k = 0; k < K; ++k) beta[k] = Wy[k];
        res->rank = K;
        res->dof_model = K;
        double* r = (double*)malloc((size_t)n * sizeof(double));
        double rss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rss) if (n >= 65536)
        for (int64_t i = 0; i < n; ++i) {
            double s = y[i];
            for (int64_t k = 0; k < K; ++k) s -= W[k * n + i] * beta[k];
            r[i] = s;
            rss += s * s;
        }
        double den = (double)(n - K);
        if (den < 1.0) den = 1.0;
        double sigma2 = rss / den;
        res->rss = rss;
        res->dof_resid = n - K;
        int cl_ok = 0;
        if (ncd > 0) {
            /* compute_clustered_se_cpu (:385-483) */
            cl_ok = 1;
            double* acc = (double*)calloc((size_t)(K * K), sizeof(double));
            double* meat = (double*)malloc((size_t)(K * K) * sizeof(double));
            double* tmp = (double*)malloc((size_t)(K * K) * sizeof(double));
            int32_t* ids = (int32_t*)malloc((size_t)n * sizeof(int32_t));
            int64_t min_single = INT32_MAX;
            for (int64_t mask = 1; mask < (1LL << ncd); ++mask) {
                int32_t sd[3], ss[3];
                int len = 0;
                for (int64_t j = 0; j < ncd; ++j)
                    if (mask & (1LL << j)) {
                        sd[len] = (int32_t)res->cluster_dims[j];
                        ss[len] = G[sd[len] - 1];
                        ++len;
                    }
                int32_t ncl = 0;
                int st = fo_cluster_ids_dimmajor(fe, n, sd, ss, len, ids, &ncl);
                if (st != 0 || ncl <= 0) {
                    cl_ok = 0;
                    break;
                }
                if (len == 1 && ncl < min_single) min_single = ncl;
                res->n_clusters[mask - 1] = ncl;
                fo_meat(ids, ncl, r, W, n, K, meat);
                double w = (len % 2 == 1) ? 1.0 : -1.0;
                for (int64_t a = 0; a < K; ++a)
                    for (int64_t b = 0; b < K; ++b) {
                        double s = 0.0;
                        for (int64_t c = 0; c < K; ++c) s += meat[c * K + a] * Qi[b * K + c];
                        tmp[b * K + a] = s;
                    }
                for (int64_t a = 0; a < K; ++a)
                    for (int64_t b = 0; b < K; ++b) {
                        double s = 0.0;
                        for (int64_t c = 0; c < K; ++c) s += Qi[c * K + a] * tmp[b * K + c];
                        acc[b * K + a] += w * s;
                    }
            }
            if (cl_ok) {
                if (min_single < INT32_MAX) res->dof_resid = (min_single - 1 > 1) ? min_single - 1 : 1;
                double dn = (double)((n - K) > 1 ? (n - K) : 1);
                double scalar = (double)((n - 1) > 1 ? (n - 1) : 1) / dn;
                if (min_single > 1 && min_single < INT32_MAX) scalar *= (double)min_single / (double)(min_single - 1);
                for (int64_t a = 0; a < K; ++a)
                    for (int64_t b = a; b < K; ++b) {
                        double s = 0.5 * (acc[b * K + a] + acc[a * K + b]) * scalar;
                        acc[b * K + a] = s;
                        acc[a * K + b] = s;
                    }
                for (int64_t k = 0; k < K; ++k) se[k] = sqrt(fmax(0.0, acc[k * K + k]));
                res->min_single_cluster = min_single;
                if (vcov) memcpy(vcov, acc, (size_t)(K * K) * sizeof(double));
            } else {
                res->n_cluster_dims = 0;
                for (int64_t m = 0; m < 7; ++m) res->n_clusters[m] = 0;
            }
            free(acc);
            free(meat);
            free(tmp);
            free(ids);
        }
        if (!cl_ok) {
            for (int64_t k = 0; k < K; ++k) se[k] = sqrt(fmax(0.0, sigma2 * Qi[k * K + k]));
            if (vcov)
                for (int64_t e = 0; e < K * K; ++e) vcov[e] = sigma2 * Qi[e];
        }
        free(r);
        free(WW);
        free(Wy);
        free(Qi);
    }
    fo_finalize_stats(res, n, n_fe > 0);
    return 0;
}

int fo_estimate(int64_t n, int64_t K, int64_t L, int64_t n_fe, double* y, double* W, double* Z, const int32_t* fe,
                const int32_t* G, const fo_options* opt, const int32_t* cluster_req, int64_t n_cluster_req,
                const int32_t* iv_reg, int64_t n_iv_reg, const int32_t* iv_inst, int64_t n_iv_inst, double* beta,
                double* se, double* vcov, fo_result* res) {
    memset(res, 0, sizeof(*res));
    if (n_fe > 64) return 2;
    if (opt->conv == 1)
        return fo_estimate_cpu(n, K, n_fe, y, W, fe, G, opt, cluster_req, n_cluster_req, beta, se, vcov, res);
    return fo_estimate_gpu(n, K, L, n_fe, y, W, Z, fe, G, opt, cluster_req, n_cluster_req, iv_reg, n_iv_reg, iv_inst,
                           n_iv_inst, beta, se, vcov, res);
}
