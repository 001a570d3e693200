/*
 * oracle.c -- CPU restatement of speedynn's greedy kNN wrapper feature selection.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * speedynn_amd/: only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The engine never calls it.
 *
 * It restates, in plain C with 64-bit indexing and no n threshold, the
 * semantics of the reference's CPU module:
 *   forward_selection   speedynn/lib/nearest_neighbors_feature_selection.c:243-394
 *   backward_selection  speedynn/lib/nearest_neighbors_feature_selection.c:76-233
 *   optimal_performance speedynn/lib/nearest_neighbors_feature_selection.c:446-465
 *   get_smallest_validation_errors_cpu  speedynn/lib/cpu.c:21-126
 *   get_k_nearest_neighbors_cpu         speedynn/lib/cpu.c:128-168
 *   compute_distance_matrix_cpu         speedynn/lib/cpu.c:185-215 (dist: util.c:10-17)
 *   update_current_dist_matrix_cpu      speedynn/lib/cpu.c:217-231 (util.c:19-21)
 *   get_mode_array                      speedynn/lib/util.c:165-188
 *   get_loss and the nine losses        speedynn/lib/loss_functions.c:6-183
 *
 * Arithmetic: IEEE binary32 in the reference's operation order, compiled with
 * -ffp-contract=off so no multiply-add is fused.  Parallelism (OpenMP) only
 * splits independent queries; every reduction runs sequentially in the
 * reference order, so results do not depend on the thread count.
 *
 * Parity pins (tests/test_oracle.py): golden vectors produced by the compiled
 * reference C (tests/golden/), brute-force sort-based top-k on tiny inputs,
 * closed-form losses, and the enumerated get_mode_array tie cases.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_FLOAT 3.402823466e+38f /* speedynn/lib/types.h:9 */

enum { L_RMSE = 0, L_RMSE_NORM, L_MAD, L_MAD_NORM, L_ZERO_ONE, L_STD,
       L_RMSE_NORM_MAD_NORM_STD, L_RMSE_NORM_MAD_NORM_GALAXIES, L_RMSE_STD };

/* ---------------------------------------------------------------- distances */

/* util.c:10-17: sum over i = dim-1 ... 0 of (a_i - b_i)^2, left to right. */
static float dist_all(const float *a, const float *b, int dim) {
    float d = 0.0f;
    for (int i = dim; i--;) {
        float t = a[i] - b[i];
        d += t * t;
    }
    return d;
}

/* util.c:19-21 */
static float dist_one(const float *a, const float *b, int f) {
    float t = a[f] - b[f];
    return t * t;
}

/* Full n x n matrix in reference order (cpu.c:185-215). */
void oracle_distance_matrix(const float *X, int64_t n, int d, float *D, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) {
        D[i * n + i] = 0.0f;
        for (int64_t j = i + 1; j < n; j++) {
            float v = dist_all(X + i * d, X + j * d, d);
            D[i * n + j] = v;
            D[j * n + i] = v;
        }
    }
}

/* D += mult * (x_f - x_f)^2 over the strict upper triangle, mirrored (cpu.c:217-231). */
void oracle_update(const float *X, int64_t n, int d, float *D, int f, int mult, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = i + 1; j < n; j++) {
            float q = dist_one(X + i * d, X + j * d, f);
            D[i * n + j] = D[i * n + j] + (float)mult * q;
            D[j * n + i] = D[j * n + i] + (float)mult * q;
        }
    }
}

/* One row of the current matrix, computed from scratch in reference order.
 * forward:  0 + q_{o0} + q_{o1} + ... (selection order)
 * backward: dist over all d (descending feature), then - q_{o0} - q_{o1} ...
 * (removal order).  Used to spot-check large inputs without an n^2 buffer. */
void oracle_distance_row(const float *X, int64_t n, int d, int64_t q, int method,
                         const int *order, int s, float *row) {
    const float *xq = X + q * d;
    for (int64_t j = 0; j < n; j++) {
        const float *xj = X + j * d;
        float v;
        if (method == 1) {
            v = 0.0f;
            for (int t = 0; t < s; t++) v = v + dist_one(xq, xj, order[t]);
        } else {
            v = (j == q) ? 0.0f : dist_all(xq, xj, d);
            for (int t = 0; t < s; t++) v = v - dist_one(xq, xj, order[t]);
        }
        row[j] = (j == q) ? 0.0f : v;
    }
}

/* Exact (double) squared distances over a feature subset, for accuracy reports. */
void oracle_distance_row_f64(const float *X, int64_t n, int d, int64_t q,
                             const int *feats, int s, double *row) {
    for (int64_t j = 0; j < n; j++) {
        double v = 0.0;
        for (int t = 0; t < s; t++) {
            double a = (double)X[q * d + feats[t]] - (double)X[j * d + feats[t]];
            v += a * a;
        }
        row[j] = v;
    }
}

/* ----------------------------------------------------------------- top-k */

/* cpu.c:128-168: scan i ascending, skip the query's fold, strict '>' entry and
 * strict '<' bubble; unfilled slots keep (MAX_FLOAT, index 0). */
static void knn_row(const float *Drow, const float *X, int64_t n, int d, int64_t q,
                    int64_t npf, int f, int mult, int kmax, int *idx, float *dist) {
    for (int j = 0; j < kmax; j++) { idx[j] = 0; dist[j] = ORACLE_MAX_FLOAT; }
    int64_t fold = q / npf;
    const float *xq = X + q * d;
    for (int64_t i = 0; i < n; i++) {
        if (i / npf == fold) continue;
        float dd;
        if (f >= 0) dd = Drow[i] + (float)mult * dist_one(xq, X + i * d, f);
        else dd = Drow[i];
        int j = kmax - 1;
        if (dist[j] > dd) {
            dist[j] = dd;
            idx[j] = (int)i;
            for (; j > 0; j--) {
                if (dist[j] < dist[j - 1]) {
                    float t = dist[j]; dist[j] = dist[j - 1]; dist[j - 1] = t;
                    int ti = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = ti;
                } else break;
            }
        }
    }
}

/* util.c:165-188: repetition counts land on the first occurrence; the first
 * position with the strictly largest count wins. */
float oracle_mode(const float *a, int size) {
    int rep[1024];
    int *r = size <= 1024 ? rep : (int *)calloc((size_t)size, sizeof(int));
    for (int i = 0; i < size; ++i) {
        r[i] = 0;
        int j = 0;
        while (j < i && a[i] != a[j]) ++j;
        ++r[j];
    }
    int best = 0;
    for (int i = 1; i < size; ++i)
        if (r[i] > r[best]) best = i;
    float out = a[best];
    if (r != rep) free(r);
    return out;
}

/* ---------------------------------------------------------------- losses */

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    if (x - y < 0) return -1;
    if (x - y > 0) return 1;
    return 0;
}

/* loss_functions.c:173-183 */
static float median_of(const float *e, int64_t num) {
    float *s = (float *)malloc((size_t)num * sizeof(float));
    memcpy(s, e, (size_t)num * sizeof(float));
    qsort(s, (size_t)num, sizeof(float), cmp_float);
    float m = s[num / 2];
    free(s);
    return m;
}

/* loss_functions.c:41-50 */
static float loss_rmse(const float *y, const float *p, int64_t num) {
    float sum = 0.0f;
    for (int64_t i = 0; i < num; i++) { float t = y[i] - p[i]; sum += t * t; }
    sum /= (float)num;
    return (float)sqrt((double)sum);
}

/* loss_functions.c:52-62 */
static float loss_rmse_norm(const float *y, const float *p, int64_t num) {
    float sum = 0.0f;
    for (int64_t i = num; i--;) {
        float tv = y[i];
        float t = (float)((double)(tv - p[i]) / (1.0 + (double)tv));
        sum += t * t;
    }
    sum /= (float)num;
    return (float)sqrt((double)sum);
}

/* loss_functions.c:67-111 */
static float loss_mad(const float *y, const float *p, int64_t num, int norm) {
    float *e = (float *)malloc((size_t)num * sizeof(float));
    for (int64_t i = 0; i < num; i++) {
        float tv = y[i];
        e[i] = norm ? (float)((double)(tv - p[i]) / (1.0 + (double)tv)) : (tv - p[i]);
    }
    float m1 = median_of(e, num);
    for (int64_t i = 0; i < num; i++) e[i] = (float)fabs((double)(e[i] - m1));
    float m2 = median_of(e, num);
    free(e);
    return m2;
}

/* loss_functions.c:113-127: abs() is the int abs, so the float difference is
 * truncated toward zero first; an error is |trunc(y - p)| >= 1. */
static float loss_zero_one(const float *y, const float *p, int64_t num) {
    float sum = 0.0f;
    for (int64_t i = 0; i < num; i++) {
        int a = abs((int)(y[i] - p[i]));
        if (a < 10E-8) sum += 0.0f; else sum += 1.0f;
    }
    sum /= (float)num;
    return sum;
}

/* loss_functions.c:129-145, including the assignment at :141 that keeps only
 * the last squared deviation. */
static float loss_std(const float *y, const float *p, int64_t num) {
    float mean = 0.0f;
    for (int64_t i = 0; i < num; i++) mean += (y[i] - p[i]);
    mean /= (float)num;
    float sum = 0.0f;
    for (int64_t i = 0; i < num; i++) {
        float t = (y[i] - p[i]) - mean;
        sum = t * t;
    }
    sum /= (float)(num - 1);
    return (float)sqrt((double)sum);
}

/* loss_functions.c:6-38 */
float oracle_loss(const float *y, const float *p, int64_t num, int loss) {
    switch (loss) {
    case L_RMSE: return loss_rmse(y, p, num);
    case L_RMSE_NORM: return loss_rmse_norm(y, p, num);
    case L_MAD: return loss_mad(y, p, num, 0);
    case L_MAD_NORM: return loss_mad(y, p, num, 1);
    case L_ZERO_ONE: return loss_zero_one(y, p, num);
    case L_STD: return loss_std(y, p, num);
    case L_RMSE_NORM_MAD_NORM_STD: {
        float a = loss_rmse_norm(y, p, num), b = loss_mad(y, p, num, 1), c = loss_std(y, p, num);
        return (float)((double)a / 0.19 + (double)b / 0.041 + (double)c / 0.16);
    }
    case L_RMSE_NORM_MAD_NORM_GALAXIES: {
        float a = loss_rmse_norm(y, p, num), b = loss_mad(y, p, num, 1);
        return (float)((double)a / 0.019 + (double)b / 0.009);
    }
    case L_RMSE_STD: return loss_rmse(y, p, num) + loss_std(y, p, num);
    default: return 0.0f;
    }
}

/* ------------------------------------------------------ candidate predictions */

/* Predictions of all k in kvals for one candidate f over the given queries
 * (cpu.c:44-95): regression = nearest-first float sum / k, classification =
 * get_mode_array over the first k labels.  preds is [nk][nq]. */
void oracle_candidate_preds(const float *D, int64_t dstride, const float *X, int64_t n, int d,
                            const float *y, const int *kvals, int nk, int64_t k_fold, int f,
                            int mult, int model_type, const int64_t *queries, int64_t nq,
                            float *preds, int nthreads) {
    int kmax = -2147483647;
    for (int m = 0; m < nk; m++) if (kvals[m] > kmax) kmax = kvals[m];
    int64_t npf = n / k_fold;
#pragma omp parallel num_threads(nthreads)
    {
        int *idx = (int *)malloc((size_t)kmax * sizeof(int));
        float *dist = (float *)malloc((size_t)kmax * sizeof(float));
        float *lab = (float *)malloc((size_t)kmax * sizeof(float));
#pragma omp for schedule(dynamic, 8)
        for (int64_t t = 0; t < nq; t++) {
            int64_t q = queries ? queries[t] : t;
            const float *row = D + (queries ? t : q) * dstride;
            knn_row(row, X, n, d, q, npf, f, mult, kmax, idx, dist);
            for (int m = 0; m < nk; m++) {
                int k = kvals[m];
                float pred;
                if (model_type == 0) {
                    pred = 0.0f;
                    for (int j = 0; j < k; j++) pred += y[idx[j]];
                    pred /= (float)k;
                } else {
                    for (int j = 0; j < k; j++) lab[j] = y[idx[j]];
                    pred = oracle_mode(lab, k);
                }
                preds[(int64_t)m * nq + t] = pred;
            }
        }
        free(idx); free(dist); free(lab);
    }
}

/* ------------------------------------------------------------- selection */

typedef struct {
    int num_final;
    int best_k;
} oracle_summary;

/* Greedy selection.  method 1 = forward, 0 = backward.  Outputs follow the
 * reference's getters: sel_dims[d] mask, ordering[d], errors[nl*d] indexed
 * [l*d + step] (forward) or [l*d + remaining] (backward).  cand_errors, if not
 * NULL, receives the first-loss validation error of every candidate at every
 * step as [step][d] (NaN where not a candidate); cand_bestk likewise the chosen
 * k.  Returns num_final_features, or -1 on invalid arguments. */
int oracle_select(int method, const float *X, int64_t n, int d, const float *y,
                  const int *kvals, int nk, int num_features, int64_t k_fold,
                  const int *losses, int nl, int auto_feature, int model_type,
                  int nthreads, int *sel_dims, int *ordering, float *errors,
                  int *best_k_out, float *cand_errors, int *cand_bestk) {
    if (n <= 0 || d <= 0 || nk <= 0 || nl <= 0 || k_fold <= 0 || k_fold > n) return -1;
    int num_desired = num_features;
    int num_final = method == 1 ? 0 : d;
    if (auto_feature) num_desired = method == 1 ? d : 0;
    float *D = (float *)malloc((size_t)(n * n) * sizeof(float));
    if (!D) return -2;
    float *preds = (float *)malloc((size_t)(n * nk) * sizeof(float));
    float *val = (float *)malloc((size_t)(nl * d) * sizeof(float));
    float *kerr = (float *)malloc((size_t)(nl * nk) * sizeof(float));
    int *bestk = (int *)malloc((size_t)d * sizeof(int));
    for (int i = 0; i < nl * d; i++) { val[i] = 0.0f; errors[i] = 0.0f; }
    for (int i = 0; i < d; i++) { sel_dims[i] = method == 1 ? 0 : 1; ordering[i] = 0; bestk[i] = 0; }
    if (method == 1) memset(D, 0, (size_t)(n * n) * sizeof(float));
    else oracle_distance_matrix(X, n, d, D, nthreads);
    int mult = method == 1 ? +1 : -1;
    int check = method == 1 ? 0 : 1;
    int best_k_model = 0;
    int step = 0;
    int i = method == 1 ? 0 : d - 1;
    for (;;) {
        if (method == 1 && !(i < num_desired)) break;
        if (method == 0 && !(i >= num_desired)) break;
        num_final += method == 1 ? 1 : -1;
        for (int f = 0; f < d; f++) {
            if (cand_errors) cand_errors[(int64_t)step * d + f] = NAN;
            if (cand_bestk) cand_bestk[(int64_t)step * d + f] = 0;
            if (sel_dims[f] != check) continue;
            oracle_candidate_preds(D, n, X, n, d, y, kvals, nk, k_fold, f, mult, model_type,
                                   NULL, n, preds, nthreads);
            for (int l = 0; l < nl; l++)
                for (int m = 0; m < nk; m++)
                    kerr[l * nk + m] = oracle_loss(y, preds + (int64_t)m * n, n, losses[l]);
            float smallest = ORACLE_MAX_FLOAT;
            for (int m = 0; m < nk; m++) {
                if (kerr[m] < smallest) {
                    smallest = kerr[m];
                    bestk[f] = kvals[m];
                    for (int l = 0; l < nl; l++) val[l * d + f] = kerr[l * nk + m];
                }
            }
            if (cand_errors) cand_errors[(int64_t)step * d + f] = val[f];
            if (cand_bestk) cand_bestk[(int64_t)step * d + f] = bestk[f];
        }
        float minv = ORACLE_MAX_FLOAT;
        int mind = -1;
        for (int f = 0; f < d; f++)
            if (sel_dims[f] == check && val[f] < minv) { minv = val[f]; mind = f; }
        if (mind < 0) break; /* the reference would index best_current_k[-1] here */
        best_k_model = bestk[mind];
        sel_dims[mind] = method == 1 ? 1 : 0;
        ordering[i] = mind;
        for (int l = 0; l < nl; l++) errors[l * d + i] = val[l * d + mind];
        oracle_update(X, n, d, D, mind, mult, nthreads);
        step++;
        if (auto_feature) {
            if (method == 1) {
                if (i > 0 && errors[i - 1] < errors[i]) break;
            } else {
                if (i > 0 && i < d - 1 && errors[i + 1] < errors[i]) break;
            }
        }
        i += method == 1 ? 1 : -1;
    }
    *best_k_out = best_k_model;
    free(D); free(preds); free(val); free(kerr); free(bestk);
    return num_final;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
