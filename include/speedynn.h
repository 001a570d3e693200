/*
 * speedynn.h -- C ABI of the MI355X kNN wrapper feature-selection engine.
 *
 * Two layers:
 *
 * 1. The reference's FFI surface, unchanged.  These are exactly the functions
 *    the reference's SWIG module binds (speedynn/lib/nearest_neighbors_feature_selection.h:59-80,
 *    typemaps in speedynn/lib/nearest_neighbors_feature_selection_float_gpu.i:14-22), with the
 *    same argument order and meaning, and process-global results read back by
 *    the getters.  Difference: errors return a negative code (see
 *    snn_last_error) instead of calling exit().
 *
 * 2. The engine API (snn_*): explicit contexts, device-resident inputs, the
 *    precision form switch, per-step candidate errors and phase timings.
 *
 * All pointers are host pointers unless the name says "device".  Calls are
 * blocking.  A context is not thread-safe; use one per host thread.
 */
#ifndef SPEEDYNN_AMD_H
#define SPEEDYNN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* 1. reference-compatible entry points                                       */
/* ------------------------------------------------------------------------- */

/* Replaces train_backward_selection, speedynn/lib/nearest_neighbors_feature_selection.c:49-54.
 * Xtrain is nXtrain x dXtrain float32 row-major; Ytrain has nYtrain == nXtrain
 * labels.  Returns the number of final features, or a negative error code. */
int train_backward_selection(float *Xtrain, int nXtrain, int dXtrain, float *Ytrain, int nYtrain,
                             int *kvals, int num_kvals, int num_features, int kfold,
                             int *loss_functions, int num_loss_functions, int auto_feature,
                             int model_type);

/* Replaces train_forward_selection, speedynn/lib/nearest_neighbors_feature_selection.c:60-65. */
int train_forward_selection(float *Xtrain, int nXtrain, int dXtrain, float *Ytrain, int nYtrain,
                            int *kvals, int num_kvals, int num_features, int kfold,
                            int *loss_functions, int num_loss_functions, int auto_feature,
                            int model_type);

/* Replace the getters at speedynn/lib/nearest_neighbors_feature_selection.c:401-428: copy the
 * first num entries of the last train call's results into the caller's array. */
void get_selected_dimensions(int *selected_dimensions_extern, int num_selected_dimensions_extern);
void get_selected_ordering(int *selected_ordering_extern, int num_selected_ordering_extern);
void get_selected_errors(float *selected_errors_extern, int num_selected_errors_extern);

/* Replaces get_final_optimal_model_k, speedynn/lib/nearest_neighbors_feature_selection.c:430-432. */
int get_final_optimal_model_k(void);

/* Replaces get_loss_all, speedynn/lib/nearest_neighbors_feature_selection.c:438-440 (losses of
 * speedynn/lib/loss_functions.c:6-38, computed on the GPU). */
float get_loss_all(float *labels, int num_labels, float *predictions, int num_predictions,
                   int loss_function);

/* ------------------------------------------------------------------------- */
/* 2. engine API                                                             */
/* ------------------------------------------------------------------------- */

#define SNN_OK 0
#define SNN_EINVAL (-1)   /* invalid argument */
#define SNN_ENOMEM (-2)   /* device allocation failed (D is n*n floats) */
#define SNN_EDEVICE (-3)  /* HIP runtime error */
#define SNN_ESTATE (-4)   /* no candidate could be chosen (all errors NaN/inf) */
#define SNN_ENOTSUP (-5)  /* combination not supported */

/* Loss codes (speedynn/base.py:13-23). */
#define SNN_LOSS_RMSE 0
#define SNN_LOSS_RMSE_NORM 1
#define SNN_LOSS_MAD 2
#define SNN_LOSS_MAD_NORM 3
#define SNN_LOSS_ZERO_ONE 4
#define SNN_LOSS_STD 5
#define SNN_LOSS_RMSE_NORM_MAD_NORM_STD 6
#define SNN_LOSS_RMSE_NORM_MAD_NORM_GALAXIES 7
#define SNN_LOSS_RMSE_STD 8

/* Precision forms. */
#define SNN_FORM_PLAIN 0     /* FP32 throughout, brute-force search */
#define SNN_FORM_PRECISION 1 /* sorted first step, exact pruned + screened search */
#define SNN_FORM_MFMA 2      /* distance matrices only: BF16x3 MFMA Gram (approximate) */

typedef struct snn_ctx snn_ctx;

typedef struct {
    int method;          /* 1 forward, 0 backward (speedynn/base.py:65-70) */
    int model_type;      /* 0 regression, 1 classification (speedynn/base.py:24-27) */
    int64_t k_fold;      /* contiguous folds; k_fold == n is leave-one-out */
    int num_features;    /* desired features (ignored when auto_feature) */
    int auto_feature;    /* stop at the first error increase */
    int form;            /* SNN_FORM_* */
    int use_graph;       /* capture selection steps in hipGraphs and replay them:
                            snn_select_batch one graph per step kind; snn_select and
                            snn_select_device one graph per distinct launch sequence
                            of a step, kept in the context for later calls of the same
                            shape (no phase times in the summary then) */
    int compensated;     /* precision form: also keep D's compensation term (TwoSum /
                            TwoProduct error, one more n*n float array); the selection
                            does not depend on it */
} snn_params;

typedef struct {
    int num_final;       /* return value of train_*_selection */
    int optimal_k;       /* get_final_optimal_model_k */
    int steps;           /* selection steps executed */
    int pruned_kernel;   /* pruned searches used, bit mask: 1 a wave per (query,
                            candidate block), 2 a workgroup per query sharing the row
                            scan, 4 that scan split over two workgroups per query,
                            8 the second forward step's walk along a sorted column,
                            16 / 32 / 64 the shared scan in workgroups of 4 / 8 / 16
                            waves (the autotuner's tile widths) */
    double ms_total;     /* device time of the selection (HIP events) */
    double ms_distance;  /* (a) initial matrix */
    double ms_search;    /* (c) candidate search kernels only */
    double ms_update;    /* (b) per-step update */
    double ms_tally;     /* candidate gather, losses and arg-min */
    int64_t evals;       /* (query x candidate) kNN evaluations performed */
    int64_t pruned_point_evals; /* point evaluations skipped by exact pruning */
    double ms_bound;     /* pruning-bound seeds (part of the precision form's search) */
    double ms_search_first; /* search kernels of the first step (included in ms_search) */
} snn_summary;

/* Create a context on HIP device `device`.  Returns SNN_OK or an error code. */
int snn_create(snn_ctx **out, int device);
void snn_destroy(snn_ctx *ctx);

/* Last error message of the calling thread (never NULL). */
const char *snn_last_error(void);

/* Greedy selection over host inputs.  Outputs (host, caller-owned):
 *   selected_dimensions[d], selected_ordering[d], selected_errors[num_loss*d]
 *   with the reference's layout ([l*d + step] forward, [l*d + remaining] backward);
 *   candidate_errors (optional, may be NULL): [d steps][d] first-loss validation
 *   error of every candidate at every step, NaN for non-candidates. */
int snn_select(snn_ctx *ctx, const float *X, int64_t n, int d, const float *y,
               const int *k_vals, int num_k, const int *loss_functions, int num_loss,
               const snn_params *params, int *selected_dimensions, int *selected_ordering,
               float *selected_errors, float *candidate_errors, snn_summary *summary);

/* Same, with X (n x d row-major) and y already resident in device memory of
 * the context's device. */
int snn_select_device(snn_ctx *ctx, const float *X_device, int64_t n, int d,
                      const float *y_device, const int *k_vals, int num_k,
                      const int *loss_functions, int num_loss, const snn_params *params,
                      int *selected_dimensions, int *selected_ordering, float *selected_errors,
                      float *candidate_errors, snn_summary *summary);

/* Bytes of the fixed device work array of snn_select_device_noalloc (one per GPU). */
#define SNN_NOALLOC_SCRATCH_BYTES (64u << 20)

/* snn_select_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every device work buffer (distance matrix, candidate columns, seeds, ...)
 * is a range of one fixed array of SNN_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; a problem that needs more returns SNN_EINVAL with
 * the size it needs.  Inputs and outputs are as in snn_select_device.  The
 * summary's times stay 0 (no events are recorded) and no autotuner trials
 * run (kernels the process has already chosen for the shape, else the
 * defaults).  Calls on one GPU share the array: calls that overlap on one
 * GPU race. */
int snn_select_device_noalloc(snn_ctx *ctx, const float *X_device, int64_t n, int d,
                              const float *y_device, const int *k_vals, int num_k,
                              const int *loss_functions, int num_loss, const snn_params *params,
                              int *selected_dimensions, int *selected_ordering,
                              float *selected_errors, float *candidate_errors,
                              snn_summary *summary);

/* One selection problem of a batch: host inputs and host outputs. */
typedef struct {
    const float *X;              /* n x d row-major */
    const float *y;              /* n labels */
    int64_t n;
    int d;
    const int *k_vals;
    int num_k;
    const int *loss_functions;
    int num_loss;
    snn_params params;           /* form and use_graph must agree across the batch */
    int *selected_dimensions;    /* outputs as in snn_select; each may be NULL */
    int *selected_ordering;
    float *selected_errors;
    int status;                  /* out: SNN_OK or this problem's error code */
    int num_final;               /* out */
    int optimal_k;               /* out */
    int steps;                   /* out */
    int inputs_on_device;        /* X and y are device pointers of the context's device */
} snn_problem;

/* Many independent selections (uneven n, d and k) advanced together: every
 * step runs the candidate search of all problems with the same direction and
 * neighbour-list length in one persistent launch over a shared work queue
 * (dense in the plain form; in the precision form a forward selection's first
 * step runs over sorted columns and later steps run the pruned search), with
 * the per-problem gather, losses, arg-min and update around it; with
 * params.use_graph the step is captured once in a hipGraph and replayed.  Problems whose arguments are invalid get their status set and
 * are left out; the call returns SNN_EINVAL if any was, after running the
 * others.  No counterpart in the reference, which selects one problem per call
 * (speedynn/lib/nearest_neighbors_feature_selection.c:49-65).  summary
 * (optional) holds the batch totals. */
int snn_select_batch(snn_ctx *ctx, int num_problems, snn_problem *problems,
                     snn_summary *summary);

/* Rows of the current distance matrix after applying `s` steps of `order`
 * from the method's start state (forward: D = 0 then + each feature;
 * backward: full matrix then - each feature).  rows_out is [nrows][n];
 * lo_out (optional) receives the compensation term of the precision form. */
int snn_distance_rows(snn_ctx *ctx, const float *X, int64_t n, int d, int method,
                      const int *order, int s, int form, const int64_t *rows, int nrows,
                      float *rows_out, float *lo_out);

/* The full-d distance matrix of X (device, n x d row-major) into D_device
 * (n x n floats): form SNN_FORM_PLAIN = exact FP32 in the reference's
 * descending-feature order (the selection's backward start), SNN_FORM_MFMA =
 * the BF16x3 matrix-core Gram over centred features (within 1e-5 relative of
 * FP64, not the reference's bits).  ms (optional) receives the device time of
 * the matrix kernels.  Replaces compute_distance_matrix_gpu,
 * speedynn/lib/gpu.c:354-379 (and compute_distance_matrix_cpu, cpu.c:185-215). */
int snn_distance_matrix(snn_ctx *ctx, const float *X_device, int64_t n, int d, int form,
                        float *D_device, double *ms);

/* Loss of predictions against labels on the GPU (any SNN_LOSS_*). */
int snn_loss(snn_ctx *ctx, const float *labels, const float *predictions, int64_t n,
             int loss_function, float *out);

/* Set the form used by the reference-compatible entry points (default 1). */
void snn_set_default_form(int form);

#ifdef __cplusplus
}
#endif

#endif /* SPEEDYNN_AMD_H */
