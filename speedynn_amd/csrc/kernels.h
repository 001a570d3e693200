// Host-side launchers of the engine's kernels (kernels.hip, search.hip).
#pragma once
#include "common.h"

namespace snn {

hipError_t launch_fill(float* p, int64_t count, float v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int64_t count, hipStream_t s);
hipError_t launch_fill_i32(int* p, int64_t count, int v, hipStream_t s);
hipError_t launch_transpose(const float* X, float* XT, int64_t n, int d, hipStream_t s);
hipError_t launch_gram_exact(const float* XT, float* D, float* Dlo, float* M, int64_t n, int d,
                             hipStream_t s);
// Backward pruned search: M[i][j] = max over the state's remaining candidates
// (Xc columns c < ncand) of fl((Xc[i][c] - Xc[j][c])^2).
hipError_t launch_max_square(const float* Xc, const SelState* st, float* M, int64_t n, int cpad,
                             hipStream_t s);
// The update of the device state's pending feature fused with the seed scan
// of the next step (seeds as launch_seed_scan writes them); with n % 4 != 0
// only that (rl.idx null, defer 0).
// rl.idx set also writes the next search's row lists (RowList) from the warm
// bounds tauw (forward: and the chosen column's range frange; backward: the
// lower bounds fl(D - Mq) from Mq = M rounded up to 8 bits, launch_quantize_m).
// defer: 0 write D; 1 leave this step's term out of D (st->fdefer); 2 apply
// the left-out term and this step's, and write D (see k_update_seeds).
hipError_t launch_update_seeds(float* D, float* Dlo, const float* XT, int64_t n,
                               SelState* st, int mult, int* seeds, int64_t npf,
                               const RowList& rl, const float* tauw, const float* frange, int d,
                               const int* cand, const int* iperm, const unsigned char* Mq,
                               const int* Mexp, hipStream_t s, int defer = 0);
// Mq[i][j] = M[i][j] rounded up to 8 bits with the row scale Mexp[i] (n % 4 == 0;
// see k_quantize_m for the code).
hipError_t launch_quantize_m(const float* M, unsigned char* Mq, int* Mexp, int64_t n,
                             hipStream_t s);
hipError_t launch_update(float* D, float* Dlo, const float* XT, int64_t n, const SelState* st,
                         int f, int mult, hipStream_t s);
hipError_t launch_compact(const float* X, const int* cand, SelState* st, float* Xc, int64_t n,
                          int d, int cpad, hipStream_t s);
hipError_t launch_search_dense(const SearchProblem& a, hipStream_t s);
// Lists longer than 64 (a.kmax > 64): global-memory lists of a.kmax slots for
// `waves` persistent waves (ld, li, lab: waves * a.kmax * 64 values each;
// long_search_waves gives the count for a shape).
int64_t long_search_waves(int64_t n, int cpad, int K);
hipError_t launch_search_long(const SearchProblem& a, float* ld, int* li, float* lab,
                              int64_t waves, hipStream_t s);
// variant: 0 = by direction (forward: a wave per (query, candidate block);
// backward: a workgroup per query sharing the row scan), 1 = per wave,
// 2 = shared scan (where pruned_shared_applies), 3 = shared scan split over
// two workgroups per query (where pruned_split_applies), 5 / 6 / 7 = shared
// scan in workgroups of 4 / 8 / 16 waves, ceil(blocks / W) per query (where
// pruned_width_applies; else per wave).
hipError_t launch_search_pruned(const SearchProblem& a, int variant, hipStream_t s);
bool pruned_shared_applies(const SearchProblem& a);
bool pruned_width_applies(const SearchProblem& a, int W);
bool pruned_split_applies(const SearchProblem& a);
// Second forward step of the precision form: the walk along the first chosen
// feature's sorted column (a.skey, a.sidx; rank from launch_rank after the
// first step's arg-min).
hipError_t launch_search_walk(const SearchProblem& a, const int* rank, hipStream_t s);
hipError_t launch_rank(const int* sidx, const SelState* st, int* rank, int64_t n, hipStream_t s);
// BF16x3 MFMA Gram of the full-d distance matrix (gram.hip); scratch of
// gram_mfma_scratch_bytes(n, d) bytes.
size_t gram_mfma_scratch_bytes(int64_t n, int d);
hipError_t launch_gram_mfma(const float* X, const float* XT, int64_t n, int d, float* D,
                            void* scratch, hipStream_t s);
// First forward step over sorted columns (a.skey, a.sidx, a.cand); grid over
// max_cand candidate slots.  Returns at once where st->nonfinite is set.
hipError_t launch_search_sorted(const SearchProblem& a, int max_cand, hipStream_t s);
// Sorted columns of XT: skey[f][.] ascending, sidx[f][.] their point indices
// (stable: equal values keep ascending index); also sets st->nonfinite when X
// holds NaN or infinity.  temp: sort_columns_temp_bytes(n, d) bytes.
size_t sort_columns_temp_bytes(int64_t n, int d);
hipError_t launch_sort_columns(const float* XT, int64_t n, int d, SelState* st, float* skey,
                               int* sidx, void* temp, size_t temp_bytes, hipStream_t s);
hipError_t launch_seed_scan(const float* D, const SelState* st, int* seeds, int64_t n,
                            int64_t npf, hipStream_t s);
// scratch: median losses with NaN errors sort a copy there, 2n floats per
// (k value, candidate slot) -- nk * cpad * 2n for launch_loss, 2n for
// launch_loss_single; may be null when needs_median_scratch() is false.
// paths = loss_paths(host copy of the loss codes): bit 1 the wave-per-candidate
// kernel (zero_one, medians, composites), bit 2 the lane-per-candidate kernel
// (the sequential sums: rmse, rmse_norm, std, rmse_std).
int loss_paths(const int* codes, int count);
hipError_t launch_loss(const float* P, const float* y, const int* losses, const SelState* st,
                       float* err, int64_t n, int cpad, int nk, int nl, int max_cand,
                       float* scratch, int paths, hipStream_t s);
hipError_t launch_loss_single(const float* y, const float* p, int64_t n, int code, float* out,
                              float* scratch, hipStream_t s);
bool needs_median_scratch(const int* codes, int count);
hipError_t launch_select(SelState* st, const Problem& pb, const int* kvals, const float* err,
                         int* cand, int* sel, int* ordering, float* errors, float* val,
                         int* bestk, float* cand_err, const int* perm, hipStream_t s);
// perm[rank] = feature, iperm[feature] = rank (iperm may be null).
hipError_t launch_feature_order(const float* XT, int64_t n, int d, double* key, int* perm,
                                int* iperm, bool by_scale, hipStream_t s);
// frange = {min, max} of XT's column of the device state's last choice.
hipError_t launch_col_range(const float* XT, const SelState* st, int64_t n, float* frange,
                            hipStream_t s);
// st->nonfinite |= (some of the count values is NaN or infinite).
hipError_t launch_check_finite(const float* XT, int64_t count, SelState* st, hipStream_t s);
hipError_t launch_init_state(SelState* st, const Problem& pb, int* cand, int* sel, int* ordering,
                             float* errors, const int* perm, hipStream_t s);

// Persistent search over many problems (device array `probs`, item prefix
// sums `start[P+1]`, work counter).  mode 0: dense, 1: dense with D = 0
// (first forward step), 2: pruned; every problem of a launch has direction
// `mult` (+1 forward, -1 backward) and list size K.  batch_grid: resident
// blocks of that kernel on the current device (call before graph capture).
int64_t batch_dense_items(const SearchProblem& a);
int64_t batch_pruned_items(const SearchProblem& a);
// Seeds a backward pruned search reads for its bound (the nearest in D first):
// K + 16 rounded up to 8, at most 128; SPEEDYNN_NSEED overrides it.
int backward_seed_count(int kmax);
int batch_grid(int mode, int mult, int K);
hipError_t launch_batch_search(const SearchProblem* probs, const int64_t* start, int P,
                               unsigned long long* counter, int mode, int mult, int K, int grid,
                               hipStream_t s);

}  // namespace snn
