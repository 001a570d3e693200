// Host-side launchers of the engine's kernels (kernels.hip).
#pragma once
#include "common.h"

namespace snn {

struct SearchLaunch {
  const float* D;
  const float* Xc;
  const float* y;
  const int* kvals;
  const SelState* st;
  float* P;
  const int* seeds;  // pruned search: [n][128] bound seeds per query
  const float* M;    // backward pruned search: [n][n] max single-feature term
  unsigned long long* visited;  // pruned search: (query, point) pairs examined
  int64_t n;
  int64_t npf;
  int cpad;
  int nk;
  int kmax;
  int mult;
  int model_type;
  int zero;       // dense forward search with D known to be 0
};

int search_list_size(int kmax);

hipError_t launch_fill(float* p, int64_t count, float v, hipStream_t s);
hipError_t launch_transpose(const float* X, float* XT, int64_t n, int d, hipStream_t s);
hipError_t launch_gram_exact(const float* XT, float* D, float* Dlo, float* M, int64_t n, int d,
                             hipStream_t s);
hipError_t launch_update(float* D, float* Dlo, const float* XT, int64_t n, const SelState* st,
                         int f, int mult, hipStream_t s);
hipError_t launch_compact(const float* X, const int* cand, SelState* st, float* Xc, int64_t n,
                          int d, int cpad, hipStream_t s);
hipError_t launch_search_dense(const SearchLaunch& a, hipStream_t s);
hipError_t launch_search_pruned(const SearchLaunch& a, hipStream_t s);
hipError_t launch_seed_scan(const float* D, const SelState* st, int* seeds, int64_t n,
                            int64_t npf, hipStream_t s);
hipError_t launch_loss(const float* P, const float* y, const int* losses, const SelState* st,
                       float* err, int64_t n, int cpad, int nk, int nl, int max_cand,
                       hipStream_t s);
hipError_t launch_loss_single(const float* y, const float* p, int64_t n, int code, float* out,
                              hipStream_t s);
hipError_t launch_select(SelState* st, const Problem& pb, const int* kvals, const float* err,
                         int* cand, int* sel, int* ordering, float* errors, float* val,
                         int* bestk, float* cand_err, const int* perm, hipStream_t s);
hipError_t launch_feature_order(const float* XT, int64_t n, int d, double* key, int* perm,
                                bool by_scale, hipStream_t s);
hipError_t launch_init_state(SelState* st, const Problem& pb, int* cand, int* sel, int* ordering,
                             float* errors, const int* perm, hipStream_t s);

}  // namespace snn
