// Shared host/device definitions of the kNN selection engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace snn {

// Sentinel of unfilled neighbour slots and initial minima: the reference's
// MAX_FLOAT_TYPE (speedynn/lib/types.h:9) rounded to float, i.e. FLT_MAX.
constexpr float kMaxF = 3.402823466e+38f;

// Loss codes, speedynn/lib/loss_functions.h:11-19.
enum Loss : int {
  kRmse = 0, kRmseNorm = 1, kMad = 2, kMadNorm = 3, kZeroOne = 4, kStd = 5,
  kRmseNormMadNormStd = 6, kRmseNormMadNormGalaxies = 7, kRmseStd = 8
};

// Device-resident selection state.  Every per-step kernel reads it, so a step
// needs no host round trip and can be captured in a graph.
struct SelState {
  int step;       // completed steps
  int stopped;    // auto-stop fired or no candidate left
  int ncand;      // candidates of the current step
  int fstar;      // feature chosen by the last completed step
  int best_k;     // its optimal k
  int num_final;  // features after the last completed step
  int error;      // 1: no candidate had a finite error
  int pos;        // record position of the next step
  int pending;    // fstar chosen but not yet applied to D
  int fslot;      // candidate slot of fstar in the step that chose it
  int nonfinite;  // X holds NaN or infinity (the sorted first step does not apply)
  int fdefer;     // feature whose term D in memory still holds (deferred write; the
                  // search's SearchProblem::defer says whether it applies)
};

// Fixed description of one selection problem on the device.
struct Problem {
  int64_t n;      // points
  int64_t npf;    // points per fold = n / k_fold
  int d;          // features
  int cpad;       // candidate slots, multiple of 64
  int nk;         // number of k values
  int nl;         // number of losses
  int kmax;       // max k value
  int method;     // 1 forward, 0 backward
  int model_type; // 0 regression, 1 classification
  int auto_feature;
  int max_steps;  // steps to run at most
  int num_desired;
};

// Per-row lists of the points a pruned search may visit, written by the
// distance update that precedes the search: for row q the entries
// e < count[q] of idx/val[q * cap + e] are the eligible points j (outside q's
// fold), ascending, whose lower bound (forward D_qj; backward fl(D_qj - M_qj),
// kept in lb) is at most bound[q]; val holds D_qj.  bound[q] is the largest
// warm bound of the candidate blocks the list serves (leading blocks whose
// bound is far above the rest's are left out, so that one loose block does
// not make every list too long); a search wave whose bound exceeds bound[q]
// scans the row instead.  count[q] = -1: no list (no finite warm bound, or
// more than cap points), scan the row.
struct RowList {
  int* idx;
  float* val;
  float* lb;  // backward only (null forward: the lower bound is val)
  int* count;
  float* bound;
  int64_t cap;
};

// One selection problem as the search kernels see it (passed by value to a
// single-problem launch, or as an array to the persistent batched launch).
struct SearchProblem {
  const float* D;        // [n][n] current distances
  const float* M;        // [n][n] max single-feature term (backward pruning)
  const float* Xc;       // [n][cpad] candidate columns
  const float* y;        // [n] labels
  const int* kvals;      // [nk]
  const SelState* st;    // device state
  const int* seeds;      // [n][128] pruning seeds
  float* P;              // [nk][cpad][n] predictions
  unsigned long long* visited;  // point evaluations performed (pruned search): kVisitSlots
                                // counters kVisitStride apart, see count_visits
  int64_t n;
  int64_t npf;           // points per fold
  int cpad;
  int nk;
  int kmax;
  int mult;              // +1 forward, -1 backward
  int model_type;
  int zero;              // D is known to be 0 (first forward step)
  int sorted;            // first forward step: sorted columns serve finite X
  int nseed;             // backward pruned search: seeds read for the bound (<= 128)
  const float* skey;     // [d][n] each feature's values, ascending
  const int* sidx;       // [d][n] their point indices
  const int* cand;       // [d] feature of each candidate slot
  // Pruned searches: the launch covers queries q0 + i * qstep, i < nq (> 0);
  // qstep = 0: every point.  The autotuner's trial runs take disjoint parts.
  int64_t q0, nq, qstep;
  // Warm bounds (pruned searches; null tauw: none).  tauw[q][r] is the K-th
  // distance of the feature of rank r (iperm) for query q at the last step
  // that searched it (kMaxF: none); forward bounds add the largest square the
  // feature chosen since can contribute, from frange = {min, max} of its
  // column in XT.
  float* tauw;
  const int* iperm;
  const float* XT;
  const float* frange;
  int d;
  RowList rows;  // short lists instead of D (and M) rows where count >= 0 (idx null: none)
  // Candidate blocks a single-problem launch covers: those holding live
  // candidates this step (every step moves exactly one feature, so step s of
  // a selection over d features has d - s); 0: all cpad / 64.  Launching only
  // these keeps workgroups from holding LDS for waves with nothing to do.
  int nblk;
  // Backward pruned searches: D in memory still holds the term of feature
  // st->fdefer (the update before this search deferred its write); every D
  // value read from the rows is first reduced by it, fl(D - fl(t^2)) with
  // t = x_q - x_j of that feature, as the update would have written it.  The
  // row lists hold the reduced values already.
  int defer;
};

// Counters of the point evaluations a pruned search performs: one counter
// per wave would be as many same-address atomics as waves (a backward step
// at n = 32768 launches half a million), which serialise at one L2 channel;
// waves add to one of kVisitSlots counters, each on its own 128-byte line,
// and the host sums them.
constexpr int kVisitSlots = 512;
constexpr int kVisitStride = 16;  // unsigned long longs per counter: 128 bytes
__device__ __forceinline__ void count_visits(unsigned long long* visited, unsigned int wave_id,
                                             unsigned long long v) {
  atomicAdd(visited + (size_t)(wave_id & (kVisitSlots - 1)) * kVisitStride, v);
}

__host__ __device__ inline int live_blocks(const SearchProblem& a) {
  return a.nblk > 0 ? a.nblk : a.cpad / 64;
}

// Queries a pruned-search launch covers, and the point of the i-th.
__host__ __device__ inline int64_t launch_queries(const SearchProblem& a) {
  return a.qstep ? a.nq : a.n;
}
__host__ __device__ inline int64_t query_point(const SearchProblem& a, int64_t i) {
  return a.qstep ? a.q0 + i * a.qstep : i;
}

// Neighbour-list length compiled for a largest k value (-1: above 64).
__host__ __device__ inline int search_list_size(int kmax) {
  constexpr int sizes[] = {2, 4, 5, 8, 10, 15, 16, 20, 24, 32, 50, 64};
  for (int v : sizes)
    if (kmax <= v) return v;
  return -1;
}

// Wave-uniform value (SGPR) of a value known to be uniform.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Pruning seeds of one query, kept by one wave: each lane holds the two
// smallest D it was offered (b0 <= b1, point indices i0, i1; -1 and +inf when
// fewer).  offer() skips the query's fold [flo, fhi) and keeps the earlier
// point among equal values.
struct SeedPair {
  float b0, b1;
  int i0, i1;
  __device__ __forceinline__ void reset() {
    b0 = b1 = __int_as_float(0x7f800000);
    i0 = i1 = -1;
  }
  __device__ __forceinline__ void offer(float v, int64_t j, int64_t flo, int64_t fhi) {
    if (j >= flo && j < fhi) return;
    offer_eligible(v, j);
  }
  // j known to lie outside the fold
  __device__ __forceinline__ void offer_eligible(float v, int64_t j) {
    if (v < b1) {
      if (v < b0) {
        b1 = b0;
        i1 = i0;
        b0 = v;
        i0 = (int)j;
      } else {
        b1 = v;
        i1 = (int)j;
      }
    }
  }
  // The wave's 128 entries written to row[0..128) by ascending D (ties by
  // slot: lane k's pair is slots k and 64 + k), on a total order of the float
  // bits, so the search can stop at the first seed whose D is too large.
  __device__ __forceinline__ void write_ranked(int* row) const {
    const int lane = threadIdx.x & 63;
    auto sortable = [](float v) {
      const unsigned int b = __float_as_uint(v);
      return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    };
    const unsigned int k0 = sortable(b0), k1 = sortable(b1);
    int r0 = 0, r1 = 0;
    for (int k = 0; k < 64; ++k) {
      const unsigned int a0 = __shfl(k0, k), a1 = __shfl(k1, k);  // slots k and 64 + k
      r0 += (a0 < k0 || (a0 == k0 && k < lane)) ? 1 : 0;
      r0 += a1 < k0 ? 1 : 0;
      r1 += a0 <= k1 ? 1 : 0;
      r1 += (a1 < k1 || (a1 == k1 && k < lane)) ? 1 : 0;
    }
    row[r0] = i0;
    row[r1] = i1;
  }
};

}  // namespace snn
