// Candidate search kernels (c) of the kNN selection engine, gfx950 wave64.
// See kernels.hip for the data layout.  The work of one (query block,
// candidate block) or one (query, candidate block) pair is a device function,
// called by the single-problem kernels and by the persistent kernels that
// serve many problems from one work queue.
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "common.h"
#include "kernels.h"

// Waves per SIMD the forward per-wave pruned search is compiled for (the
// register budget of its body): lists up to 10 fit six waves (80 registers);
// longer ones run five (96 registers: at 80 they spill, which cost the
// n = 65536 forward selection 40% of its late-step search time).
#ifndef SNN_PRUNED_WAVES
#define SNN_PRUNED_WAVES(K) ((K) <= 10 ? 6 : 5)
#endif

namespace snn {

// ------------------------------------------------------ (c) candidate search
//
// Top-K insertion into a per-lane list sorted by (distance, index).  Points
// are offered in ascending index order, so a strict '<' gives the reference's
// tie rule (speedynn/lib/cpu.c:149-165: strict '>' entry, strict '<' bubble):
// an equal distance never displaces an earlier, lower index.  Branch-free
// network over compile-time slots, so the list lives in VGPRs.
// Callers pass only dd < Ld[K - 1] (so dd is not NaN).  Slot m's new distance
// is dd clamped to [Ld[m-1], Ld[m]] (one v_med3: the list is sorted), which is
// the value the compare-and-shift gives; its index follows the comparisons.
template <int K>
__device__ __forceinline__ void topk_insert(float (&Ld)[K], int (&Li)[K], float dd, int j) {
  bool c[K];  // dd < Ld[m] on the list as it stands
#pragma unroll
  for (int m = 0; m < K; ++m) c[m] = dd < Ld[m];
#pragma unroll
  for (int m = K - 1; m > 0; --m) {
    Li[m] = c[m - 1] ? Li[m - 1] : (c[m] ? j : Li[m]);
    Ld[m] = __builtin_amdgcn_fmed3f(Ld[m - 1], dd, Ld[m]);
  }
  if (c[0]) {
    Ld[0] = dd;
    Li[0] = j;
  }
}

// Prediction for the first k neighbours whose labels sit in the column
// `lab` (stride 64 floats per slot: LDS, or device memory in k_search_long).  Regression: nearest-first float sum / k
// (speedynn/lib/cpu.c:56-66).  Classification: get_mode_array
// (speedynn/lib/util.c:165-188): a label's count goes to its first
// occurrence, and the first position with the strictly largest count wins.
__device__ __forceinline__ float predict(const float* lab, int k, int model_type) {
  if (model_type == 0) {
    float s = 0.0f;
    for (int i = 0; i < k; ++i) s += lab[i * 64];
    return s / (float)k;
  }
  float best_v = lab[0];
  int best_cnt = 0;
  for (int i = 0; i < k; ++i) {
    const float v = lab[i * 64];
    bool first = true;
    int cnt = 0;
    for (int j = 0; j < k; ++j) {
      const bool eq = (j == i) || (lab[j * 64] == v);
      cnt += eq ? 1 : 0;
      if (j < i && eq) first = false;
    }
    if (first && cnt > best_cnt) {
      best_cnt = cnt;
      best_v = v;
    }
  }
  return best_v;
}

// Classification vote over the first k of the K neighbours' labels lv
// (k <= K), held in registers: the rule of predict(), every comparison on
// registers instead of LDS reads (lists up to 24, where the K x K compare
// network fits the kernels' register budgets).
template <int K>
__device__ __forceinline__ float vote_regs(const float (&lv)[K], int k) {
  float best_v = lv[0];
  int best_cnt = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < k) {
      const float v = lv[i];
      bool first = true;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        if (j < k) {
          const bool eq = (j == i) || (lv[j] == v);
          cnt += eq ? 1 : 0;
          if (j < i && eq) first = false;
        }
      }
      if (first && cnt > best_cnt) {
        best_cnt = cnt;
        best_v = v;
      }
    }
  }
  return best_v;
}

// Epilogue of one (query, candidate) lane: the prediction of every k value,
// written to P[m][c][q].  Classification with lists up to 24 votes on
// registers; otherwise the labels go to the wave's LDS column `lab`.
template <int K>
__device__ __forceinline__ void finish_query(const SearchProblem& a, const int (&Li)[K],
                                             float* lab, int64_t q, int c, bool live) {
  if constexpr (K <= 24) {
    if (a.model_type == 1) {
      float lv[K];
#pragma unroll
      for (int m = 0; m < K; ++m) lv[m] = a.y[Li[m]];
      for (int m = 0; m < a.nk; ++m) {
        const float pred = vote_regs<K>(lv, a.kvals[m]);
        if (live) a.P[((int64_t)m * a.cpad + c) * a.n + q] = pred;
      }
      return;
    }
  }
#pragma unroll
  for (int m = 0; m < K; ++m) lab[m * 64] = a.y[Li[m]];
  for (int m = 0; m < a.nk; ++m) {
    const float pred = predict(lab, a.kvals[m], a.model_type);
    if (live) a.P[((int64_t)m * a.cpad + c) * a.n + q] = pred;
  }
}

// Dense search of one item: queries [qb*WAVES*QW, +WAVES*QW) x candidate
// block cb, one lane per (query, candidate): a wave owns QW queries and the 64
// candidates.  Per chunk of 64 points the block stages the candidates' values
// Xc[j0:j0+64][cb*64:+64] and its queries' D values in LDS, double-buffered:
// the next chunk's loads are in flight while this one is scanned, and each
// staged value serves WAVES*QW queries.  Per point:
// d = fl(D_qj + mult*fl((x_qf - x_jf)^2)) (speedynn/lib/cpu.c:142-144); points
// in the query's fold are skipped (cpu.c:139: LOO when k_fold == n).  ZERO: D
// is known to be 0 (first forward step), so d = fl((x_qf - x_jf)^2) exactly
// and D is not read.
// A point whose d is below the lane's threshold (its list's K-th distance at
// the last flush) goes to the lane's pending buffer in LDS.  When some lane's
// buffer may overflow, the wave inserts every lane's entries in arrival
// order, i.e. ascending index, each tested strictly against the current list,
// so the lists equal those of offering every point in index order, and the
// K-slot insertion network runs once per entry of the busiest lane instead of
// once per point on which any of the 64 lanes inserts.
// Called by the whole block; returns early only block-uniformly.
constexpr int kDensePend = 12;  // pending (distance, index) slots per lane and query
constexpr int kDenseBatch = 4;  // points per scan batch (one 16-byte D read)

template <int K, int QW, int WAVES, bool ZERO>
constexpr int dense_lds_words() {
  constexpr int scan = 2 * 64 * 64 + (ZERO ? 0 : 2 * WAVES * QW * 64) +
                       WAVES * QW * 2 * kDensePend * 64;
  constexpr int epi = WAVES * K * 64;
  return scan > epi ? scan : epi;
}

template <int K, int QW, int WAVES, int MULT, bool ZERO>
__device__ __forceinline__ void dense_block(const SearchProblem& a, float* lds, int64_t qb,
                                            int cb) {
  constexpr int J = 64, S = kDensePend, B = kDenseBatch;
  constexpr int NQ = WAVES * QW;               // queries of the block
  constexpr int XV = J * 16 / (64 * WAVES);    // float4 of the X tile per thread
  static_assert(XV * 64 * WAVES == J * 16, "X tile split");
  static_assert(B == 4, "one float4 of D per batch");
  if (a.st->stopped) return;
  if (ZERO && a.sorted && !a.st->nonfinite) return;  // k_search_sorted serves this step
  const int ncand = a.st->ncand;
  if (cb * 64 >= ncand) return;
  const int64_t n = a.n, npf = a.npf;
  const int cpad = a.cpad;
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int c = cb * 64 + lane;
  const int64_t qbase = qb * NQ + wave * QW;  // this wave's first query

  float* xt = lds;                  // [2][J][64]: X tiles
  float* dt = lds + 2 * J * 64;     // [2][NQ][J]: D tiles
  float* pend = dt + (ZERO ? 0 : 2 * NQ * J) + wave * QW * (2 * S * 64) + lane;

  float xq[QW], thr[QW];
  float Ld[QW][K];
  int Li[QW][K];
  int pc[QW];
  int64_t flo[QW], fhi[QW];
#pragma unroll
  for (int qi = 0; qi < QW; ++qi) {
    const int64_t q = qbase + qi;
    xq[qi] = q < n ? a.Xc[q * cpad + c] : 0.0f;
    const int64_t fold = q / npf;
    flo[qi] = q < n ? fold * npf : n;
    fhi[qi] = q < n ? (fold + 1) * npf : n;
    thr[qi] = kMaxF;
    pc[qi] = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      Ld[qi][m] = kMaxF;
      Li[qi][m] = 0;
    }
  }

  // the next chunk's values, loaded at the top of an iteration and written to
  // LDS at its end
  float4 xr[XV];
  float dr[QW];
  const float* __restrict__ Xblk = a.Xc + cb * 64;
  auto fetch = [&](int64_t j0) {
#pragma unroll
    for (int v = 0; v < XV; ++v) {
      const int t = threadIdx.x + v * 64 * WAVES;
      const int64_t j = j0 + (t >> 4);
      xr[v] = j < n ? *reinterpret_cast<const float4*>(Xblk + j * cpad + (t & 15) * 4)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (!ZERO) {
      const int64_t j = j0 + lane;
#pragma unroll
      for (int qi = 0; qi < QW; ++qi) {
        const int64_t q = qbase + qi;
        dr[qi] = j < n ? a.D[(q < n ? q : n - 1) * n + j] : 0.0f;
      }
    }
  };
  auto commit = [&](int b) {
    float* dst = xt + b * (J * 64);
#pragma unroll
    for (int v = 0; v < XV; ++v) {
      const int t = threadIdx.x + v * 64 * WAVES;
      *reinterpret_cast<float4*>(dst + (t >> 4) * 64 + (t & 15) * 4) = xr[v];
    }
    if (!ZERO) {
#pragma unroll
      for (int qi = 0; qi < QW; ++qi) dt[(b * NQ + wave * QW + qi) * J + lane] = dr[qi];
    }
  };
  auto flush = [&](int qi) {
    int top = pc[qi];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) top = max(top, __shfl_xor(top, o));
    const float* pb = pend + qi * (2 * S * 64);
    for (int sl = 0; sl < top; ++sl) {
      if (sl < pc[qi]) {
        const float dd = pb[sl * 64];
        const int jj = __float_as_int(pb[(S + sl) * 64]);
        if (dd < Ld[qi][K - 1]) topk_insert<K>(Ld[qi], Li[qi], dd, jj);
      }
    }
    pc[qi] = 0;
    thr[qi] = Ld[qi][K - 1];
  };
  // one batch of B points from jj of the chunk at j0; CHECK: mask points
  // past n and in a query's fold
  auto batch = [&](const float* xs, const float* dsq, int64_t j0, int jj, bool check) {
    float xj[B];
#pragma unroll
    for (int u = 0; u < B; ++u) xj[u] = xs[(jj + u) * 64];
#pragma unroll
    for (int qi = 0; qi < QW; ++qi) {
      float dv[B] = {0.f, 0.f, 0.f, 0.f};
      if (!ZERO) {
        const float4 d4 = *reinterpret_cast<const float4*>(dsq + qi * J + jj);
        dv[0] = d4.x;
        dv[1] = d4.y;
        dv[2] = d4.z;
        dv[3] = d4.w;
      }
      float dd[B];
      bool p[B];
      bool any = false;
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const float t = xq[qi] - xj[u];
        const float pp = t * t;
        dd[u] = ZERO ? pp : (MULT > 0 ? dv[u] + pp : dv[u] - pp);
        p[u] = dd[u] < thr[qi];
        if (check) {
          const int64_t j = j0 + jj + u;
          p[u] = p[u] && j < n && (j < flo[qi] || j >= fhi[qi]);
        }
        any |= p[u];
      }
      if (__ballot(any)) {
        float* pb = pend + qi * (2 * S * 64);
#pragma unroll
        for (int u = 0; u < B; ++u) {
          if (p[u]) {
            pb[pc[qi] * 64] = dd[u];
            pb[(S + pc[qi]) * 64] = __int_as_float((int)(j0 + jj + u));
            ++pc[qi];
          }
        }
        if (__ballot(pc[qi] > S - B)) flush(qi);
      }
    }
  };

  __syncthreads();  // the previous item of a persistent block may still read lds
  fetch(0);
  commit(0);
  __syncthreads();
  int buf = 0;
  for (int64_t j0 = 0; j0 < n; j0 += J) {
    const bool more = j0 + J < n;
    if (more) fetch(j0 + J);
    const float* xs = xt + buf * (J * 64) + lane;
    const float* dsq = dt + (buf * NQ + wave * QW) * J;
    bool check = n - j0 < J;
#pragma unroll
    for (int qi = 0; qi < QW; ++qi) check |= (flo[qi] < j0 + J) && (fhi[qi] > j0);
    if (check) {
      for (int jj = 0; jj < J; jj += B) batch(xs, dsq, j0, jj, true);
    } else {
      for (int jj = 0; jj < J; jj += B) batch(xs, dsq, j0, jj, false);
    }
    if (more) commit(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int qi = 0; qi < QW; ++qi) flush(qi);
  __syncthreads();  // tiles and pending areas are read; the labels reuse them

  float* lab = lds + wave * (K * 64) + lane;
  const bool live = c < ncand;
#pragma unroll
  for (int qi = 0; qi < QW; ++qi) {
    const int64_t q = qbase + qi;
    if (q >= n) break;
    finish_query<K>(a, Li[qi], lab, q, c, live);
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Values-only sorted insertion (the bound needs the K-th value, not indices).
template <int K>
__device__ __forceinline__ void kth_insert(float (&B)[K], float v) {
#pragma unroll
  for (int m = K - 1; m > 0; --m) B[m] = v < B[m - 1] ? B[m - 1] : (v < B[m] ? v : B[m]);
  B[0] = v < B[0] ? v : B[0];
}

// Smallest float above a finite v.
__device__ __forceinline__ float next_up(float v) {
  const unsigned int b = __float_as_uint(v);
  if (b == 0x80000000u) return __uint_as_float(1u);  // -0: the smallest subnormal
  return __uint_as_float((int)b >= 0 ? b + 1u : b - 1u);
}

// Initial screen limit of a lane with bound tau (the K-th smallest exact
// distance over K eligible points, so no point with d > tau is in its final
// list and such points need not be offered; +inf: no bound, -inf: padding
// lane).  Forward screens compare fma(t, t, D) with tau widened by 2^-20 (plus
// 1e-36), which passes every d <= tau; backward tests are exact, d <= tau.
template <int MULT>
__device__ __forceinline__ float screen_limit(float tau) {
  const float kInf = __int_as_float(0x7f800000);
  if (tau == kInf || tau == -kInf) return tau;
  return MULT > 0 ? fmaf(tau, 1.00000095367431640625f, 1.0e-36f) : next_up(tau);
}

// Warm bound of candidate rank r for query q, from the K-th distance tau of
// the last step that searched it (+inf where none is known or X is not
// finite).  The K points of that list are eligible and distinct, so their
// distances now bound the new K-th distance:
//   backward: D only loses terms, D' = fl(D - q*) <= D, and by monotone
//     rounding fl(D' - t^2) <= fl(D - t^2) <= tau: tau itself is a bound;
//   forward: D' = fl(D + q*) with q* = fl((x_qg - x_jg)^2) <= Q, the square of
//     the feature g chosen since, largest over all j (from the column's min
//     and max; rounding is monotone), so fl(D' + t^2) <= (tau + Q)(1 + 3.01u)
//     (u = 2^-24; fl(D + t^2) <= tau gives D + t^2 <= tau(1 + u)): the bound
//     is fl(tau + Q) widened by 2^-20, plus 1e-36 for the subnormal range.
template <int MULT>
__device__ __forceinline__ float warm_bound(const SearchProblem& a, int64_t q, int r) {
  const float kInf = __int_as_float(0x7f800000);
  if (!a.tauw || a.st->nonfinite) return kInf;
  const float tau = a.tauw[q * a.d + r];
  if (!(tau < kMaxF)) return kInf;
  if (MULT < 0) return tau;
  const int g = a.st->fstar;
  if (g < 0) return kInf;
  const float xg = a.XT[(int64_t)g * a.n + q];
  const float lo = xg - a.frange[0], hi = a.frange[1] - xg;
  const float qm = fmaxf(lo * lo, hi * hi);
  const float w = fmaf(tau + qm, 1.00000095367431640625f, 1.0e-36f);
  return w < kInf ? w : kInf;
}

// Rank (iperm) of the feature in candidate slot c, 0 for padding slots.
__device__ __forceinline__ int cand_rank(const SearchProblem& a, int c, bool live) {
  return (a.tauw && live) ? a.iperm[a.cand[c]] : 0;
}

// Bound of one lane (candidate): the K-th smallest exact distance over the
// query's seeds, which k_seed_scan lists by ascending D.  Forward distances
// are >= D, so once the next seed's D exceeds every live lane's current K-th
// value no later seed can lower a bound, and the scan stops; a backward
// selection reads the first a.nseed seeds.  `warm` is a bound known already
// (warm_bound, +inf if none): the result is the smaller, and a forward scan
// also stops at a seed whose D exceeds every live lane's warm bound.  Returns
// +inf for an unfilled list without a warm bound, -inf for padding lanes.
//
// PIPE: batches of 8 seeds in two register sets, the next batch's index, D
// and value loads in flight while this batch's distances are inserted (up to
// 30 more registers; the persistent batch kernel keeps the one-set loop).
template <int K, int MULT, bool PIPE = true>
__device__ __forceinline__ float seed_bound(const SearchProblem& a, const float* __restrict__ Xl,
                                            const float* __restrict__ Dq, float xq, int64_t q,
                                            int cpad, bool live, float warm) {
  const float kInf = __int_as_float(0x7f800000);
  // backward: the warm bound is the previous step's K-th distance itself,
  // which the seeds rarely improve; where every live lane has one, the 32
  // seed gathers are skipped
  if (MULT < 0 && !__ballot(live && !(warm < kInf))) return live ? warm : -kInf;
  // a deferred term of D (SearchProblem::defer): subtracted from each seed's D
  const float* __restrict__ xpr =
      MULT < 0 && a.defer ? a.XT + (int64_t)a.st->fdefer * a.n : nullptr;
  const float xqp = xpr ? xpr[q] : 0.0f;
  auto dval = [&](int jj) {
    float v = Dq[jj];
    if (MULT < 0 && xpr) {
      const float t = xqp - xpr[jj];
      v = v - t * t;
    }
    return v;
  };
  float B[K];
#pragma unroll
  for (int m = 0; m < K; ++m) B[m] = kInf;
  const int* __restrict__ sq = a.seeds + q * 128;
  const int nseed = MULT > 0 ? 128 : a.nseed;  // a multiple of 8
  if constexpr (!PIPE) {
    for (int s0 = 0; s0 < nseed; s0 += 8) {
      int js[8];
      float xs[8], dqs[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        js[u] = sq[s0 + u];
        const int jj = js[u] < 0 ? 0 : js[u];
        xs[u] = Xl[(int64_t)jj * cpad];
        dqs[u] = dval(jj);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (js[u] >= 0) {
          const float t = xq - xs[u];
          const float dd = MULT > 0 ? dqs[u] + t * t : dqs[u] - t * t;
          if (dd < B[K - 1]) kth_insert<K>(B, dd);
        }
      }
      if (MULT > 0 && s0 + 8 < nseed) {
        const int jn = sq[s0 + 8];
        const float bmax = wave_max(live ? fminf(B[K - 1], warm) : -kInf);
        if (jn < 0 || Dq[jn] > bmax) break;  // seeds ascend in D; missing ones come last
      }
    }
    return live ? fminf(B[K - 1], warm) : -kInf;
  }
  int js[2][8];
  float xs[2][8], dqs[2][8];
  auto load = [&](int s0, int (&j_)[8], float (&x_)[8], float (&d_)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      j_[u] = sq[s0 + u];
      const int jj = j_[u] < 0 ? 0 : j_[u];
      x_[u] = Xl[(int64_t)jj * cpad];
      d_[u] = dval(jj);
    }
  };
  auto insert = [&](const int (&j_)[8], const float (&x_)[8], const float (&d_)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (j_[u] >= 0) {
        const float t = xq - x_[u];
        const float dd = MULT > 0 ? d_[u] + t * t : d_[u] - t * t;
        if (dd < B[K - 1]) kth_insert<K>(B, dd);
      }
    }
  };
  // seeds ascend in D and missing ones (-1) come last: a forward scan stops
  // at a seed whose D exceeds every live lane's current K-th value
  auto stop = [&](const int (&j_)[8], const float (&d_)[8]) {
    if (MULT < 0) return false;
    const float bmax = wave_max(live ? fminf(B[K - 1], warm) : -kInf);
    return j_[0] < 0 || d_[0] > bmax;
  };
  load(0, js[0], xs[0], dqs[0]);
  for (int s0 = 0; s0 < nseed; s0 += 16) {
    const bool more1 = s0 + 8 < nseed;
    if (more1) load(s0 + 8, js[1], xs[1], dqs[1]);
    __builtin_amdgcn_sched_barrier(0);
    insert(js[0], xs[0], dqs[0]);
    if (!more1 || stop(js[1], dqs[1])) break;
    const bool more2 = s0 + 16 < nseed;
    if (more2) load(s0 + 16, js[0], xs[0], dqs[0]);
    __builtin_amdgcn_sched_barrier(0);
    insert(js[1], xs[1], dqs[1]);
    if (!more2 || stop(js[0], dqs[0])) break;
  }
  return live ? fminf(B[K - 1], warm) : -kInf;
}

// Top-K insertion ordered by (distance, index) whatever the offering order:
// the list ends as the K smallest pairs, which is what the ascending-index
// scan with strict comparisons produces.
__device__ __forceinline__ bool pair_less(float a, int ja, float b, int jb) {
  return a < b || (a == b && ja < jb);
}

template <int K>
__device__ __forceinline__ void topk_insert_pair(float (&Ld)[K], int (&Li)[K], float dd, int j) {
#pragma unroll
  for (int m = K - 1; m > 0; --m) {
    const bool cp = pair_less(dd, j, Ld[m - 1], Li[m - 1]);
    const bool ch = pair_less(dd, j, Ld[m], Li[m]);
    const float nd = cp ? Ld[m - 1] : (ch ? dd : Ld[m]);
    const int ni = cp ? Li[m - 1] : (ch ? j : Li[m]);
    Ld[m] = nd;
    Li[m] = ni;
  }
  if (pair_less(dd, j, Ld[0], Li[0])) {
    Ld[0] = dd;
    Li[0] = j;
  }
}

// Pending insertions of one lane: up to pend_slots<K>() (distance, index)
// pairs that passed the exact test against the lane's list as it stood, held
// in the wave's LDS area (slot s of lane l at word s*64 + l, distances then
// indices), so the K-slot insertion network runs once per pending entry of
// the busiest lane at a flush instead of once per point on which any of the
// 64 lanes inserts.  A flush inserts a lane's entries in arrival order, which
// is ascending point index.  A visit batch pushes at most 8 entries per lane,
// so a flush comes when a lane holds more than slots - 8: 12 slots where the
// list length already limits the kernel to six waves per SIMD (K >= 15), 10
// below, so that LDS still admits eight 4-wave blocks per CU.
template <int K>
constexpr int pend_slots() {
  return K >= 15 ? 12 : 10;
}

// Words of LDS one wave of the per-wave pruned kernels uses: the pending
// buffer during the scan, the epilogue labels (K slots x 64 lanes) after it,
// and then the ring of visited points (kRing (offset, D) pairs from word
// ring_offset<K>(), then their kRing point indices).
constexpr int kRing = 128;
template <int K>
constexpr int ring_offset() {
  return 64 * (K > 2 * pend_slots<K>() ? K : 2 * pend_slots<K>());
}
template <int K>
constexpr int wave_lds_words() {
  return ring_offset<K>() + 3 * kRing;
}

// Flush of the pending buffer `buf` (the lane's slot 0; `pc` entries): insert
// the entries in arrival order, then tighten the lane's screen limit (forward:
// widened for the fma screen; backward: exact) and its bound thr, and return
// the wave's bound, the largest thr.  `top` is the wave's largest pc.
template <int K, int MULT>
__device__ __forceinline__ float flush_pending(const float* buf, int& pc, float (&Ld)[K],
                                               int (&Li)[K], float& lim, float& thr, float T,
                                               bool live) {
  constexpr float kScreen = 1.00000095367431640625f;  // 1 + 2^-20
  constexpr float kTiny = 1.0e-36f;
  int top = pc;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) top = max(top, __shfl_xor(top, o));
  for (int sl = 0; sl < top; ++sl) {
    if (sl < pc) {
      const float dd = buf[sl * 64];
      const int jj = __float_as_int(buf[(pend_slots<K>() + sl) * 64]);
      if (dd < Ld[K - 1]) topk_insert<K>(Ld, Li, dd, jj);
    }
  }
  pc = 0;
  if (top == 0) return T;
  if (Ld[K - 1] < kMaxF)
    lim = fminf(lim, MULT > 0 ? fmaf(Ld[K - 1], kScreen, kTiny) : Ld[K - 1]);
  thr = live ? fminf(thr, Ld[K - 1]) : -__int_as_float(0x7f800000);
  return wave_max(thr);
}

// Pruned search of one (query q, candidate block cb) item by one wave
// (precision form: forward steps after the first, every backward step).
//
// Bound.  For a candidate f, the K-th smallest d_qjf over any K distinct
// eligible points is an upper bound tau_f of its final K-th distance; here it
// is the smaller of the K-th smallest over the 128 seeds of k_seed_scan,
// evaluated exactly, and the warm bound from the previous step's K-th
// distance (warm_bound).
// Each point has a lower bound lb_qj of d_qjf valid for every candidate:
//   forward  (MULT = +1): d = fl(D + fl(t^2)) >= D, so lb = D;
//   backward (MULT = -1): M_qj >= fl(t^2) for every feature (max single-
//     feature term, computed with the full matrix), so d = fl(D - fl(t^2)) >=
//     fl(D - M) = lb by monotone rounding.
// A point with lb > T = max over the wave's candidates of tau_f can enter no
// lane's final top-K and is skipped; T tightens to the lists' own K-th
// distance as they fill.  Points are still offered in ascending index order,
// so every list equals the dense search's bit for bit.
//
// Per 64 points the wave takes the ballot of lb <= T outside the query's fold
// (the D/M values of the next four chunks are already in flight; a forward
// step whose update wrote the row's list (RowList) scans that list instead of
// the row: the same points that can pass, in the same order), then for
// each such point, ascending, gathers the candidates' values Xc[j][64cb+lane]
// (256 B, coalesced) and evaluates d.  Forward tests use a screen fma(t,t,D)
// against the K-th distance widened by 2^-20 (plus 1e-36 for the subnormal
// range): |fl(D+fl(t*t)) - fma(t,t,D)| <= 3.01*2^-24*(D+t*t) for D, t*t >= 0,
// so a rejected point cannot pass the exact test; passes recompute d in the
// reference's order.  Backward tests are exact (no screen: the fma error is
// not bounded relative to D - t^2).  Points that beat the list's last entry
// go to the lane's pending buffer (pend_slots<K>() slots in `buf`, the wave's
// LDS area offset by the lane), flushed into the list when a lane's buffer
// may overflow and at the end; the bound and screen limit tighten at each
// flush.
// No barriers.  Registers decide the speed: the search is latency-bound, and
// at K = 20 this body runs five waves per SIMD (SNN_PRUNED_WAVES).
template <int K, int MULT, bool PIPE = true>
__device__ __forceinline__ void pruned_wave(const SearchProblem& a, float* buf, int64_t q,
                                            int cb) {
  if (a.st->stopped) return;
  const int ncand = a.st->ncand;
  if (cb * 64 >= ncand) return;
  const int64_t n = a.n, npf = a.npf;
  const int cpad = a.cpad;
  const int lane = threadIdx.x & 63;
  const int c = cb * 64 + lane;
  const bool live = c < ncand;

  const float* __restrict__ Dq = a.D + q * n;
  const float* __restrict__ Mq = MULT > 0 ? nullptr : a.M + q * n;
  const float* __restrict__ Xl = a.Xc + cb * 64 + lane;
  const float xq = Xl[q * cpad];
  const int64_t flo = (q / npf) * npf;
  const int64_t fhi = flo + npf < n ? flo + npf : n;
  const float kInf = __int_as_float(0x7f800000);
  auto dist = [](float dq, float p) { return MULT > 0 ? dq + p : dq - p; };

  // the row list's count and bound, loaded before the seed bound needs the wave
  const int lcnt0 = a.rows.idx ? a.rows.count[q] : -1;
  const float lbound = a.rows.idx ? a.rows.bound[q] : -kInf;
  const int rk = cand_rank(a, c, live);
  float thr = seed_bound<K, MULT, PIPE>(a, Xl, Dq, xq, q, cpad, live,
                                        live ? warm_bound<MULT>(a, q, rk) : kInf);
  float T = wave_max(thr);

  float Ld[K];
  int Li[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    Ld[m] = kMaxF;
    Li[m] = 0;
  }
  // screen limit: d <= thr until the list is full, then its K-th distance;
  // padding lanes never pass (-inf)
  float lim = screen_limit<MULT>(thr);
  unsigned long long seen = 0;
  int pc = 0;  // this lane's pending entries

  auto flush = [&]() { T = flush_pending<K, MULT>(buf, pc, Ld, Li, lim, thr, T, live); };

  // D (and M) of 64 points per chunk, four chunks in flight in slots that
  // stay in place: the scan loop is unrolled by four, because rotating the
  // window through register copies would wait for the load just issued.  A
  // lane past the last point loads point n - 1 (its ballot bit is off): a
  // load that a branch may skip leaves the compiler unable to count the
  // loads in flight, and it waits for all of them.  The rows stream through
  // once (non-temporal), leaving the caches to Xc.
  const int n32 = (int)n;  // n < 2^31 (checked on the host)
  auto ldD = [&](int j) { return __builtin_nontemporal_load(Dq + (j < n32 ? j : n32 - 1)); };
  auto ldM = [&](int j) {
    return MULT < 0 ? __builtin_nontemporal_load(Mq + (j < n32 ? j : n32 - 1)) : 0.0f;
  };
  const uint32_t ucpad = (uint32_t)cpad;  // n * cpad < 2^31 (checked on the host)
  // Visits.  Each chunk's points to visit (the set bits of its ballot) are
  // appended, in ascending point order, to a ring of (offset j*cpad, D) pairs
  // in the wave's LDS area (and their indices j beside it) by one
  // lane-parallel store, and evaluated eight at a time from the ring's head:
  // four broadcast 16-byte LDS reads whose values the gathers and the screen
  // use as they are (no move to scalar registers), the gathers of the 64
  // candidates' values (256 B each, coalesced) and the screen; the index is
  // read only for a point that passes.  Batches stay full across chunk
  // boundaries.  At most 7 + 64 entries are pending, so the ring (128) never
  // overflows and full batches start at multiples of 8.  The order of
  // evaluation is the order of the points, so the pending buffer still
  // receives ascending indices.
  int2* __restrict__ ring = reinterpret_cast<int2*>(buf - lane + ring_offset<K>());
  int* __restrict__ ringj = reinterpret_cast<int*>(buf - lane + ring_offset<K>() + 2 * kRing);
  int rhead = 0, rtail = 0;  // wave-uniform: entries [rhead, rtail) are pending
  constexpr int BATCH = 8;
  // screen of a batch (cnt valid entries); the exact test runs only if some
  // lane may pass
  auto screen = [&](int base, const float (&dqs)[BATCH], float (&ts)[BATCH], int cnt) {
    bool pass[BATCH];
    bool any = false;
#pragma unroll
    for (int u = 0; u < BATCH; ++u) {
      ts[u] = xq - ts[u];
      const float scr = MULT > 0 ? fmaf(ts[u], ts[u], dqs[u]) : dist(dqs[u], ts[u] * ts[u]);
      pass[u] = u < cnt && scr < lim;
      any |= pass[u];
    }
    if (__ballot(any)) {
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        if (pass[u]) {
          const float dd = dist(dqs[u], ts[u] * ts[u]);
          if (dd < Ld[K - 1]) {
            buf[pc * 64] = dd;
            buf[(pend_slots<K>() + pc) * 64] = __int_as_float(ringj[(base + u) & (kRing - 1)]);
            ++pc;
          }
        }
      }
      if (__ballot(pc > pend_slots<K>() - BATCH)) flush();
    }
  };
  auto eval8 = [&]() {  // the next eight entries (rhead % 8 == 0)
    const int4* r4 = reinterpret_cast<const int4*>(ring + (rhead & (kRing - 1)));
    int4 e[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) e[h] = r4[h];  // broadcast reads, all issued first
    uint32_t os[BATCH];
    float dqs[BATCH], ts[BATCH];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      os[2 * h] = (uint32_t)e[h].x;
      dqs[2 * h] = __int_as_float(e[h].y);
      os[2 * h + 1] = (uint32_t)e[h].z;
      dqs[2 * h + 1] = __int_as_float(e[h].w);
    }
#pragma unroll
    for (int u = 0; u < BATCH; ++u) ts[u] = Xl[os[u]];
    const int base = rhead;
    rhead += BATCH;
    screen(base, dqs, ts, BATCH);
  };
  auto eval_tail = [&]() {  // the last (< 8) entries
    const int cnt = rtail - rhead;
    float dqs[BATCH], ts[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; ++u) {
      const int2 e = ring[(rhead + (u < cnt ? u : 0)) & (kRing - 1)];
      dqs[u] = __int_as_float(e.y);
      ts[u] = u < cnt ? Xl[(uint32_t)e.x] : 0.0f;
    }
    const int base = rhead;
    rhead = rtail;
    screen(base, dqs, ts, cnt);
  };
  // a chunk's points to visit: lane l holds point jv with D value dcur, and
  // bit l of mask says whether to visit it
  auto visit = [&](int jv, float dcur, unsigned long long mask) {
    if ((mask >> lane) & 1ull) {
      const int r = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      const int pos = (rtail + r) & (kRing - 1);
      ring[pos] = make_int2((int)((uint32_t)jv * ucpad), __float_as_int(dcur));
      ringj[pos] = jv;
    }
    __builtin_amdgcn_wave_barrier();  // the stores above precede the reads of other lanes
    rtail += (int)__popcll(mask);
    while (rtail - rhead >= BATCH) eval8();
  };
  // the row's list serves this wave if its bound covers the wave's
  const int lcnt = T > lbound ? -1 : lcnt0;
  if (lcnt >= 0) {
    // the row's list (RowList): its entries ascend in point index, lie
    // outside q's fold and hold every point whose lower bound is <= T; the
    // same chunked scan with two chunks of 64 entries in flight
    const int64_t lbase = q * a.rows.cap;
    const int* __restrict__ li = a.rows.idx + lbase;
    const float* __restrict__ lv = a.rows.val + lbase;
    const float* __restrict__ ll = MULT > 0 ? lv : a.rows.lb + lbase;
    const int last = lcnt > 0 ? lcnt - 1 : 0;
    // cached loads: every candidate block of the query reads the same list
    auto ldI = [&](int e) { return li[e < lcnt ? e : last]; };
    auto ldV = [&](int e) { return lv[e < lcnt ? e : last]; };
    auto ldL = [&](int e) { return MULT > 0 ? 0.0f : ll[e < lcnt ? e : last]; };
    int jsl[2];
    float dsl[2], lsl[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      jsl[u] = ldI(64 * u + lane);
      dsl[u] = ldV(64 * u + lane);
      lsl[u] = ldL(64 * u + lane);
    }
    for (int g0 = 0; g0 < lcnt; g0 += 128) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = g0 + 64 * u + lane;
        const float lbe = MULT > 0 ? dsl[u] : lsl[u];
        const unsigned long long mask = __ballot(e < lcnt && lbe <= T);
        if (mask) {
          seen += __popcll(mask);
          visit(jsl[u], dsl[u], mask);
        }
        jsl[u] = ldI(e + 128);
        dsl[u] = ldV(e + 128);
        lsl[u] = ldL(e + 128);
      }
    }
  } else {
  // the row scan; DF: D in memory still holds a deferred term
  // (SearchProblem::defer), subtracted from every D value as it arrives
  auto scan_row = [&](auto df_tag) {
    constexpr bool DF = MULT < 0 && decltype(df_tag)::value;
    const float* __restrict__ xpr = DF ? a.XT + (int64_t)a.st->fdefer * n : nullptr;
    const float xqp = DF ? xpr[q] : 0.0f;
    auto ldP = [&](int j) { return DF ? xpr[j < n32 ? j : n32 - 1] : 0.0f; };
    float dsl[4], msl[4], psl[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dsl[u] = ldD(64 * u + lane);
      msl[u] = ldM(64 * u + lane);
      psl[u] = ldP(64 * u + lane);
    }
    const int flo32 = (int)flo, fhi32 = (int)fhi;
    for (int g0 = 0; g0 < n32; g0 += 256) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jl = g0 + 64 * u + lane;
        float dv = dsl[u];
        if (DF) {
          const float t = xqp - psl[u];
          dv = dv - t * t;
        }
        const float lb = MULT > 0 ? dv : dv - msl[u];
        const bool ok = jl < n32 && (jl < flo32 || jl >= fhi32) && lb <= T;
        const unsigned long long mask = __ballot(ok);
        if (mask) {
          seen += __popcll(mask);
          visit(jl, dv, mask);
        }
        // the slot's next chunk, issued once its value is dead, so the load
        // lands in the slot's own register
        dsl[u] = ldD(jl + 256);
        if (MULT < 0) msl[u] = ldM(jl + 256);
        if (DF) psl[u] = ldP(jl + 256);
      }
    }
  };
  if constexpr (MULT < 0) {
    if (a.defer) scan_row(std::true_type{});
    else scan_row(std::false_type{});
  } else {
    scan_row(std::false_type{});
  }
  }
  if (rtail > rhead) eval_tail();
  flush();
  if (a.visited && lane == 0) {
    const int nlive = ncand - cb * 64 < 64 ? ncand - cb * 64 : 64;
    count_visits(a.visited, blockIdx.x * 4 + (threadIdx.x >> 6),
                 seen * (unsigned long long)nlive);  // point evaluations performed
  }
  if (a.tauw && live) a.tauw[q * a.d + rk] = Ld[K - 1];  // the next step's warm bound
  finish_query<K>(a, Li, buf, q, c, live);
}

// First step of a forward selection (D = 0) over sorted columns: one lane per
// (sorted position p of candidate f's column, f).  The query is the point at
// p; walking outward, d = fl((x_q - x_j)^2) is non-decreasing on each side,
// so the two sides are merged by distance and a side stops at the first
// point whose d exceeds the list's K-th distance (points of equal distance
// still enter when their index is lower).  Points of the query's fold are
// passed over.  Consecutive lanes hold neighbouring positions, so their walks
// read the same cache lines.  Needs finite X (st->nonfinite == 0).
template <int K>
__global__ __launch_bounds__(256) void k_search_sorted(const SearchProblem a) {
  __shared__ float lds[4 * K * 64];
  const SelState* st = a.st;
  if (st->stopped || st->nonfinite) return;
  const int c = blockIdx.y;
  if (c >= st->ncand) return;
  const int f = a.cand[c];
  const int64_t n = a.n, npf = a.npf;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool live = p < n;
  const int64_t pp = live ? p : n - 1;
  const float* __restrict__ key = a.skey + (int64_t)f * n;
  const int* __restrict__ idx = a.sidx + (int64_t)f * n;
  const float xq = key[pp];
  const int q = idx[pp];
  const int64_t flo = ((int64_t)q / npf) * npf, fhi = flo + npf;
  const float kInf = __int_as_float(0x7f800000);

  float Ld[K];
  int Li[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    Ld[m] = kMaxF;
    Li[m] = 0;
  }
  auto dist = [&](int64_t t) {
    const float u = xq - key[t];
    return u * u;
  };
  int64_t l = pp - 1, r = pp + 1;
  float dl = l >= 0 ? dist(l) : kInf, dr = r < n ? dist(r) : kInf;
  int jl = l >= 0 ? idx[l] : 0, jr = r < n ? idx[r] : 0;
  for (;;) {
    const float lim = Ld[K - 1];
    const bool lok = dl <= lim, rok = dr <= lim;
    if (!lok && !rok) break;
    const bool left = lok && (!rok || dl <= dr);
    const float dd = left ? dl : dr;
    const int j = left ? jl : jr;
    if (j < flo || j >= fhi) topk_insert_pair<K>(Ld, Li, dd, j);
    if (left) {
      --l;
      dl = l >= 0 ? dist(l) : kInf;
      jl = l >= 0 ? idx[l] : 0;
    } else {
      ++r;
      dr = r < n ? dist(r) : kInf;
      jr = r < n ? idx[r] : 0;
    }
  }
  const int wave = threadIdx.x >> 6;
  finish_query<K>(a, Li, lds + wave * (K * 64) + (threadIdx.x & 63), q, c, live);
}

// ----------------------------------------------------- single-problem kernels

template <int K, int QW, int WAVES, int MULT, bool ZERO>
__global__ __launch_bounds__(64 * WAVES) void k_search_dense(const SearchProblem a) {
  __shared__ __align__(16) float lds[dense_lds_words<K, QW, WAVES, ZERO>()];
  dense_block<K, QW, WAVES, MULT, ZERO>(a, lds, blockIdx.x, blockIdx.y);
}

// Every wave is an independent (query, 64-candidate block) item.  Blocks are
// dealt to the 8 XCDs round-robin; XCD x takes the x-th contiguous slice of
// the queries.  cb_major: all queries of a slice for candidate block 0, then
// block 1, ... -- concurrent waves gather from one 64-column slice of Xc
// (measured 1.6x faster on the n=65536 forward selection).  Otherwise the
// nblk blocks of one query are consecutive waves that read the query's D (and
// M) row once into L2.
template <int K, int MULT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(K <= 20 && MULT > 0 ? SNN_PRUNED_WAVES(K) : 1)))
void k_search_pruned(const SearchProblem a, int cb_major) {
  __shared__ float lds[4 * wave_lds_words<K>()];  // pending buffers, then epilogue labels
  const int wave = uniform(threadIdx.x >> 6);
  const int nblk = live_blocks(a);
  const int64_t nq = launch_queries(a);
  const int64_t per_xcd = (nq + 7) / 8;
  const int64_t local = (int64_t)(blockIdx.x / 8) * 4 + wave;
  const int64_t qi = cb_major ? local % per_xcd : local / nblk;
  const int cb = (int)(cb_major ? local / per_xcd : local % nblk);
  const int64_t i = (int64_t)(blockIdx.x % 8) * per_xcd + qi;
  if (i >= nq || cb >= nblk || qi >= per_xcd) return;  // whole wave; no barriers
  pruned_wave<K, MULT>(a, lds + wave * wave_lds_words<K>() + (threadIdx.x & 63),
                       query_point(a, i), cb);
}

// Pruned search with the row scan shared by the candidate blocks of a query:
// one workgroup per query, wave w owns candidate block w (waves past the
// last block only help scanning).  Bounds, the exact path and the pending
// buffers (one LDS area per wave) are those of pruned_wave.  Per chunk of R*64*W points each lane tests R points against
// the largest bound of the workgroup's waves; the passing points go, in
// ascending index order, to a queue in LDS (index, D, lower bound), and every
// wave visits the queue, skipping entries above its own bound.  The row is
// read once per query instead of once per candidate block.  Bounds only
// shrink, so a wave reading another wave's bound while it is being lowered
// still gets a valid (larger) bound.  Barriers: counts written -> queue
// written -> queue read (the next chunk's first barrier).
template <int K, int MULT, int W>
__global__ __launch_bounds__(64 * W) void k_search_pruned_shared(const SearchProblem a) {
  constexpr int R = W >= 16 ? 2 : 4;  // points per lane per chunk (16 waves: 128 registers)
  constexpr int SUB = 64 * W;
  constexpr int CH = R * SUB;
  constexpr int LDS_Q = (MULT > 0 ? 2 : 3) * CH;  // queue: index, D, lower bound (backward)
  // epilogue labels, 4 waves at a time (only for lists longer than a wave's
  // pending area, which otherwise holds them)
  constexpr int LDS_E = K <= 2 * pend_slots<K>() ? 0 : 4 * K * 64;
  __shared__ float lds[LDS_Q > LDS_E ? LDS_Q : LDS_E];
  __shared__ float pend[W * 2 * pend_slots<K>() * 64];  // pending buffers, one area per wave
  __shared__ float Tw[W];
  __shared__ int wcount[R * W];
  int* qj = reinterpret_cast<int*>(lds);
  float* qd = lds + CH;
  float* qlb = lds + 2 * CH;  // backward only: forward lower bounds are the D values

  const SelState* st = a.st;
  if (st->stopped) return;
  const int ncand = st->ncand;
  const int n = (int)a.n;  // n < 2^31 (checked on the host)
  const int q = (int)query_point(a, blockIdx.x);
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // several workgroups per query split its candidate blocks: blockIdx.y = part,
  // so all queries' first parts run before the second parts and the gathers
  // of each half of the run stay within half of Xc's columns
  const int cb = wave + W * blockIdx.y;
  if ((int)blockIdx.y * W * 64 >= ncand) return;  // no live block in this part: whole workgroup
  const bool active = cb * 64 < ncand && cb < live_blocks(a);
  const int c = cb * 64 + lane;
  const bool live = active && c < ncand;
  const int cpad = a.cpad;
  const int npf = (int)a.npf;
  const int flo = (q / npf) * npf;
  const int fspan = flo + npf < n ? npf : n - flo;  // fold of q: [flo, flo + fspan)
  const float* __restrict__ Dq = a.D + (int64_t)q * n;
  const float* __restrict__ Mq = MULT > 0 ? nullptr : a.M + (int64_t)q * n;
  const float* __restrict__ Xl = a.Xc + (active ? cb * 64 : 0) + lane;
  const float kInf = __int_as_float(0x7f800000);
  auto dist = [](float dq, float p) { return MULT > 0 ? dq + p : dq - p; };
  const float xq = active ? Xl[(int64_t)q * cpad] : 0.0f;

  const int rk = cand_rank(a, c, live);
  float thr = active ? seed_bound<K, MULT>(a, Xl, Dq, xq, q, cpad, live,
                                           live ? warm_bound<MULT>(a, q, rk) : kInf)
                     : -kInf;
  float T = wave_max(thr);
  if (lane == 0) Tw[wave] = T;

  float Ld[K];
  int Li[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    Ld[m] = kMaxF;
    Li[m] = 0;
  }
  float lim = screen_limit<MULT>(thr);  // as in pruned_wave
  unsigned long long seen = 0;
  float* buf = pend + wave * (2 * pend_slots<K>() * 64) + lane;
  int pc = 0;  // this lane's pending entries
  auto flush = [&]() {
    T = flush_pending<K, MULT>(buf, pc, Ld, Li, lim, thr, T, live);
    if (lane == 0) Tw[wave] = T;
  };

  const int n32 = (int)n;  // n < 2^31 (checked on the host)
  auto ldD = [&](int j) { return j < n32 ? __builtin_nontemporal_load(Dq + j) : kInf; };
  auto ldM = [&](int j) {
    return (MULT < 0 && j < n32) ? __builtin_nontemporal_load(Mq + j) : 0.0f;
  };
  // a deferred term of D (SearchProblem::defer), subtracted as D arrives
  const bool df = MULT < 0 && a.defer;
  const float* __restrict__ xpr = df ? a.XT + (int64_t)a.st->fdefer * n : nullptr;
  const float xqp = df ? xpr[q] : 0.0f;
  auto ldP = [&](int j) { return df && j < n32 ? xpr[j] : 0.0f; };
  const int t0 = threadIdx.x;
  float dn[R], mn[R], pn[R];  // the next chunk's D (and M), loaded one chunk ahead
#pragma unroll
  for (int r = 0; r < R; ++r) {
    dn[r] = ldD(r * SUB + t0);
    mn[r] = ldM(r * SUB + t0);
    pn[r] = ldP(r * SUB + t0);
  }
  const uint32_t ucpad = (uint32_t)cpad;  // n * cpad < 2^31 (checked on the host)
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += CH) {
    float Tmax = -kInf;
#pragma unroll
    for (int w = 0; w < W; ++w) Tmax = fmaxf(Tmax, Tw[w]);
    float dv[R], lbv[R];
    bool ok[R];
    unsigned long long m[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = c0 + r * SUB + t0;
      dv[r] = dn[r];
      if (df) {
        const float t = xqp - pn[r];
        dv[r] = dv[r] - t * t;
      }
      lbv[r] = MULT > 0 ? dv[r] : dv[r] - mn[r];
      dn[r] = ldD(j + CH);
      if (MULT < 0) mn[r] = ldM(j + CH);
      if (df) pn[r] = ldP(j + CH);
      ok[r] = j < n && (uint32_t)(j - flo) >= (uint32_t)fspan && lbv[r] <= Tmax;
      m[r] = __ballot(ok[r]);
      if (lane == 0) wcount[r * W + wave] = (int)__popcll(m[r]);
    }
    __syncthreads();
    // queue order (sub-chunk, wave, lane) is ascending point index
    int base[R], total = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      base[r] = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int cw = wcount[r * W + w];
        if (w == wave) base[r] = total;
        total += cw;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (ok[r]) {
        const int pos = base[r] + (int)__builtin_amdgcn_mbcnt_hi(
                                      (uint32_t)(m[r] >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)m[r], 0u));
        qj[pos] = c0 + r * SUB + t0;
        qd[pos] = dv[r];
        if (MULT < 0) qlb[pos] = lbv[r];
      }
    }
    __syncthreads();
    if (!active) continue;
    // the queue, 64 entries at a time: ballot of the entries within this
    // wave's bound, then the visit loop of pruned_wave over its set bits
    for (int e0 = 0; e0 < total; e0 += 64) {
      const int e = e0 + lane;
      const bool has = e < total;
      const int je = has ? qj[e] : 0;
      const float de = has ? qd[e] : 0.0f;
      const float lbe = has ? (MULT > 0 ? de : qlb[e]) : kInf;
      unsigned long long mask = __ballot(has && lbe <= T);
      if (!mask) continue;
      seen += __popcll(mask);
      while (mask) {
        constexpr int BATCH = 8;
        int js[BATCH];
        float dqs[BATCH], ts[BATCH];
        bool pass[BATCH];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
          if (mask) {
            const int b = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            js[u] = __builtin_amdgcn_readlane(je, b);
            dqs[u] = lane_value(de, b);
            ts[u] = Xl[(uint32_t)js[u] * ucpad];
            cnt = u + 1;
          }
        }
        bool any = false;
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
          ts[u] = xq - ts[u];
          const float scr = MULT > 0 ? fmaf(ts[u], ts[u], dqs[u]) : dist(dqs[u], ts[u] * ts[u]);
          pass[u] = u < cnt && scr < lim;
          any |= pass[u];
        }
        if (__ballot(any)) {
#pragma unroll
          for (int u = 0; u < BATCH; ++u) {
            if (pass[u]) {
              const float dd = dist(dqs[u], ts[u] * ts[u]);
              if (dd < Ld[K - 1]) {
                buf[pc * 64] = dd;
                buf[(pend_slots<K>() + pc) * 64] = __int_as_float(js[u]);
                ++pc;
              }
            }
          }
          if (__ballot(pc > pend_slots<K>() - BATCH)) flush();
        }
      }
    }
  }
  flush();
  if (a.visited && lane == 0 && active) {
    const int nlive = ncand - cb * 64 < 64 ? ncand - cb * 64 : 64;
    count_visits(a.visited, (blockIdx.x + blockIdx.y * gridDim.x) * W + wave,
                 seen * (unsigned long long)nlive);
  }
  if (a.tauw && live) a.tauw[(int64_t)q * a.d + rk] = Ld[K - 1];
  // epilogue: the labels go to the wave's own pending area (empty after the
  // last flush) when it holds K slots, so all waves finish at once; longer
  // lists take four waves at a time through the label area (aliases the queue)
  if constexpr (K <= 2 * pend_slots<K>()) {
    if (active) finish_query<K>(a, Li, buf, q, c, live);
  } else {
    for (int g = 0; g < W; g += 4) {
      __syncthreads();
      if (active && wave >= g && wave < g + 4)
        finish_query<K>(a, Li, lds + (wave - g) * (K * 64) + lane, q, c, live);
    }
  }
}

// Second forward step (one feature g chosen, so D_qj = fl((x_qg - x_jg)^2)
// exactly, the first step's update of D = 0): walking outward from the
// query's position in g's sorted column visits the points in ascending D
// (non-decreasing on each side, the two sides merged), so the wave's bound
// tightens after few points and the walk stops at the first point with
// D > T = the largest K-th distance of the wave's live lanes (d >= D).  No D
// row is read.  Insertion is by (distance, index), so the walk order does not
// change the lists.  Windows of 64 sorted positions per side sit in VGPRs and
// the next point is taken with readlane.  Needs finite X (st->nonfinite == 0;
// otherwise the wave runs pruned_wave).  The wave mapping is k_search_pruned's.
template <int K>
__global__ __launch_bounds__(256) void k_search_walk(const SearchProblem a,
                                                     const int* __restrict__ rank) {
  __shared__ float lds[4 * wave_lds_words<K>()];
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nblk = live_blocks(a);
  const int64_t nq = launch_queries(a);
  const int64_t per_xcd = (nq + 7) / 8;
  const int64_t local = (int64_t)(blockIdx.x / 8) * 4 + wave;
  const int64_t qi = local % per_xcd;
  const int cb = (int)(local / per_xcd);
  const int64_t i = (int64_t)(blockIdx.x % 8) * per_xcd + qi;
  if (i >= nq || cb >= nblk) return;  // whole wave; no barriers
  const int64_t q = query_point(a, i);
  float* lab = lds + wave * wave_lds_words<K>() + lane;
  const SelState* st = a.st;
  if (st->stopped) return;
  if (st->nonfinite) {
    pruned_wave<K, 1>(a, lab, q, cb);
    return;
  }
  const int ncand = st->ncand;
  if (cb * 64 >= ncand) return;
  const int n = (int)a.n;
  const int c = cb * 64 + lane;
  const bool live = c < ncand;
  const int g = st->fstar;  // the feature the first step chose
  const float* __restrict__ key = a.skey + (int64_t)g * n;
  const int* __restrict__ idx = a.sidx + (int64_t)g * n;
  const int pos = rank[q];
  const float xg = key[pos];
  const float* __restrict__ Xl = a.Xc + cb * 64 + lane;
  const uint32_t ucpad = (uint32_t)a.cpad;
  const float xq = Xl[(uint32_t)q * ucpad];
  const int npf = (int)a.npf;
  const int flo = ((int)q / npf) * npf;
  const int fspan = flo + npf < n ? npf : n - flo;
  const float kInf = __int_as_float(0x7f800000);
  constexpr float kScreen = 1.00000095367431640625f;  // 1 + 2^-20
  constexpr float kTiny = 1.0e-36f;

  float Ld[K];
  int Li[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    Ld[m] = kMaxF;
    Li[m] = 0;
  }
  const int rk = cand_rank(a, c, live);
  const float warm = live ? warm_bound<1>(a, q, rk) : -kInf;  // per lane; -inf: padding
  float lim = screen_limit<1>(warm);
  float T = wave_max(warm);
  unsigned long long seen = 0;
  // Windows of 64 sorted positions, one lane per position: left windows end
  // at lo - 1 (lane 63 nearest the query), right windows start at hi (lane 0
  // nearest).  A window's D values come from the sorted keys; its points
  // within T outside the query's fold are visited as in the row scan (a
  // ballot, then eight gathers in flight).  A side stops at the first window
  // whose nearest point has D > T, since D does not decrease outward.
  int lo = pos, hi = pos + 1;
  bool lact = lo > 0, ract = hi < n;
  while (lact || ract) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (side == 0 ? !lact : !ract) continue;
      const int p = side == 0 ? lo - 64 + lane : hi + lane;
      const bool inr = p >= 0 && p < n;
      const float kv = inr ? key[p] : 0.0f;
      const int jv = inr ? idx[p] : 0;
      const float tg = xg - kv;
      const float Dv = tg * tg;
      if (side == 0) {
        lo -= 64;
        lact = lo > 0;
      } else {
        hi += 64;
        ract = hi < n;
      }
      if (!(lane_value(Dv, side == 0 ? 63 : 0) <= T)) {  // nearest point beyond T
        if (side == 0) lact = false;
        else ract = false;
        continue;
      }
      const bool ok = inr && Dv <= T && (uint32_t)(jv - flo) >= (uint32_t)fspan;
      unsigned long long mask = __ballot(ok);
      seen += __popcll(mask);
      bool changed = false;
      while (mask) {
        constexpr int BATCH = 8;
        int js[BATCH];
        float dqs[BATCH], ts[BATCH];
        bool pass[BATCH];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
          if (mask) {
            const int b = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            js[u] = __builtin_amdgcn_readlane(jv, b);
            dqs[u] = lane_value(Dv, b);
            ts[u] = Xl[(uint32_t)js[u] * ucpad];
            cnt = u + 1;
          }
        }
        bool any = false;
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
          ts[u] = xq - ts[u];
          pass[u] = u < cnt && fmaf(ts[u], ts[u], dqs[u]) < lim;
          any |= pass[u];
        }
        if (__ballot(any)) {
#pragma unroll
          for (int u = 0; u < BATCH; ++u) {
            if (pass[u]) {
              const float dd = dqs[u] + ts[u] * ts[u];
              if (pair_less(dd, js[u], Ld[K - 1], Li[K - 1])) {
                topk_insert_pair<K>(Ld, Li, dd, js[u]);
                if (Ld[K - 1] < kMaxF) lim = fminf(lim, fmaf(Ld[K - 1], kScreen, kTiny));
                changed = true;
              }
            }
          }
        }
      }
      if (__ballot(changed)) T = wave_max(live ? fminf(Ld[K - 1], warm) : -kInf);
    }
  }
  if (a.visited && lane == 0) {
    const int nlive = ncand - cb * 64 < 64 ? ncand - cb * 64 : 64;
    count_visits(a.visited, blockIdx.x * 4 + wave, seen * (unsigned long long)nlive);
  }
  if (a.tauw && live) a.tauw[q * a.d + rk] = Ld[K - 1];
  finish_query<K>(a, Li, lab, q, c, live);
}

// rank[sidx[g][p]] = p for the feature g the first step chose.
__global__ __launch_bounds__(256) void k_rank(const int* __restrict__ sidx,
                                              const SelState* __restrict__ st,
                                              int* __restrict__ rank, int64_t n) {
  if (st->stopped || st->nonfinite) return;
  const int* col = sidx + (int64_t)st->fstar * n;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x)
    rank[col[p]] = (int)p;
}

// ------------------------------------------------------ persistent batch kernels
//
// Many independent problems per launch: a global counter hands out items
// (problem, query block or query, candidate block) in problem order; the item
// ranges of the problems are in `start` (prefix sums, P+1 entries).  The
// problems of one launch share the list size K and the direction: registers
// are allocated per kernel for its longest list, so one kernel for all list
// sizes would run the short lists at the long list's occupancy.  The counter
// is reset by a kernel before every launch (no memset nodes in a captured
// graph).

__device__ __forceinline__ int find_problem(const int64_t* start, int P, int64_t item) {
  int lo = 0, hi = P;  // largest p with start[p] <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Dense items of a problem: (candidate block, query block of QW*WAVES
// queries), candidate-block major; the body of k_search_dense.
template <int K, int QW, int WAVES, int MULT, bool ZERO>
__global__ __launch_bounds__(64 * WAVES) void k_batch_dense(
    const SearchProblem* __restrict__ probs, const int64_t* __restrict__ start, int P,
    unsigned long long* __restrict__ counter) {
  __shared__ __align__(16) float lds[dense_lds_words<K, QW, WAVES, ZERO>()];
  __shared__ long long s_item;
  const int64_t total = start[P];
  for (;;) {
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t item = uniform64(s_item);
    __syncthreads();
    if (item >= total) break;
    const int p = find_problem(start, P, item);
    const SearchProblem a = probs[p];
    const int64_t local = item - start[p];
    const int64_t qblocks = (a.n + QW * WAVES - 1) / (QW * WAVES);
    dense_block<K, QW, WAVES, MULT, ZERO>(a, lds, local % qblocks, (int)(local / qblocks));
  }
}

// Pruned items of a problem: (candidate block, query), candidate-block major;
// one wave per item, no barriers; the register budget of k_search_pruned.
template <int K, int MULT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(K <= 20 && MULT > 0 ? SNN_PRUNED_WAVES(K) : 1)))
void k_batch_pruned(const SearchProblem* __restrict__ probs, const int64_t* __restrict__ start,
                    int P, unsigned long long* __restrict__ counter) {
  __shared__ float lds[4 * wave_lds_words<K>()];
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  float* lab = lds + wave * wave_lds_words<K>() + lane;
  const int64_t total = start[P];
  for (;;) {
    unsigned long long it = 0;
    if (lane == 0) it = atomicAdd(counter, 1ull);
    const int64_t item = uniform64(__shfl((long long)it, 0));
    if (item >= total) break;
    const int p = find_problem(start, P, item);
    const SearchProblem a = probs[p];
    const int64_t local = item - start[p];
    pruned_wave<K, MULT, false>(a, lab, local % a.n, (int)(local / a.n));
  }
}

__global__ void k_reset_counter(unsigned long long* c) { *c = 0ull; }

// Seeds of the pruning bound: for each query, lane L of one wave keeps the two
// smallest eligible D_qj over the residue class j = L (mod 64), ties to the
// lower index; 128 distinct points, -1 where a class has fewer than two.
__global__ __launch_bounds__(256) void k_seed_scan(const float* __restrict__ D,
                                                   const SelState* __restrict__ st,
                                                   int* __restrict__ seeds, int64_t n,
                                                   int64_t npf) {
  if (st->stopped) return;
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6);
  if (q >= n) return;
  const float* __restrict__ Dq = D + q * n;
  const int64_t flo = (q / npf) * npf;
  const int64_t fhi = flo + npf < n ? flo + npf : n;
  SeedPair sp;
  sp.reset();
  int64_t j = lane;
  for (; j + 7 * 64 < n; j += 8 * 64) {  // eight loads in flight per lane
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(Dq + j + u * 64);
#pragma unroll
    for (int u = 0; u < 8; ++u) sp.offer(v[u], j + u * 64, flo, fhi);
  }
  for (; j < n; j += 64) sp.offer(Dq[j], j, flo, fhi);
  sp.write_ranked(seeds + q * 128);
}

// Search with lists longer than the register kernels compile (k values above
// 64; the reference takes any k, speedynn/lib/cpu.c:50-51).  One lane per
// (query, candidate) as in dense_block, but the list of K (distance, index)
// pairs lives in global memory, slot m of a lane at [wave][m][lane], and a
// point enters by the reference's own steps (cpu.c:149-165): strict '<'
// against the last slot, then it moves toward the front while strictly below
// its neighbour, so equal distances keep the lower index.  The labels of the
// final list go beside it for predict().  Persistent: wave w of W serves the
// items (query, candidate block) w, w + W, ...
__global__ __launch_bounds__(256) void k_search_long(SearchProblem a, int K, float* __restrict__ ld,
                                                     int* __restrict__ li, float* __restrict__ lab_g,
                                                     int64_t waves_total) {
  if (a.st->stopped) return;
  const int ncand = a.st->ncand;
  const int nblk = (ncand + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wid >= waves_total) return;
  const int64_t base = wid * (int64_t)K * 64 + lane;
  float* Ld = ld + base;  // slot m at Ld[m * 64]
  int* Li = li + base;
  float* lab = lab_g + base;
  const int64_t n = a.n, npf = a.npf;
  const int cpad = a.cpad;
  const int64_t items = n * nblk;
  for (int64_t it = wid; it < items; it += waves_total) {
    const int64_t q = it / nblk;
    const int c = (int)(it - q * nblk) * 64 + lane;
    const bool live = c < ncand;
    const float xq = a.Xc[q * cpad + c];
    for (int m = 0; m < K; ++m) {
      Ld[m * 64] = kMaxF;
      Li[m * 64] = 0;
    }
    float last = kMaxF;
    const int64_t flo = (q / npf) * npf, fhi = flo + npf;
    const float* __restrict__ Drow = a.D + q * n;
    const float* __restrict__ xc = a.Xc + c;
    for (int64_t j = 0; j < n; ++j) {
      if (j == flo) {  // the query's fold: cpu.c:139
        j = fhi - 1;
        continue;
      }
      const float t = xq - xc[j * cpad];
      const float p = t * t;
      const float dd = a.mult > 0 ? Drow[j] + p : Drow[j] - p;
      if (dd < last) {
        int m = K - 1;
        for (; m > 0; --m) {
          const float prev = Ld[(m - 1) * 64];
          if (!(dd < prev)) break;
          Ld[m * 64] = prev;
          Li[m * 64] = Li[(m - 1) * 64];
        }
        Ld[m * 64] = dd;
        Li[m * 64] = (int)j;
        last = Ld[(K - 1) * 64];
      }
    }
    for (int m = 0; m < K; ++m) lab[m * 64] = a.y[Li[m * 64]];
    for (int m = 0; m < a.nk; ++m) {
      const float pred = predict(lab, a.kvals[m], a.model_type);
      if (live) a.P[((int64_t)m * cpad + c) * n + q] = pred;
    }
  }
}

// ============================================================ launchers

int64_t long_search_waves(int64_t n, int cpad, int K) {
  // enough waves to fill the device several times over, with the lists
  // (12 bytes per slot and lane) within 512 MiB
  int64_t w = n * (cpad / 64);
  const int64_t cap = ((int64_t)512 << 20) / ((int64_t)K * 64 * 12);
  if (w > 8192) w = 8192;
  if (w > cap) w = cap;
  return w < 1 ? 1 : w;
}

hipError_t launch_search_long(const SearchProblem& a, float* ld, int* li, float* lab,
                              int64_t waves, hipStream_t s) {
  const int K = a.kmax;
  const dim3 g((unsigned)((waves + 3) / 4)), b(256);
  hipLaunchKernelGGL(k_search_long, g, b, 0, s, a, K, ld, li, lab, waves);
  return hipGetLastError();
}

template <int K, int QW, int WAVES>
static hipError_t launch_dense_k(const SearchProblem& a, hipStream_t s) {
  const int qpb = QW * WAVES;
  dim3 g((unsigned)((a.n + qpb - 1) / qpb), (unsigned)live_blocks(a));
  dim3 b(64 * WAVES);
  if (a.mult < 0)
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, -1, false>), g, b, 0, s, a);
  else if (a.zero)
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, 1, true>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, 1, false>), g, b, 0, s, a);
  return hipGetLastError();
}

int backward_seed_count(int kmax) {
  static const int forced = [] {
    const char* e = std::getenv("SPEEDYNN_NSEED");
    return e ? std::atoi(e) : 0;
  }();
  // K + 16: the bound is the K-th smallest exact distance over these seeds
  // (c4, K = 15: 32 seeds run the search 3% faster than 48, same selections)
  int s = forced > 0 ? forced : search_list_size(kmax) + 16;
  s = (s + 7) / 8 * 8;
  return s < 8 ? 8 : (s > 128 ? 128 : s);
}

template <int K, int W>
static void launch_pruned_shared(const SearchProblem& a, hipStream_t s) {
  // workgroups per query: enough W-wave groups to cover the candidate blocks
  const dim3 g((unsigned)launch_queries(a), (unsigned)((live_blocks(a) + W - 1) / W)), b(64 * W);
  if (a.mult > 0) hipLaunchKernelGGL((k_search_pruned_shared<K, 1, W>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_search_pruned_shared<K, -1, W>), g, b, 0, s, a);
}

template <int K>
static hipError_t launch_pruned_k(const SearchProblem& a, int variant, hipStream_t s) {
  const int nblk = live_blocks(a);
  // the shared-scan kernel: up to 16 candidate blocks (16 waves need lists of
  // at most 20 to stay within 128 registers); its label area needs K <= 50
  const int W = nblk <= 4 ? 4 : nblk <= 8 ? 8 : nblk <= 16 && K <= 20 ? 16 : 0;
  // variant 0: forward takes the per-wave kernel (the candidate blocks'
  // bounds differ widely when feature scales do, and a shared workgroup waits
  // for its loosest block), backward the shared scan (two rows per query,
  // similar bounds)
  const int v = variant;
  // variants 5, 6, 7: the shared scan with workgroups of W = 4, 8, 16 waves,
  // ceil(nblk / W) workgroups per query (the tile of one workgroup: one query
  // x W candidate blocks)
  if (v >= 5 && v <= 7) {
    const int Wv = 4 << (v - 5);
    if constexpr (K <= 50) {
      if (pruned_width_applies(a, Wv)) {
        if (Wv == 4) launch_pruned_shared<K, 4>(a, s);
        else if (Wv == 8) launch_pruned_shared<K, 8>(a, s);
        else if constexpr (K <= 20) launch_pruned_shared<K, 16>(a, s);
        return hipGetLastError();
      }
    }
  }
  const bool shared = K <= 50 && W > 0 && (v == 2 || (v == 0 && a.mult < 0));
  if (shared) {
    if constexpr (K <= 50) {
      if (W == 4) launch_pruned_shared<K, 4>(a, s);
      else if (W == 8) launch_pruned_shared<K, 8>(a, s);
      else if constexpr (K <= 20) launch_pruned_shared<K, 16>(a, s);
    }
    return hipGetLastError();
  }
  // variant 3: the shared scan split over two workgroups per query (half the
  // candidate blocks each; the row is read twice, but two workgroups share a
  // CU, so one's barrier waits hide behind the other's work, and each half of
  // the run gathers from half of Xc's columns)
  if (v == 3 && K <= 50 && nblk > 4 && nblk <= 16) {
    if constexpr (K <= 50) {
      if (nblk <= 8) launch_pruned_shared<K, 4>(a, s);
      else launch_pruned_shared<K, 8>(a, s);
    }
    return hipGetLastError();
  }
  // 4 independent waves per block; each XCD slice holds per_xcd*nblk waves.
  // forward: gather locality (candidate-block major); backward: the D and M
  // rows dominate (query major)
  const int64_t per_xcd = (launch_queries(a) + 7) / 8;
  const int64_t slice_blocks = (per_xcd * nblk + 3) / 4;
  const dim3 g((unsigned)(8 * slice_blocks)), b(256);
  if (a.mult > 0) hipLaunchKernelGGL((k_search_pruned<K, 1>), g, b, 0, s, a, 1);
  else hipLaunchKernelGGL((k_search_pruned<K, -1>), g, b, 0, s, a, 0);
  return hipGetLastError();
}

#define SNN_FOR_EACH_LIST_SIZE(X) \
  X(2, 1, 16) X(4, 1, 16) X(5, 1, 16) X(8, 1, 16) X(10, 1, 16) X(15, 1, 16) X(16, 1, 16) \
  X(20, 1, 16) X(24, 1, 8) X(32, 1, 8) X(50, 1, 4) X(64, 1, 4)

hipError_t launch_search_dense(const SearchProblem& a, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_dense_k<K, QW, W>(a, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

template <int K>
static hipError_t launch_sorted_k(const SearchProblem& a, int max_cand, hipStream_t s) {
  const dim3 g((unsigned)((a.n + 255) / 256), (unsigned)max_cand), b(256);
  hipLaunchKernelGGL(k_search_sorted<K>, g, b, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_search_sorted(const SearchProblem& a, int max_cand, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_sorted_k<K>(a, max_cand, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_search_pruned(const SearchProblem& a, int variant, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_pruned_k<K>(a, variant, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

bool pruned_shared_applies(const SearchProblem& a) {
  const int K = search_list_size(a.kmax), nblk = live_blocks(a);
  return K > 0 && K <= 50 && (nblk <= 8 || (nblk <= 16 && K <= 20));
}

template <int K>
static hipError_t launch_walk_k(const SearchProblem& a, const int* rank, hipStream_t s) {
  const int64_t per_xcd = (launch_queries(a) + 7) / 8;
  const int64_t slice_blocks = (per_xcd * live_blocks(a) + 3) / 4;
  hipLaunchKernelGGL(k_search_walk<K>, dim3((unsigned)(8 * slice_blocks)), dim3(256), 0, s, a,
                     rank);
  return hipGetLastError();
}

hipError_t launch_search_walk(const SearchProblem& a, const int* rank, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_walk_k<K>(a, rank, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rank(const int* sidx, const SelState* st, int* rank, int64_t n, hipStream_t s) {
  const int64_t g = (n + 255) / 256;
  hipLaunchKernelGGL(k_rank, dim3((unsigned)(g < 4096 ? g : 4096)), dim3(256), 0, s, sidx, st,
                     rank, n);
  return hipGetLastError();
}

bool pruned_width_applies(const SearchProblem& a, int W) {
  const int K = search_list_size(a.kmax);
  return K > 0 && K <= 50 && (W == 4 || W == 8 || (W == 16 && K <= 20));
}

bool pruned_split_applies(const SearchProblem& a) {
  const int K = search_list_size(a.kmax), nblk = live_blocks(a);
  return K > 0 && K <= 50 && nblk > 4 && nblk <= 16;
}

// Dense configuration (queries per wave, waves per block) of a list size.
static void dense_config(int K, int* qw, int* waves) {
  *qw = 1;
  *waves = 4;
  switch (K) {
#define SNN_CASE(KK, QW, W) \
  case KK: *qw = QW; *waves = W; break;
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: break;
  }
}

int64_t batch_dense_items(const SearchProblem& a) {
  int qw, w;
  dense_config(search_list_size(a.kmax), &qw, &w);
  return ((a.n + qw * w - 1) / (qw * w)) * (a.cpad / 64);
}

int64_t batch_pruned_items(const SearchProblem& a) { return a.n * (a.cpad / 64); }

// Resident blocks per CU x CUs: a persistent grid fills the chip once.
template <class Kernel>
static int resident_grid(Kernel kernel, int threads) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess ||
      per < 1)
    per = 1;
  return cus * per;
}

template <int K, int QW, int W>
static int batch_grid_k(int mode, int mult) {
  static const int g[4] = {
      resident_grid(k_batch_dense<K, QW, W, 1, false>, 64 * W),
      resident_grid(k_batch_dense<K, QW, W, 1, true>, 64 * W),
      resident_grid(k_batch_pruned<K, 1>, 256), resident_grid(k_batch_pruned<K, -1>, 256)};
  if (mode == 2) return g[mult > 0 ? 2 : 3];
  return g[mode == 1 ? 1 : 0];
}

int batch_grid(int mode, int mult, int K) {
  switch (K) {
#define SNN_CASE(KK, QW, W) \
  case KK: return batch_grid_k<KK, QW, W>(mode, mult);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return 0;
  }
}

template <int K, int QW, int W>
static void launch_batch_k(const SearchProblem* probs, const int64_t* start, int P,
                           unsigned long long* counter, int mode, int mult, int grid,
                           hipStream_t s) {
  const dim3 g(grid), bp(256), bd(64 * W);
  if (mode == 2) {
    if (mult > 0) hipLaunchKernelGGL((k_batch_pruned<K, 1>), g, bp, 0, s, probs, start, P, counter);
    else hipLaunchKernelGGL((k_batch_pruned<K, -1>), g, bp, 0, s, probs, start, P, counter);
  } else if (mult < 0) {
    hipLaunchKernelGGL((k_batch_dense<K, QW, W, -1, false>), g, bd, 0, s, probs, start, P,
                       counter);
  } else if (mode == 1) {
    hipLaunchKernelGGL((k_batch_dense<K, QW, W, 1, true>), g, bd, 0, s, probs, start, P, counter);
  } else {
    hipLaunchKernelGGL((k_batch_dense<K, QW, W, 1, false>), g, bd, 0, s, probs, start, P,
                       counter);
  }
}

hipError_t launch_batch_search(const SearchProblem* probs, const int64_t* start, int P,
                               unsigned long long* counter, int mode, int mult, int K, int grid,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, counter);
  switch (K) {
#define SNN_CASE(KK, QW, W) \
  case KK: launch_batch_k<KK, QW, W>(probs, start, P, counter, mode, mult, grid, s); break;
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_seed_scan(const float* D, const SelState* st, int* seeds, int64_t n,
                            int64_t npf, hipStream_t s) {
  hipLaunchKernelGGL(k_seed_scan, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, D, st, seeds, n,
                     npf);
  return hipGetLastError();
}

}  // namespace snn
