// Candidate search kernels (c) of the kNN selection engine, gfx950 wave64.
// See kernels.hip for the data layout.
#include "common.h"
#include "kernels.h"

namespace snn {

// ------------------------------------------------------ (c) candidate search
//
// Top-K insertion into a per-lane list sorted by (distance, index).  Points
// are offered in ascending index order, so a strict '<' gives the reference's
// tie rule (speedynn/lib/cpu.c:149-165: strict '>' entry, strict '<' bubble):
// an equal distance never displaces an earlier, lower index.  Branch-free
// network over compile-time slots, so the list lives in VGPRs.
template <int K>
__device__ __forceinline__ void topk_insert(float (&Ld)[K], int (&Li)[K], float dd, int j) {
#pragma unroll
  for (int m = K - 1; m > 0; --m) {
    const bool cp = dd < Ld[m - 1];
    const bool ch = dd < Ld[m];
    const float nd = cp ? Ld[m - 1] : (ch ? dd : Ld[m]);
    const int ni = cp ? Li[m - 1] : (ch ? j : Li[m]);
    Ld[m] = nd;
    Li[m] = ni;
  }
  if (dd < Ld[0]) {
    Ld[0] = dd;
    Li[0] = j;
  }
}

// Prediction for the first k neighbours whose labels sit in LDS column
// `lab` (stride 64 floats per slot).  Regression: nearest-first float sum / k
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

// Epilogue of one (query, candidate) lane: the list's labels go to the
// wave's LDS column `lab`, then the prediction of every k value is written to
// P[m][c][q].  The final list (index, or -1 for an unfilled slot) goes to NN
// when requested: the next forward step seeds its pruning bound with the list
// of the chosen candidate.
template <int K>
__device__ __forceinline__ void finish_query(const float (&Ld)[K], const int (&Li)[K], float* lab,
                                             const float* __restrict__ y,
                                             const int* __restrict__ kvals, int nk,
                                             int model_type, float* __restrict__ P,
                                             int* __restrict__ NN, int64_t q, int c, int cpad,
                                             int64_t n, bool live) {
#pragma unroll
  for (int m = 0; m < K; ++m) lab[m * 64] = y[Li[m]];
  if (NN && live) {
    int* out = NN + (q * cpad + c) * K;
#pragma unroll
    for (int m = 0; m < K; ++m) out[m] = Ld[m] < kMaxF ? Li[m] : -1;
  }
  for (int m = 0; m < nk; ++m) {
    const int k = kvals[m];
    const float pred = predict(lab, k, model_type);
    if (live) P[((int64_t)m * cpad + c) * n + q] = pred;
  }
}

// Dense search, one lane per (query, candidate): a wave owns QW queries and
// the 64 candidates of block blockIdx.y; a block owns WAVES*QW consecutive
// queries.  Per step of 64 points the block stages the candidates' columns
// Xc[j0:j0+64][cb*64:+64] in LDS (double-buffered), so each staged value
// serves WAVES*QW queries; the query's D row is wave-uniform and read with
// scalar loads.  Per point: d = fl(D_qj + mult*fl((x_qf - x_jf)^2))
// (speedynn/lib/cpu.c:142-144), points in the query's fold are skipped
// (cpu.c:139: LOO when k_fold == n).  ZERO: D is known to be 0 (first forward
// step), so d = fl((x_qf - x_jf)^2) exactly and D is not read.
template <int K, int QW, int WAVES, int MULT, bool ZERO>
__global__ __launch_bounds__(64 * WAVES) void k_search_dense(
    const float* __restrict__ D, const float* __restrict__ Xc, const float* __restrict__ y,
    const int* __restrict__ kvals, const SelState* __restrict__ st, float* __restrict__ P,
    int* __restrict__ NN, int64_t n, int64_t npf, int cpad, int nk, int model_type) {
  constexpr int J = 64;
  constexpr int LDS_SCAN = 2 * J * 64;
  constexpr int LDS_EPI = WAVES * K * 64;
  __shared__ float lds[LDS_SCAN > LDS_EPI ? LDS_SCAN : LDS_EPI];

  if (st->stopped) return;
  const int ncand = st->ncand;
  const int cb = blockIdx.y;
  if (cb * 64 >= ncand) return;
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int c = cb * 64 + lane;
  const int64_t qbase = ((int64_t)blockIdx.x * WAVES + wave) * QW;

  float xq[QW];
  float Ld[QW][K];
  int Li[QW][K];
  int64_t flo[QW], fhi[QW];
  const float* __restrict__ Drow[QW];
#pragma unroll
  for (int qi = 0; qi < QW; ++qi) {
    const int64_t q = qbase + qi;
    // queries past n (last block) scan row n-1 and never store a result
    Drow[qi] = D + (q < n ? q : n - 1) * n;
    xq[qi] = q < n ? Xc[q * cpad + c] : 0.0f;
    const int64_t fold = q / npf;
    flo[qi] = q < n ? fold * npf : n;
    fhi[qi] = q < n ? (fold + 1) * npf : n;
#pragma unroll
    for (int m = 0; m < K; ++m) {
      Ld[qi][m] = kMaxF;
      Li[qi][m] = 0;
    }
  }

  const float* __restrict__ Xblk = Xc + cb * 64;
  auto stage = [&](int buf, int64_t j0) {
    float* dst = lds + buf * (J * 64);
    for (int t = threadIdx.x; t < J * 16; t += 64 * WAVES) {
      const int r = t >> 4, c4 = (t & 15) * 4;
      const int64_t j = j0 + r;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (j < n) v = *reinterpret_cast<const float4*>(Xblk + j * cpad + c4);
      *reinterpret_cast<float4*>(dst + r * 64 + c4) = v;
    }
  };

  stage(0, 0);
  __syncthreads();
  int buf = 0;
  for (int64_t j0 = 0; j0 < n; j0 += J) {
    if (j0 + J < n) stage(buf ^ 1, j0 + J);
    const float* xs = lds + buf * (J * 64);
    const int jn = (int)((n - j0) < J ? (n - j0) : J);
    bool skip_any = false;
#pragma unroll
    for (int qi = 0; qi < QW; ++qi)
      skip_any |= (flo[qi] < j0 + jn) && (fhi[qi] > j0);
    if (!skip_any && jn == J) {
      // no fold overlap in this chunk: unchecked inner loop
      for (int jj = 0; jj < J; jj += 4) {
        float dq[QW][4];
#pragma unroll
        for (int qi = 0; qi < QW; ++qi)
#pragma unroll
          for (int u = 0; u < 4; ++u) dq[qi][u] = ZERO ? 0.0f : Drow[qi][j0 + jj + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float xj = xs[(jj + u) * 64 + lane];
          const int j = (int)(j0 + jj + u);
#pragma unroll
          for (int qi = 0; qi < QW; ++qi) {
            const float t = xq[qi] - xj;
            const float p = t * t;
            const float dd = ZERO ? p : (MULT > 0 ? dq[qi][u] + p : dq[qi][u] - p);
            if (dd < Ld[qi][K - 1]) topk_insert<K>(Ld[qi], Li[qi], dd, j);
          }
        }
      }
    } else {
      for (int jj = 0; jj < jn; ++jj) {
        const float xj = xs[jj * 64 + lane];
        const int64_t j = j0 + jj;
#pragma unroll
        for (int qi = 0; qi < QW; ++qi) {
          if (j >= flo[qi] && j < fhi[qi]) continue;
          const float dqv = ZERO ? 0.0f : Drow[qi][j];
          const float t = xq[qi] - xj;
          const float p = t * t;
          const float dd = ZERO ? p : (MULT > 0 ? dqv + p : dqv - p);
          if (dd < Ld[qi][K - 1]) topk_insert<K>(Ld[qi], Li[qi], dd, (int)j);
        }
      }
    }
    __syncthreads();
    buf ^= 1;
  }

  float* lab = lds + wave * (K * 64) + lane;
  const bool live = c < ncand;
#pragma unroll
  for (int qi = 0; qi < QW; ++qi) {
    const int64_t q = qbase + qi;
    if (q >= n) break;
    finish_query<K>(Ld[qi], Li[qi], lab, y, kvals, nk, model_type, P, NN, q, c, cpad, n, live);
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

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
  const float kInf = __int_as_float(0x7f800000);
  float b0 = kInf, b1 = kInf;
  int i0 = -1, i1 = -1;
  for (int64_t j = lane; j < n; j += 64) {
    if (j >= flo && j < fhi) continue;
    const float v = Dq[j];
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
  seeds[q * 128 + lane] = i0;
  seeds[q * 128 + 64 + lane] = i1;
}

// Values-only sorted insertion (the bound needs the K-th value, not indices).
template <int K>
__device__ __forceinline__ void kth_insert(float (&B)[K], float v) {
#pragma unroll
  for (int m = K - 1; m > 0; --m) B[m] = v < B[m - 1] ? B[m - 1] : (v < B[m] ? v : B[m]);
  B[0] = v < B[0] ? v : B[0];
}

// Pruned forward search with an exact bound (forward steps after the first).
//
// Forward distances only grow: d_qjf = fl(D_qj + fl(t^2)) >= D_qj.  For a
// candidate f, the K-th smallest d_qjf over any K distinct eligible points is
// an upper bound tau_f of its final K-th distance; here it is the K-th
// smallest over the 128 seeds of k_seed_scan, evaluated exactly.  A point
// with D_qj > T = max over the wave's candidates of tau_f can enter no lane's
// final top-K and is skipped before any per-candidate work; T tightens to the
// lists' own K-th distance as they fill.  Points are still offered in
// ascending index order, so the lists equal the dense search's bit for bit.
//
// One block = one query; wave w = candidate slots [64w, 64w+64), so the D row
// comes from HBM once per query and the other waves hit in cache.  Per 64
// points: one coalesced load of D_q[j0:j0+64], a ballot of the points with
// D_qj <= T outside the query's fold, then for each such point (ascending)
// the candidates' values Xc[j][64w+lane] (256 B, coalesced) and the distance.
// A fused multiply-add screen fma(t,t,D) is compared against the K-th
// distance widened by 2^-20 (plus 1e-36 for the subnormal range);
// |fl(D+fl(t*t)) - fma(t,t,D)| <= 3.01*2^-24*(D+t*t) for D, t*t >= 0, so a
// point the screen rejects cannot pass the exact test.  Screen passes
// recompute d in the reference's order before inserting.
template <int K>
__global__ __launch_bounds__(1024) void k_search_pruned(
    const float* __restrict__ D, const float* __restrict__ Xc, const float* __restrict__ y,
    const int* __restrict__ kvals, const SelState* __restrict__ st,
    const int* __restrict__ seeds, float* __restrict__ P, unsigned long long* __restrict__ visited,
    int64_t n, int64_t npf, int cpad, int nk, int model_type) {
  extern __shared__ float lds[];  // epilogue labels: waves x K x 64
  if (st->stopped) return;
  const int ncand = st->ncand;
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int cb = blockIdx.y * (blockDim.x >> 6) + wave;
  if (cb * 64 >= ncand) return;  // whole wave; this kernel has no block barriers
  const int c = cb * 64 + lane;
  const bool live = c < ncand;
  const int64_t q = blockIdx.x;

  const float* __restrict__ Dq = D + q * n;
  const float* __restrict__ Xl = Xc + cb * 64 + lane;
  const float xq = Xl[q * cpad];
  const int64_t flo = (q / npf) * npf;
  const int64_t fhi = flo + npf < n ? flo + npf : n;
  const float kInf = __int_as_float(0x7f800000);
  constexpr float kScreen = 1.00000095367431640625f;  // 1 + 2^-20
  constexpr float kTiny = 1.0e-36f;

  // bound: K-th smallest exact distance over the seeds
  float B[K];
#pragma unroll
  for (int m = 0; m < K; ++m) B[m] = kInf;
  const int* __restrict__ sq = seeds + q * 128;
  for (int s0 = 0; s0 < 128; s0 += 8) {
    int js[8];
    float xs[8], dqs[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      js[u] = sq[s0 + u];
      const int jj = js[u] < 0 ? 0 : js[u];
      xs[u] = Xl[(int64_t)jj * cpad];
      dqs[u] = Dq[jj];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (js[u] >= 0) {
        const float t = xq - xs[u];
        const float p = t * t;
        const float dd = dqs[u] + p;
        if (dd < B[K - 1]) kth_insert<K>(B, dd);
      }
    }
  }
  float thr = live ? B[K - 1] : -kInf;
  float T = wave_max(thr);

  float Ld[K];
  int Li[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    Ld[m] = kMaxF;
    Li[m] = 0;
  }
  // screen limit: live lanes pass every point until their list is full
  // (+inf); padding lanes (c >= ncand) never pass (-inf)
  float lim = live ? kInf : -kInf;
  unsigned long long seen = 0;
  for (int64_t j0 = 0; j0 < n; j0 += 64) {
    const int64_t jl = j0 + lane;
    const float Dv = jl < n ? Dq[jl] : kInf;
    const bool ok = jl < n && (jl < flo || jl >= fhi) && Dv <= T;
    unsigned long long mask = __ballot(ok);
    if (!mask) continue;
    seen += __popcll(mask);
    bool changed = false;
    while (mask) {
      constexpr int BATCH = 8;
      int js[BATCH];
      float dqs[BATCH], xs[BATCH];
      int cnt = 0;
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        if (mask) {
          const int b = __ffsll((long long)mask) - 1;
          mask &= mask - 1;
          js[u] = (int)(j0 + b);
          dqs[u] = lane_value(Dv, b);
          xs[u] = Xl[(int64_t)js[u] * cpad];
          cnt = u + 1;
        }
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        if (u < cnt) {
          const float t = xq - xs[u];
          const float screen = fmaf(t, t, dqs[u]);
          if (screen < lim) {
            const float p = t * t;
            const float dd = dqs[u] + p;
            if (dd < Ld[K - 1]) {
              topk_insert<K>(Ld, Li, dd, js[u]);
              lim = Ld[K - 1] < kMaxF ? fmaf(Ld[K - 1], kScreen, kTiny) : kInf;
              changed = true;
            }
          }
        }
      }
    }
    if (__ballot(changed)) {
      thr = live ? fminf(thr, Ld[K - 1]) : -kInf;
      T = wave_max(thr);
    }
  }
  if (visited && lane == 0) {
    const int nlive = ncand - cb * 64 < 64 ? ncand - cb * 64 : 64;
    atomicAdd(visited, seen * (unsigned long long)nlive);  // point evaluations performed
  }
  float* lab = lds + wave * (K * 64) + lane;
  finish_query<K>(Ld, Li, lab, y, kvals, nk, model_type, P, nullptr, q, c, cpad, n, live);
}

// ============================================================ launchers

template <int K, int QW, int WAVES>
static hipError_t launch_dense_k(const SearchLaunch& a, hipStream_t s) {
  const int qpb = QW * WAVES;
  dim3 g((unsigned)((a.n + qpb - 1) / qpb), (unsigned)(a.cpad / 64));
  dim3 b(64 * WAVES);
  if (a.mult < 0)
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, -1, false>), g, b, 0, s, a.D, a.Xc, a.y,
                       a.kvals, a.st, a.P, a.NN, a.n, a.npf, a.cpad, a.nk, a.model_type);
  else if (a.zero)
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, 1, true>), g, b, 0, s, a.D, a.Xc, a.y,
                       a.kvals, a.st, a.P, a.NN, a.n, a.npf, a.cpad, a.nk, a.model_type);
  else
    hipLaunchKernelGGL((k_search_dense<K, QW, WAVES, 1, false>), g, b, 0, s, a.D, a.Xc, a.y,
                       a.kvals, a.st, a.P, a.NN, a.n, a.npf, a.cpad, a.nk, a.model_type);
  return hipGetLastError();
}

template <int K>
static hipError_t launch_pruned_k(const SearchLaunch& a, hipStream_t s) {
  // waves of a block share the query; epilogue LDS (waves*K*256 B) stays <= 64 KB
  const int nblk = a.cpad / 64;
  int waves = nblk < 16 ? nblk : 16;
  if (waves > 256 / K) waves = 256 / K;
  const unsigned gy = (unsigned)((nblk + waves - 1) / waves);
  const size_t lds = sizeof(float) * (size_t)waves * K * 64;
  hipLaunchKernelGGL((k_search_pruned<K>), dim3((unsigned)a.n, gy), dim3(64 * waves), lds, s, a.D,
                     a.Xc, a.y, a.kvals, a.st, a.seeds, a.P, a.visited, a.n, a.npf, a.cpad, a.nk,
                     a.model_type);
  return hipGetLastError();
}

int search_list_size(int kmax) {
  static const int sizes[] = {2, 4, 5, 8, 10, 15, 16, 20, 24, 32, 50, 64};
  for (int v : sizes)
    if (kmax <= v) return v;
  return -1;
}

#define SNN_FOR_EACH_LIST_SIZE(X) \
  X(2, 2, 8) X(4, 2, 8) X(5, 2, 8) X(8, 2, 8) X(10, 2, 8) X(15, 2, 8) X(16, 2, 8) X(20, 2, 8) \
  X(24, 2, 8) X(32, 1, 8) X(50, 1, 4) X(64, 1, 4)

hipError_t launch_search_dense(const SearchLaunch& a, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_dense_k<K, QW, W>(a, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_search_pruned(const SearchLaunch& a, hipStream_t s) {
  switch (search_list_size(a.kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_pruned_k<K>(a, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_seed_scan(const float* D, const SelState* st, int* seeds, int64_t n,
                            int64_t npf, hipStream_t s) {
  hipLaunchKernelGGL(k_seed_scan, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, D, st, seeds, n,
                     npf);
  return hipGetLastError();
}

}  // namespace snn
