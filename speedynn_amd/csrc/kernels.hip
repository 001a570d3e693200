// Device kernels of the kNN wrapper feature-selection engine (gfx950, wave64).
//
// Data layout in HBM (one selection problem):
//   X   [n][d]     float32, row-major (the caller's layout)
//   XT  [d][n]     float32, feature-major copy: one feature column is contiguous
//   D   [n][n]     float32, current squared distances, row-major and symmetric
//   Dlo [n][n]     float32, compensation term of D (precision form only)
//   Xc  [n][cpad]  float32, the current candidates' columns gathered per point
//   P   [nk][cpad][n] float32, LOO prediction per (k value, candidate, query)
//   err [nl][nk][cpad] float32, loss per (loss, k value, candidate)
//
// Arithmetic of the plain paths is IEEE binary32 in the reference's operation
// order; the build uses -ffp-contract=off, so a*b+c is never fused unless a
// kernel calls fmaf() on purpose.
#include "common.h"
#include "kernels.h"

namespace snn {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------- utility

__global__ void k_fill(float* __restrict__ p, int64_t count, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n4 = count >> 2;
  float4* p4 = reinterpret_cast<float4*>(p);
  for (int64_t t = i; t < n4; t += stride) p4[t] = make_float4(v, v, v, v);
  for (int64_t t = (n4 << 2) + i; t < count; t += stride) p[t] = v;
}

// XT[f][j] = X[j][f], 32x32 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ X,
                                                   float* __restrict__ XT, int64_t n, int d) {
  __shared__ float tile[32][33];
  const int64_t j0 = (int64_t)blockIdx.x * 32;
  const int f0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t j = j0 + r;
    const int f = f0 + tx;
    tile[r][tx] = (j < n && f < d) ? X[j * d + f] : 0.0f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int f = f0 + r;
    const int64_t j = j0 + tx;
    if (f < d && j < n) XT[(int64_t)f * n + j] = tile[tx][r];
  }
}

// Error of one compensated step s = fl(acc + ps), ps = +-fl(t*t), t = fl(a - b):
// the exact value acc + (a-b)^2 (or minus it) equals s plus the returned
// term up to O(u^2): TwoDiff error e of the difference (contributing
// 2*t*e + e*e to the square), TwoProduct error of t*t (exact via fma) and
// TwoSum error of the addition.  neg: ps = -p (backward update).
__device__ __forceinline__ float comp_error(float a, float b, float t, float p, float acc,
                                            float ps, float s, bool neg) {
  const float bb = t - a;
  const float e = (a - (t - bb)) + (-b - bb);          // a - b = t + e exactly
  const float pe = fmaf(t, t, -p) + fmaf(2.0f * t, e, e * e);  // (a-b)^2 - p
  const float bp = s - acc;
  const float se = (acc - (s - bp)) + (ps - bp);       // acc + ps = s + se exactly
  return se + (neg ? -pe : pe);
}

// ------------------------------------------------- (a) full distance matrix
//
// D[i][j] = sum over k = d-1 ... 0 of fl((x_ik - x_jk)^2), accumulated left to
// right in float32 (speedynn/lib/util.c:10-17, cpu.c:185-215).  The diagonal
// comes out as exactly 0 and the matrix exactly symmetric.  With COMP the
// rounding error of every square and every addition is accumulated in Dlo
// (TwoProduct/TwoSum error terms), so D + Dlo carries ~2x the precision.
// 64x64 output tile per 256-thread block, 4x4 per thread, X^T chunks of 16
// features staged in LDS.
template <bool COMP>
__global__ __launch_bounds__(256) void k_gram_exact(const float* __restrict__ XT,
                                                    float* __restrict__ D, float* __restrict__ Dlo,
                                                    int64_t n, int d) {
  constexpr int KC = 16;
  __shared__ float As[KC][64];
  __shared__ float Bs[KC][64];
  const int64_t i0 = (int64_t)blockIdx.y * 64;
  const int64_t j0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[4][4], lo[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) { acc[a][b] = 0.0f; lo[a][b] = 0.0f; }

  for (int kc = d; kc > 0; kc -= KC) {
    const int kb = kc - KC > 0 ? kc - KC : 0;
    const int cnt = kc - kb;
    for (int t = threadIdx.x; t < KC * 64; t += 256) {
      const int kk = t >> 6, r = t & 63;
      float av = 0.0f, bv = 0.0f;
      if (kk < cnt) {
        const float* row = XT + (int64_t)(kb + kk) * n;
        if (i0 + r < n) av = row[i0 + r];
        if (j0 + r < n) bv = row[j0 + r];
      }
      As[kk][r] = av;
      Bs[kk][r] = bv;
    }
    __syncthreads();
    for (int kk = cnt - 1; kk >= 0; --kk) {
      float a4[4], b4[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) a4[a] = As[kk][ty * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) b4[b] = Bs[kk][tx * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float t = a4[a] - b4[b];
          const float p = t * t;
          const float s = acc[a][b] + p;
          if (COMP) lo[a][b] += comp_error(a4[a], b4[b], t, p, acc[a][b], p, s, false);
          acc[a][b] = s;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + ty * 4 + a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + tx * 4 + b;
      if (j < n) {
        D[i * n + j] = acc[a][b];
        if (COMP) Dlo[i * n + j] = lo[a][b];
      }
    }
  }
}

// --------------------------------------------------------- (b) update
//
// D[i][j] = fl(D[i][j] + mult * fl((x_if - x_jf)^2)) over the full matrix
// (speedynn/lib/cpu.c:217-231 does the upper triangle and mirrors it; the
// square is symmetric under the sign of the difference, and the diagonal
// receives + 0, so both give the same bits).  The feature is the argument f
// when f >= 0, else the device state's pending choice.  Streaming: float4
// along each row, x_if read once per row.  COMP keeps the TwoSum/TwoProduct
// error in Dlo; D itself stays bit-identical to the plain update.
template <bool COMP, bool VEC>
__global__ __launch_bounds__(256) void k_update(float* __restrict__ D, float* __restrict__ Dlo,
                                                const float* __restrict__ XT, int64_t n,
                                                const SelState* __restrict__ st, int f, int mult) {
  if (f < 0) {
    if (!st->pending) return;
    f = st->fstar;
  }
  const float* __restrict__ xf = XT + (int64_t)f * n;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const float xi = xf[i];
    float* __restrict__ row = D + i * n;
    float* __restrict__ rlo = COMP ? Dlo + i * n : nullptr;
    if (VEC) {
      for (int64_t j = (int64_t)threadIdx.x * 4; j < n; j += 256 * 4) {
        const float4 xj = *reinterpret_cast<const float4*>(xf + j);
        float4 dv = *reinterpret_cast<float4*>(row + j);
        float4 lv;
        if (COMP) lv = *reinterpret_cast<float4*>(rlo + j);
        float* dp = &dv.x;
        float* lp = &lv.x;
        const float* xp = &xj.x;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float t = xi - xp[e];
          const float p = t * t;
          const float ps = mult > 0 ? p : -p;
          const float s = dp[e] + ps;
          if (COMP) lp[e] += comp_error(xi, xp[e], t, p, dp[e], ps, s, mult < 0);
          dp[e] = s;
        }
        *reinterpret_cast<float4*>(row + j) = dv;
        if (COMP) *reinterpret_cast<float4*>(rlo + j) = lv;
      }
    } else {
      for (int64_t j = threadIdx.x; j < n; j += 256) {
        const float xj = xf[j];
        const float t = xi - xj;
        const float p = t * t;
        const float ps = mult > 0 ? p : -p;
        const float dv = row[j];
        const float s = dv + ps;
        if (COMP) rlo[j] += comp_error(xi, xj, t, p, dv, ps, s, mult < 0);
        row[j] = s;
      }
    }
  }
}

// ------------------------------------------------------- candidate gather

// Xc[j][c] = X[j][cand[c]] for c < ncand, 0 for the padding slots.
__global__ __launch_bounds__(256) void k_compact(const float* __restrict__ X,
                                                 const int* __restrict__ cand,
                                                 SelState* __restrict__ st,
                                                 float* __restrict__ Xc, int64_t n, int d,
                                                 int cpad) {
  // the previous step's update has completed (stream order): retire it
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pending = 0;
  if (st->stopped) return;
  const int nc = st->ncand;
  const int64_t total = n * cpad;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / cpad;
    const int c = (int)(t - j * cpad);
    Xc[t] = c < nc ? X[j * d + cand[c]] : 0.0f;
  }
}

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

// Pruned forward search with an exact bound (forward steps after the first).
//
// Forward distances only grow: d_qjf = fl(D_qj + fl(t^2)) >= D_qj.  The
// previous step's final list for (q, chosen feature) gives K distinct
// eligible points ("seeds"); the K-th smallest d_qjf over all points is at
// most the largest d_qjf over the seeds, tau_f.  A point with
// D_qj > T = max over the wave's candidates of tau_f therefore cannot enter any
// lane's final list and is skipped before any per-candidate work; T tightens
// to the lists' own K-th distance as they fill.  Points are still offered in
// ascending index order, so the lists are identical to the dense search's.
//
// One wave = one query x 64 candidates.  Per 64 points: one coalesced load of
// D_q[j0:j0+64], a ballot of the points with D_qj <= T outside the query's
// fold, then for each such point (ascending) the candidates' values
// Xc[j][cb*64+lane] (256 B, coalesced) and the distance.  A fused
// multiply-add screen fma(t,t,D) is compared against the K-th distance
// widened by 2^-20; |fl(D+fl(t*t)) - fma(t,t,D)| <= 3.01*2^-24*(D+t*t) for
// D, t*t >= 0, so a point the screen rejects cannot pass the exact test.
// Screen passes recompute d in the reference's order before inserting.
template <int K, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_search_pruned(
    const float* __restrict__ D, const float* __restrict__ Xc, const float* __restrict__ y,
    const int* __restrict__ kvals, const SelState* __restrict__ st,
    const int* __restrict__ seeds, float* __restrict__ P, int* __restrict__ NN,
    unsigned long long* __restrict__ visited, int64_t n, int64_t npf, int cpad, int nk,
    int model_type) {
  __shared__ float lds[WAVES * K * 64];
  if (st->stopped) return;
  const int ncand = st->ncand;
  const int cb = blockIdx.y;
  if (cb * 64 >= ncand) return;
  const int lane = threadIdx.x & 63;
  const int wave = uniform(threadIdx.x >> 6);
  const int c = cb * 64 + lane;
  const bool live = c < ncand;
  const int64_t q = (int64_t)blockIdx.x * WAVES + wave;
  if (q >= n) return;  // whole wave; this kernel has no block barriers

  const float* __restrict__ Dq = D + q * n;
  const float* __restrict__ Xl = Xc + cb * 64 + lane;
  const float xq = Xl[q * cpad];
  const int64_t fold = q / npf;
  const int64_t flo = fold * npf;
  const int64_t fhi = flo + npf < n ? flo + npf : n;
  const float kInf = __int_as_float(0x7f800000);
  constexpr float kScreen = 1.00000095367431640625f;  // 1 + 2^-20
  constexpr float kTiny = 1.0e-36f;  // absolute slack: the relative bound fails near subnormals

  // seed bound
  float tau = -kInf;
  bool bounded = true;
  for (int m = 0; m < K; ++m) {
    const int j = seeds[q * K + m];
    if (j < 0) {
      bounded = false;
      break;
    }
    const float t = xq - Xl[(int64_t)j * cpad];
    const float p = t * t;
    const float dd = Dq[j] + p;
    tau = fmaxf(tau, dd);
  }
  float thr = live ? (bounded ? tau : kInf) : -kInf;
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
      constexpr int B = 8;
      int js[B];
      float dqs[B], xs[B];
      int cnt = 0;
#pragma unroll
      for (int u = 0; u < B; ++u) {
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
      for (int u = 0; u < B; ++u) {
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
  finish_query<K>(Ld, Li, lab, y, kvals, nk, model_type, P, NN, q, c, cpad, n, live);
}

// seeds[q][m] = NN[q][fslot][m]: the chosen candidate's final lists.
template <int K>
__global__ __launch_bounds__(256) void k_seeds(const int* __restrict__ NN,
                                              const SelState* __restrict__ st,
                                              int* __restrict__ seeds, int64_t n, int cpad) {
  if (!st->pending) return;
  const int slot = st->fslot;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * K;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = t / K;
    const int m = (int)(t - q * K);
    seeds[t] = NN[(q * cpad + slot) * K + m];
  }
}

// ------------------------------------------------------------------ losses

// Sequential float sum over q = 0..n-1 (or n-1..0) of per-point terms, in
// exactly that order, one wave: lanes compute 64 terms, then the terms are
// added one by one (readlane) so the rounding matches a scalar loop.
template <bool REV, class Term>
__device__ float seq_sum(int64_t n, Term term) {
  const int lane = threadIdx.x & 63;
  float s = 0.0f;
  if (!REV) {
    for (int64_t q0 = 0; q0 < n; q0 += 64) {
      const int64_t q = q0 + lane;
      const float t = q < n ? term(q) : 0.0f;
      const int cnt = (int)((n - q0) < 64 ? (n - q0) : 64);
      for (int i = 0; i < cnt; ++i) s = s + lane_value(t, i);
    }
  } else {
    for (int64_t q1 = n; q1 > 0; q1 -= 64) {
      const int64_t base = q1 - 64 > 0 ? q1 - 64 : 0;
      const int64_t q = base + lane;
      const float t = q < q1 ? term(q) : 0.0f;
      for (int i = (int)(q1 - base) - 1; i >= 0; --i) s = s + lane_value(t, i);
    }
  }
  return s;
}

// loss_functions.c:113-127.  abs() there is the int abs, so an error is
// (int)(y - p) != 0; x86 converts NaN and out-of-range values to INT_MIN,
// whose abs() is negative and so counts as correct: mirrored here.
__device__ __forceinline__ int zero_one_err(float yv, float pv) {
  const float v = yv - pv;
  if (!(fabsf(v) < 2147483648.0f)) return 0;
  return ((int)v != 0) ? 1 : 0;
}

__device__ float loss_value(int code, const float* __restrict__ y, const float* __restrict__ p,
                            int64_t n) {
  const int lane = threadIdx.x & 63;
  switch (code) {
    case kZeroOne: {
      int cnt = 0;
      for (int64_t q = lane; q < n; q += 64) cnt += zero_one_err(y[q], p[q]);
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
      return (float)cnt / (float)n;
    }
    case kRmse: {
      const float s = seq_sum<false>(n, [&](int64_t q) {
        const float t = y[q] - p[q];
        return t * t;
      });
      return (float)sqrt((double)(s / (float)n));
    }
    case kRmseNorm: {
      const float s = seq_sum<true>(n, [&](int64_t q) {
        const float tv = y[q];
        const float t = (float)((double)(tv - p[q]) / (1.0 + (double)tv));
        return t * t;
      });
      return (float)sqrt((double)(s / (float)n));
    }
    case kStd: {
      float mean = seq_sum<false>(n, [&](int64_t q) { return y[q] - p[q]; });
      mean = mean / (float)n;
      const float t = (y[n - 1] - p[n - 1]) - mean;
      const float s = (t * t) / (float)(n - 1);
      return (float)sqrt((double)s);
    }
    case kRmseStd: {
      const float a = loss_value(kRmse, y, p, n);
      const float b = loss_value(kStd, y, p, n);
      return a + b;
    }
    default:
      return __int_as_float(0x7fc00000);
  }
}

// err[l][m][c] for c < ncand: one wave per (candidate, k value).
__global__ __launch_bounds__(64) void k_loss(const float* __restrict__ P,
                                             const float* __restrict__ y,
                                             const int* __restrict__ losses,
                                             const SelState* __restrict__ st,
                                             float* __restrict__ err, int64_t n, int cpad, int nk,
                                             int nl) {
  if (st->stopped) return;
  const int c = blockIdx.x, m = blockIdx.y;
  if (c >= st->ncand) return;
  const float* p = P + ((int64_t)m * cpad + c) * n;
  for (int l = 0; l < nl; ++l) {
    const float e = loss_value(losses[l], y, p, n);
    if (threadIdx.x == 0) err[((int64_t)l * nk + m) * cpad + c] = e;
  }
}

// Loss of one prediction vector (get_loss_all).
__global__ __launch_bounds__(64) void k_loss_single(const float* __restrict__ y,
                                                    const float* __restrict__ p, int64_t n,
                                                    int code, float* __restrict__ out) {
  const float e = loss_value(code, y, p, n);
  if (threadIdx.x == 0) *out = e;
}

// ------------------------------------------------------------- arg-min step
//
// One thread: best k per candidate (first k value, in k_vals order, with the
// strictly smallest first-loss error; speedynn/lib/cpu.c:110-120), arg-min
// over candidates in ascending feature order with strict '<'
// (nearest_neighbors_feature_selection.c:318-329 forward, :155-167 backward),
// records (:336-342 / :174-181), auto-stop (:446-465) and the next step's
// candidate list.
__global__ void k_select(SelState* __restrict__ st, const Problem pb,
                         const int* __restrict__ kvals, const float* __restrict__ err,
                         int* __restrict__ cand, int* __restrict__ sel, int* __restrict__ ordering,
                         float* __restrict__ errors, float* __restrict__ val,
                         int* __restrict__ bestk, float* __restrict__ cand_err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->stopped) return;
  const int nc = st->ncand;
  const int cpad = pb.cpad, nk = pb.nk, nl = pb.nl, d = pb.d;
  const float qnan = __int_as_float(0x7fc00000);
  for (int c = 0; c < nc; ++c) {
    float smallest = kMaxF;
    int bm = -1;
    for (int m = 0; m < nk; ++m) {
      const float e = err[(int64_t)m * cpad + c];
      if (e < smallest) {
        smallest = e;
        bm = m;
      }
    }
    for (int l = 0; l < nl; ++l)
      val[l * cpad + c] = bm >= 0 ? err[((int64_t)l * nk + bm) * cpad + c] : qnan;
    bestk[c] = bm >= 0 ? kvals[bm] : 0;
    if (cand_err) cand_err[(int64_t)st->step * d + cand[c]] = val[c];
  }
  float minv = kMaxF;
  int best = -1;
  for (int c = 0; c < nc; ++c)
    if (val[c] < minv) {
      minv = val[c];
      best = c;
    }
  if (best < 0) {
    st->error = 1;
    st->stopped = 2;
    return;
  }
  const int f = cand[best];
  const int pos = st->pos;
  const bool fwd = pb.method == 1;
  sel[f] = fwd ? 1 : 0;
  ordering[pos] = f;
  for (int l = 0; l < nl; ++l) errors[l * d + pos] = val[l * cpad + best];
  st->best_k = bestk[best];
  st->fstar = f;
  st->fslot = best;
  st->pending = 1;
  st->num_final += fwd ? 1 : -1;
  st->step += 1;
  bool stop = false;
  if (pb.auto_feature) {
    if (fwd) stop = pos > 0 && errors[pos - 1] < errors[pos];
    else stop = pos > 0 && pos < d - 1 && errors[pos + 1] < errors[pos];
  }
  const int next = fwd ? pos + 1 : pos - 1;
  if (fwd ? !(next < pb.num_desired) : !(next >= pb.num_desired)) stop = true;
  st->pos = next;
  const int check = fwd ? 0 : 1;
  int nc2 = 0;
  for (int g = 0; g < d; ++g)
    if (sel[g] == check) cand[nc2++] = g;
  st->ncand = nc2;
  if (nc2 == 0) stop = true;
  st->stopped = stop ? 1 : 0;
}

// Initial state: all features unselected (forward) or selected (backward).
__global__ void k_init_state(SelState* __restrict__ st, const Problem pb, int* __restrict__ cand,
                             int* __restrict__ sel, int* __restrict__ ordering,
                             float* __restrict__ errors) {
  const int t = threadIdx.x;
  const bool fwd = pb.method == 1;
  for (int g = t; g < pb.d; g += blockDim.x) {
    sel[g] = fwd ? 0 : 1;
    ordering[g] = 0;
    cand[g] = g;
  }
  for (int g = t; g < pb.nl * pb.d; g += blockDim.x) errors[g] = 0.0f;
  if (t == 0) {
    st->step = 0;
    st->fstar = -1;
    st->pending = 0;
    st->best_k = 0;
    st->error = 0;
    st->ncand = pb.d;
    st->num_final = fwd ? 0 : pb.d;
    st->pos = fwd ? 0 : pb.d - 1;
    const bool none = fwd ? !(0 < pb.num_desired) : !(pb.d - 1 >= pb.num_desired);
    st->stopped = none ? 1 : 0;
  }
}

// ============================================================ launchers

static inline int grid_1d(int64_t work, int block, int cap) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  return (int)(g > cap ? cap : g);
}

hipError_t launch_fill(float* p, int64_t count, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_1d(count / 4, 256, 16384)), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_transpose(const float* X, float* XT, int64_t n, int d, hipStream_t s) {
  dim3 g((unsigned)((n + 31) / 32), (unsigned)((d + 31) / 32));
  hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, s, X, XT, n, d);
  return hipGetLastError();
}

hipError_t launch_gram_exact(const float* XT, float* D, float* Dlo, int64_t n, int d,
                             hipStream_t s) {
  dim3 g((unsigned)((n + 63) / 64), (unsigned)((n + 63) / 64));
  if (Dlo) hipLaunchKernelGGL(k_gram_exact<true>, g, dim3(256), 0, s, XT, D, Dlo, n, d);
  else hipLaunchKernelGGL(k_gram_exact<false>, g, dim3(256), 0, s, XT, D, Dlo, n, d);
  return hipGetLastError();
}

hipError_t launch_update(float* D, float* Dlo, const float* XT, int64_t n, const SelState* st,
                         int f, int mult, hipStream_t s) {
  const int g = (int)(n < 8192 ? n : 8192);
  const bool vec = (n % 4) == 0;
  if (Dlo) {
    if (vec) hipLaunchKernelGGL((k_update<true, true>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st, f, mult);
    else hipLaunchKernelGGL((k_update<true, false>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st, f, mult);
  } else {
    if (vec) hipLaunchKernelGGL((k_update<false, true>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st, f, mult);
    else hipLaunchKernelGGL((k_update<false, false>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st, f, mult);
  }
  return hipGetLastError();
}

hipError_t launch_compact(const float* X, const int* cand, SelState* st, float* Xc,
                          int64_t n, int d, int cpad, hipStream_t s) {
  hipLaunchKernelGGL(k_compact, dim3(grid_1d(n * cpad, 256, 16384)), dim3(256), 0, s, X, cand, st,
                     Xc, n, d, cpad);
  return hipGetLastError();
}

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
  constexpr int WAVES = 4;
  dim3 g((unsigned)((a.n + WAVES - 1) / WAVES), (unsigned)(a.cpad / 64));
  hipLaunchKernelGGL((k_search_pruned<K, WAVES>), g, dim3(64 * WAVES), 0, s, a.D, a.Xc, a.y,
                     a.kvals, a.st, a.seeds, a.P, a.NN, a.visited, a.n, a.npf, a.cpad, a.nk,
                     a.model_type);
  return hipGetLastError();
}

template <int K>
static hipError_t launch_seeds_k(const int* NN, const SelState* st, int* seeds, int64_t n,
                                 int cpad, hipStream_t s) {
  hipLaunchKernelGGL((k_seeds<K>), dim3(grid_1d(n * K, 256, 4096)), dim3(256), 0, s, NN, st,
                     seeds, n, cpad);
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

hipError_t launch_seeds(const int* NN, const SelState* st, int* seeds, int64_t n, int cpad,
                        int kmax, hipStream_t s) {
  switch (search_list_size(kmax)) {
#define SNN_CASE(K, QW, W) \
  case K: return launch_seeds_k<K>(NN, st, seeds, n, cpad, s);
    SNN_FOR_EACH_LIST_SIZE(SNN_CASE)
#undef SNN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_loss(const float* P, const float* y, const int* losses, const SelState* st,
                       float* err, int64_t n, int cpad, int nk, int nl, int max_cand,
                       hipStream_t s) {
  dim3 g((unsigned)max_cand, (unsigned)nk);
  hipLaunchKernelGGL(k_loss, g, dim3(64), 0, s, P, y, losses, st, err, n, cpad, nk, nl);
  return hipGetLastError();
}

hipError_t launch_loss_single(const float* y, const float* p, int64_t n, int code, float* out,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_loss_single, dim3(1), dim3(64), 0, s, y, p, n, code, out);
  return hipGetLastError();
}

hipError_t launch_select(SelState* st, const Problem& pb, const int* kvals, const float* err,
                         int* cand, int* sel, int* ordering, float* errors, float* val,
                         int* bestk, float* cand_err, hipStream_t s) {
  hipLaunchKernelGGL(k_select, dim3(1), dim3(64), 0, s, st, pb, kvals, err, cand, sel, ordering,
                     errors, val, bestk, cand_err);
  return hipGetLastError();
}

hipError_t launch_init_state(SelState* st, const Problem& pb, int* cand, int* sel, int* ordering,
                             float* errors, hipStream_t s) {
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(256), 0, s, st, pb, cand, sel, ordering, errors);
  return hipGetLastError();
}

}  // namespace snn
