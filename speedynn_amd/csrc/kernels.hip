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
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "qsort_order.h"

namespace snn {

// ---------------------------------------------------------------- utility

__global__ void k_fill(float* __restrict__ p, int64_t count, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n4 = count >> 2;
  float4* p4 = reinterpret_cast<float4*>(p);
  for (int64_t t = i; t < n4; t += stride) p4[t] = make_float4(v, v, v, v);
  for (int64_t t = (n4 << 2) + i; t < count; t += stride) p[t] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int64_t count, int v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += stride) p[t] = 0;
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
// With MAXM, M[i][j] = max over features of fl((x_ik - x_jk)^2) is written
// too: the backward pruned search bounds d = fl(D - q_f) >= fl(D - M).
// Tile (bi, bj), bi <= bj, of the upper triangle of 64 x 64 tiles, from a
// linear block index L (column-major: L = bj (bj + 1) / 2 + bi).
__device__ __forceinline__ void tri_tile(int64_t L, int64_t& bi, int64_t& bj) {
  int64_t c = (int64_t)((sqrt(8.0 * (double)L + 1.0) - 1.0) / 2.0);
  while (c * (c + 1) / 2 > L) --c;
  while ((c + 1) * (c + 2) / 2 <= L) ++c;
  bj = c;
  bi = L - c * (c + 1) / 2;
}

// A 64 x 64 tile's values v[a][b] at (i0 + 4ty + a, j0 + 4tx + b) stored at
// their place and, for an off-diagonal tile of a symmetric matrix, mirrored
// at (j, i): each thread's mirrored values of one column j are four
// consecutive floats of row j.
__device__ __forceinline__ void store_tile(float* __restrict__ Out, int64_t n, int64_t i0,
                                           int64_t j0, int tx, int ty, const float (&v)[4][4],
                                           bool mirror) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + ty * 4 + a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + tx * 4 + b;
      if (j < n) Out[i * n + j] = v[a][b];
    }
  }
  if (!mirror) return;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int64_t j = j0 + tx * 4 + b;
    if (j >= n) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = i0 + ty * 4 + a;
      if (i < n) Out[j * n + i] = v[a][b];
    }
  }
}

template <bool COMP, bool MAXM, bool TRI>
__global__ __launch_bounds__(256) void k_gram_exact(const float* __restrict__ XT,
                                                    float* __restrict__ D, float* __restrict__ Dlo,
                                                    float* __restrict__ M, int64_t n, int d) {
  constexpr int KC = 16;
  __shared__ float As[KC][64];
  __shared__ float Bs[KC][64];
  // TRI: one tile of each symmetric pair (D and M are exactly symmetric:
  // fl(a - b) = -fl(b - a), so every term's square is the same), mirrored
  int64_t bi = blockIdx.y, bj = blockIdx.x;
  if (TRI) tri_tile(blockIdx.x, bi, bj);
  const int64_t i0 = bi * 64;
  const int64_t j0 = bj * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[4][4], lo[4][4], mx[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) { acc[a][b] = 0.0f; lo[a][b] = 0.0f; mx[a][b] = 0.0f; }

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
          if (MAXM) mx[a][b] = fmaxf(mx[a][b], p);
          acc[a][b] = s;
        }
    }
    __syncthreads();
  }
  store_tile(D, n, i0, j0, tx, ty, acc, TRI && bi != bj);
  if (COMP) store_tile(Dlo, n, i0, j0, tx, ty, lo, false);
  if (MAXM) store_tile(M, n, i0, j0, tx, ty, mx, TRI && bi != bj);
}

// M refreshed over the remaining candidates of a backward selection:
// M[i][j] = max over c < ncand of fl((Xc[i][c] - Xc[j][c])^2), the same
// squares the search subtracts, so fl(D - M) stays a lower bound of every
// candidate's distance and tightens as candidates go (the M of the full
// matrix keeps the largest term of features already removed).  Tiles as in
// k_gram_exact, from the point-major candidate columns.
__global__ __launch_bounds__(256) void k_max_square(const float* __restrict__ Xc,
                                                    const SelState* __restrict__ st,
                                                    float* __restrict__ M, int64_t n, int cpad) {
  constexpr int KC = 16;
  __shared__ float As[KC][64];
  __shared__ float Bs[KC][64];
  if (st->stopped) return;
  const int ncand = st->ncand;
  // one tile of each symmetric pair, mirrored (M is symmetric)
  int64_t bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  const int64_t i0 = bi * 64;
  const int64_t j0 = bj * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float mx[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) mx[a][b] = 0.0f;
  for (int c0 = 0; c0 < ncand; c0 += KC) {
    for (int t = threadIdx.x; t < KC * 64; t += 256) {
      const int r = t >> 4, kk = t & 15;  // 16 consecutive columns of a row
      const int c = c0 + kk;
      float av = 0.0f, bv = 0.0f;
      if (c < cpad) {
        if (i0 + r < n) av = Xc[(i0 + r) * cpad + c];
        if (j0 + r < n) bv = Xc[(j0 + r) * cpad + c];
      }
      As[kk][r] = av;
      Bs[kk][r] = bv;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; ++kk) {
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
          mx[a][b] = fmaxf(mx[a][b], t * t);
        }
    }
    __syncthreads();
  }
  store_tile(M, n, i0, j0, tx, ty, mx, bi != bj);
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
typedef float v4f __attribute__((ext_vector_type(4)));

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
      // four float4 of the row in flight per thread; D streams through once
      // (non-temporal), leaving the caches to the gathered feature values
      auto upd = [&](v4f& dv, v4f& lv, const v4f xj) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float t = xi - xj[e];
          const float p = t * t;
          const float ps = mult > 0 ? p : -p;
          const float s = dv[e] + ps;
          if (COMP) lv[e] += comp_error(xi, xj[e], t, p, dv[e], ps, s, mult < 0);
          dv[e] = s;
        }
      };
      constexpr int U = 4;
      int64_t j = (int64_t)threadIdx.x * 4;
      for (; j + (U - 1) * 1024 < n; j += U * 1024) {
        v4f dv[U], lv[U], xj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          xj[u] = *reinterpret_cast<const v4f*>(xf + j + u * 1024);
          dv[u] = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j + u * 1024));
          if (COMP) lv[u] = *reinterpret_cast<v4f*>(rlo + j + u * 1024);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          upd(dv[u], lv[u], xj[u]);
          __builtin_nontemporal_store(dv[u], reinterpret_cast<v4f*>(row + j + u * 1024));
          if (COMP) *reinterpret_cast<v4f*>(rlo + j + u * 1024) = lv[u];
        }
      }
      for (; j < n; j += 1024) {
        v4f lv = {0.f, 0.f, 0.f, 0.f};
        const v4f xj = *reinterpret_cast<const v4f*>(xf + j);
        v4f dv = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j));
        if (COMP) lv = *reinterpret_cast<v4f*>(rlo + j);
        upd(dv, lv, xj);
        __builtin_nontemporal_store(dv, reinterpret_cast<v4f*>(row + j));
        if (COMP) *reinterpret_cast<v4f*>(rlo + j) = lv;
      }
    } else {
      // n % 4 != 0: rows start at any 4-byte boundary.  A scalar head up to
      // the row's first 16-byte boundary, float4 over the body (D, and Dlo,
      // which shares the layout), a scalar tail; the feature's values are
      // read one by one (its column has an alignment of its own)
      auto one = [&](int64_t j) {
        const float xj = xf[j];
        const float t = xi - xj;
        const float p = t * t;
        const float ps = mult > 0 ? p : -p;
        const float dv = row[j];
        const float s = dv + ps;
        if (COMP) rlo[j] += comp_error(xi, xj, t, p, dv, ps, s, mult < 0);
        row[j] = s;
      };
      const int64_t mis = (i * n) & 3;
      const int64_t head = (mis ? 4 - mis : 0) < n ? (mis ? 4 - mis : 0) : n;
      const int64_t end = head + (n - head) / 4 * 4;  // [head, end): whole float4
      if ((int64_t)threadIdx.x < head) one(threadIdx.x);
      for (int64_t j = head + (int64_t)threadIdx.x * 4; j < end; j += 1024) {
        v4f dv = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j));
        v4f lv = {0.f, 0.f, 0.f, 0.f};
        if (COMP) lv = *reinterpret_cast<v4f*>(rlo + j);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float xj = xf[j + e];
          const float t = xi - xj;
          const float p = t * t;
          const float ps = mult > 0 ? p : -p;
          const float sv = dv[e] + ps;
          if (COMP) lv[e] += comp_error(xi, xj, t, p, dv[e], ps, sv, mult < 0);
          dv[e] = sv;
        }
        __builtin_nontemporal_store(dv, reinterpret_cast<v4f*>(row + j));
        if (COMP) *reinterpret_cast<v4f*>(rlo + j) = lv;
      }
      if (end + (int64_t)threadIdx.x < n) one(end + threadIdx.x);
    }
  }
}

// M rounded up to 8 bits per value with a scale per row, the bound the
// distance update reads to list a backward row's points: code c of row i
// stands for 2^(E_i - 16 + (c >> 4)) * (1 + (c & 15) / 16), the smallest such
// value >= M_ij, with E_i = floor(log2 max_j M_ij) + 2 (at least -100), so a
// row's values down to 2^-14 of its maximum keep four mantissa bits and the
// smaller ones round up to 2^(E_i - 16).  Mq >= M, so fl(D - Mq) <= fl(D - M)
// is still a lower bound of every remaining candidate's distance, at a
// quarter of M's bytes.  A row whose maximum is not finite or reaches 2^126
// has E_i = kMqInf and reads as +inf (every point listed: no list).
constexpr int kMqInf = 1 << 20;

__device__ __forceinline__ unsigned int mq_encode(float v, int E) {
  const unsigned int b = __float_as_uint(v);
  const int base = E - 16;
  const int e = (int)((b >> 23) & 0xFFu) - 127;
  if (e < base) return 0u;  // 2^base >= v
  int ef = e - base;
  unsigned int m = (b >> 19) & 15u;
  if (b & 0x7FFFFu) {
    if (++m == 16u) {
      m = 0u;
      ++ef;
    }
  }
  return ef > 15 ? 255u : ((unsigned int)ef << 4) | m;  // ef > 15: only for non-finite M
}

__device__ __forceinline__ float mq_decode(unsigned int c, int E) {
  return __uint_as_float(((unsigned int)(E - 16 + (int)(c >> 4) + 127) << 23) | ((c & 15u) << 19));
}

// Four codes (4-byte aligned): loaded once (streamed), widened to float.
__device__ __forceinline__ unsigned int mq_raw(const unsigned char* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const unsigned int*>(p));
}
__device__ __forceinline__ v4f mq_widen(unsigned int w, int E) {
  if (E == kMqInf) {
    const float inf = __int_as_float(0x7f800000);
    return v4f{inf, inf, inf, inf};
  }
  return v4f{mq_decode(w & 255u, E), mq_decode((w >> 8) & 255u, E),
             mq_decode((w >> 16) & 255u, E), mq_decode(w >> 24, E)};
}

// One wave per row (n % 4 == 0): the row's maximum, then its codes.
__global__ __launch_bounds__(256) void k_quantize_m(const float* __restrict__ M,
                                                    unsigned char* __restrict__ Mq,
                                                    int* __restrict__ Mexp, int64_t n) {
  const int lane = threadIdx.x & 63;
  for (int64_t i = (int64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6); i < n;
       i += (int64_t)gridDim.x * 4) {
    const float* __restrict__ row = M + i * n;
    float mx = 0.0f;
    bool bad = false;
    for (int64_t j = (int64_t)lane * 4; j < n; j += 256) {
      const v4f v = *reinterpret_cast<const v4f*>(row + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        bad |= !(v[e] < 0x1p126f);  // NaN, inf or too large
        mx = fmaxf(mx, v[e]);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    int E;
    if (__any(bad)) {
      E = kMqInf;
    } else {
      const int em = (int)((__float_as_uint(mx) >> 23) & 0xFFu) - 127;
      E = em + 2 > -100 ? em + 2 : -100;
    }
    if (lane == 0) Mexp[i] = E;
    unsigned char* __restrict__ out = Mq + i * n;
    for (int64_t j = (int64_t)lane * 4; j < n; j += 256) {
      const v4f v = *reinterpret_cast<const v4f*>(row + j);
      const unsigned int w = E == kMqInf ? 0u
                             : mq_encode(v[0], E) | (mq_encode(v[1], E) << 8) |
                                   (mq_encode(v[2], E) << 16) | (mq_encode(v[3], E) << 24);
      *reinterpret_cast<unsigned int*>(out + j) = w;
    }
  }
}

// The update fused with the next step's seed scan (n % 4 == 0, feature from
// the device state): one wave per row i updates the row exactly as k_update
// does and keeps, per lane, the two smallest new D_ij outside i's fold (lane
// L is offered the points 4L .. 4L + 3 of every 256); the row's 128 seeds are
// then written ranked by D, as k_seed_scan writes them.  Saves the seed
// scan's full read of D on every pruned step.
//
// With rl.idx set the row's list for the next search (RowList) is written as
// well.  Its bound is the largest warm bound over the next step's candidates
// (cand[c < ncand], k_select has listed them): forward fl(max tau + Q)
// widened as in the search's warm bound, backward max tau itself; no list
// where some candidate's tau is unknown (kMaxF) or X is not finite.  The
// backward lower bound fl(D - Mq) uses M rounded up to 8 bits (Mq >= M),
// from the rows as they stand; a later refresh only raises the bound, so the
// list still holds every point the search can visit.
//
// DEF (deferred writes, backward lists): 0 applies fstar and writes D; 1
// applies it, writes the seeds and lists but not D, and records fstar in
// st->fdefer (D in memory then lacks that one term, and the next step's
// readers subtract it themselves, in the same arithmetic); 2 applies the
// recorded feature first, then fstar, and writes D.  Alternating 1 and 2
// saves every other write of the n^2 floats, each step's values unchanged.
// SEEDS: write the seeds (false for backward lists: every backward search
// after the first has a warm bound for each candidate and reads no seed, and
// where one would, the seeds already there -- eligible points, as any K
// distinct ones bound the K-th distance -- still give a valid bound).
template <bool COMP, int DEF, bool SEEDS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COMP ? 1 : 4)))
void k_update_seeds(float* __restrict__ D,
                                                      float* __restrict__ Dlo,
                                                      const float* __restrict__ XT, int64_t n,
                                                      SelState* __restrict__ st, int mult,
                                                      int* __restrict__ seeds, int64_t npf,
                                                      RowList rl, const float* __restrict__ tauw,
                                                      const float* __restrict__ frange, int d,
                                                      const int* __restrict__ cand,
                                                      const int* __restrict__ iperm,
                                                      const unsigned char* __restrict__ Mq,
                                                      const int* __restrict__ Mexp) {
  if (!st->pending) return;
  const int f = st->fstar;
  const int lane = threadIdx.x & 63;
  const float kInf = __int_as_float(0x7f800000);
  const float* __restrict__ xf = XT + (int64_t)f * n;
  // DEF 2: the feature the previous step left out of D
  const float* __restrict__ xp = DEF == 2 ? XT + (int64_t)st->fdefer * n : nullptr;
  if (DEF == 1 && blockIdx.x == 0 && threadIdx.x == 0) st->fdefer = f;  // read by later launches
  for (int64_t i = (int64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6); i < n;
       i += (int64_t)gridDim.x * 4) {
    const float xi = xf[i];
    const float xip = DEF == 2 ? xp[i] : 0.0f;
    float* __restrict__ row = D + i * n;
    float* __restrict__ rlo = COMP ? Dlo + i * n : nullptr;
    const int64_t flo = (i / npf) * npf;
    const int64_t fhi = flo + npf < n ? flo + npf : n;
    SeedPair sp;
    sp.reset();
    // the row list's bound: the largest warm bound over the candidate blocks
    // it serves.  Leading blocks (the largest-scale candidates come first)
    // whose largest bound is more than twice that of every later block are
    // left out, and their search waves scan the row: one loose block would
    // otherwise make every list too long for the others.
    float Tq = -kInf;  // -inf: no list
    if (rl.idx && !st->nonfinite) {
      const int nc = st->ncand;
      const int nb = (nc + 63) / 64;
      // lane b: block b's largest warm bound and whether one is unknown
      float mb = 0.0f;
      bool ub = false;
      if (nb <= 16) {
        // four lanes per block, 16 candidates each, then two shuffles
        const int b = lane >> 2;
        float t = 0.0f;
        int unk = 0;
        if (b < nb) {
          // the three dependent loads of all 16 candidates in flight together
          int cs[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            const int c = b * 64 + (lane & 3) + 4 * k;
            cs[k] = c < nc ? cand[c] : -1;
          }
#pragma unroll
          for (int k = 0; k < 16; ++k) cs[k] = cs[k] >= 0 ? iperm[cs[k]] : -1;
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            if (cs[k] >= 0) {
              const float v = tauw[i * d + cs[k]];
              if (v < kMaxF) t = fmaxf(t, v);
              else unk = 1;
            }
          }
        }
        t = fmaxf(t, __shfl_xor(t, 1));
        t = fmaxf(t, __shfl_xor(t, 2));
        unk |= __shfl_xor(unk, 1);
        unk |= __shfl_xor(unk, 2);
        mb = __shfl(t, (4 * lane) & 63);
        ub = __shfl(unk, (4 * lane) & 63) != 0;
      } else
      for (int b = 0; b < nb; ++b) {
        const int c = b * 64 + lane;
        float t = 0.0f;
        bool unk = false;
        if (c < nc) {
          t = tauw[i * d + iperm[cand[c]]];
          unk = !(t < kMaxF);
          if (unk) t = 0.0f;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t = fmaxf(t, __shfl_xor(t, o));
        const bool bu = __any(unk);
        if ((b & 63) == lane) {
          mb = fmaxf(mb, t);
          ub = ub || bu;
        }
      }
      // suffix maxima over blocks b' >= b (lane b), then the first block kept
      float sfx = lane < nb ? mb : 0.0f;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const float v = __shfl_down(sfx, o);
        if (lane + o < 64) sfx = fmaxf(sfx, v);
      }
      const float after = __shfl_down(sfx, 1);  // max over blocks > b
      const bool drop = nb <= 64 && lane + 1 < nb && (ub || mb > 2.0f * after);
      const unsigned long long keep = __ballot(!drop);  // lanes >= nb count as kept
      const int b0 = keep ? (int)__builtin_ctzll(keep) : 0;
      const float m = __shfl(sfx, b0);
      const bool unknown = __any(ub && lane >= b0 && lane < nb) || nb > 64;
      if (!unknown) {
        if (mult < 0) {
          Tq = m;
        } else {
          const float lo = xi - frange[0], hi = frange[1] - xi;
          const float w = fmaf(m + fmaxf(lo * lo, hi * hi), 1.00000095367431640625f, 1.0e-36f);
          if (w < kInf) Tq = w;
        }
      }
    }
    int64_t lcount = 0;  // list entries written (wave-uniform)
    int* __restrict__ lidx = rl.idx ? rl.idx + i * rl.cap : nullptr;
    float* __restrict__ lval = rl.idx ? rl.val + i * rl.cap : nullptr;
    float* __restrict__ llb = rl.lb ? rl.lb + i * rl.cap : nullptr;
    const unsigned char* __restrict__ mrow = rl.lb ? Mq + i * n : nullptr;  // backward lists
    const int Erow = rl.lb ? Mexp[i] : 0;
    // in: the lane holds four points of the row (all lanes call this together);
    // mv: their M values (backward lists)
    v4f lbv = {0.f, 0.f, 0.f, 0.f};
    // FOLD (compile time): the lane's points may lie in row i's fold; a block
    // of points wholly outside it skips the per-point fold tests
    auto upd = [&](v4f& dv, v4f& lv, const v4f xj, const v4f mv, const v4f pj, int64_t j,
                   bool in, auto fold_tag) {
      constexpr bool FOLD = decltype(fold_tag)::value;
      int keep = 0;
      if (in) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (DEF == 2) {  // the deferred term first, as the previous step would have
            const float t0 = xip - pj[e];
            const float p0 = t0 * t0;
            dv[e] = dv[e] + (mult > 0 ? p0 : -p0);
          }
          const float t = xi - xj[e];
          const float p = t * t;
          const float ps = mult > 0 ? p : -p;
          const float s = dv[e] + ps;
          if (COMP) lv[e] += comp_error(xi, xj[e], t, p, dv[e], ps, s, mult < 0);
          dv[e] = s;
          if (SEEDS) {
            if (FOLD) sp.offer(s, j + e, flo, fhi);
            else sp.offer_eligible(s, j + e);
          }
          lbv[e] = mrow ? s - mv[e] : s;
          keep |= (lbv[e] <= Tq && (!FOLD || j + e < flo || j + e >= fhi)) ? 1 << e : 0;
        }
      }
      // most blocks of a row add nothing to its list: one ballot then
      if (Tq != -kInf && __ballot(keep != 0)) {
        // entries in ascending index: lane L's four points precede lane L+1's
        const int c = __builtin_popcount(keep);
        int pre = 0, tot = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const unsigned long long m = __ballot((c >> b) & 1);
          pre += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
          tot += (int)__popcll(m) << b;
        }
        if (lcount + tot > rl.cap) {
          Tq = -kInf;  // too many: no list for this row
          lcount = -1;
        } else {
          int pos = (int)lcount + pre;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (keep & (1 << e)) {
              lidx[pos] = (int)(j + e);
              lval[pos] = dv[e];
              if (llb) llb[pos] = lbv[e];
              ++pos;
            }
          }
          lcount += tot;
        }
      }
    };
    // every lane runs the same iterations (the list's ballots need the whole
    // wave): blocks of U * 256 points, then 256-point chunks in which lanes
    // past the row's end hold no point (n % 4 == 0: a lane's four points are
    // all in the row or all past it)
    constexpr int U = 4;
    auto block = [&](int64_t jb, auto fold_tag) {
      const int64_t j = jb + (int64_t)lane * 4;
      v4f dv[U], lv[U], xj[U], mv[U], pj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xj[u] = *reinterpret_cast<const v4f*>(xf + j + u * 256);
        dv[u] = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j + u * 256));
        if (COMP) lv[u] = *reinterpret_cast<v4f*>(rlo + j + u * 256);
        mv[u] = mrow ? mq_widen(mq_raw(mrow + j + u * 256), Erow) : v4f{0.f, 0.f, 0.f, 0.f};
        pj[u] = DEF == 2 ? *reinterpret_cast<const v4f*>(xp + j + u * 256)
                         : v4f{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        upd(dv[u], lv[u], xj[u], mv[u], pj[u], j + u * 256, true, fold_tag);
        if (DEF != 1)
          __builtin_nontemporal_store(dv[u], reinterpret_cast<v4f*>(row + j + u * 256));
        if (COMP) *reinterpret_cast<v4f*>(rlo + j + u * 256) = lv[u];
      }
    };
    int64_t jb = 0;
    for (; jb + U * 256 <= n; jb += U * 256) {
      if (jb < fhi && jb + U * 256 > flo) block(jb, std::true_type{});
      else block(jb, std::false_type{});
    }
    for (; jb < n; jb += 256) {
      const int64_t j = jb + (int64_t)lane * 4;
      const bool in = j < n;
      v4f lv = {0.f, 0.f, 0.f, 0.f}, xj = {0.f, 0.f, 0.f, 0.f}, dv = {0.f, 0.f, 0.f, 0.f};
      v4f mv = {0.f, 0.f, 0.f, 0.f}, pj = {0.f, 0.f, 0.f, 0.f};
      if (in) {
        xj = *reinterpret_cast<const v4f*>(xf + j);
        dv = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j));
        if (COMP) lv = *reinterpret_cast<v4f*>(rlo + j);
        if (mrow) mv = mq_widen(mq_raw(mrow + j), Erow);
        if (DEF == 2) pj = *reinterpret_cast<const v4f*>(xp + j);
      }
      upd(dv, lv, xj, mv, pj, j, in, std::true_type{});
      if (in) {
        if (DEF != 1) __builtin_nontemporal_store(dv, reinterpret_cast<v4f*>(row + j));
        if (COMP) *reinterpret_cast<v4f*>(rlo + j) = lv;
      }
    }
    if (SEEDS) sp.write_ranked(seeds + i * 128);
    if (rl.idx && lane == 0) {
      rl.count[i] = Tq == -kInf ? -1 : (int)lcount;
      rl.bound[i] = Tq;
    }
  }
}

// The update fused with the next step's seed scan for rows at any 4-byte
// boundary (n % 4 != 0; no row lists): one wave per row, a scalar head up to
// the row's first 16-byte boundary, float4 over the body, a scalar tail, each
// new D_ij offered to the lane's two-smallest pair as k_update_seeds offers
// it; the row's 128 seeds are written ranked by D.
template <bool COMP>
__global__ __launch_bounds__(256) void k_update_seeds_any(float* __restrict__ D,
                                                          float* __restrict__ Dlo,
                                                          const float* __restrict__ XT,
                                                          int64_t n,
                                                          const SelState* __restrict__ st,
                                                          int mult, int* __restrict__ seeds,
                                                          int64_t npf) {
  if (!st->pending) return;
  const int lane = threadIdx.x & 63;
  const float* __restrict__ xf = XT + (int64_t)st->fstar * n;
  for (int64_t i = (int64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6); i < n;
       i += (int64_t)gridDim.x * 4) {
    const float xi = xf[i];
    float* __restrict__ row = D + i * n;
    float* __restrict__ rlo = COMP ? Dlo + i * n : nullptr;
    const int64_t flo = (i / npf) * npf;
    const int64_t fhi = flo + npf < n ? flo + npf : n;
    SeedPair sp;
    sp.reset();
    auto term = [&](float xj, float dv, float& lo) {
      const float t = xi - xj;
      const float p = t * t;
      const float ps = mult > 0 ? p : -p;
      const float sv = dv + ps;
      if (COMP) lo += comp_error(xi, xj, t, p, dv, ps, sv, mult < 0);
      return sv;
    };
    auto one = [&](int64_t j) {
      float lo = COMP ? rlo[j] : 0.0f;
      const float sv = term(xf[j], row[j], lo);
      row[j] = sv;
      if (COMP) rlo[j] = lo;
      sp.offer(sv, j, flo, fhi);
    };
    const int64_t mis = (i * n) & 3;
    const int64_t head = (mis ? 4 - mis : 0) < n ? (mis ? 4 - mis : 0) : n;
    const int64_t end = head + (n - head) / 4 * 4;  // [head, end): whole float4
    if (lane < head) one(lane);
    for (int64_t j = head + (int64_t)lane * 4; j < end; j += 256) {
      v4f dv = __builtin_nontemporal_load(reinterpret_cast<v4f*>(row + j));
      v4f lv = {0.f, 0.f, 0.f, 0.f};
      if (COMP) lv = *reinterpret_cast<v4f*>(rlo + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float lo = lv[e];
        dv[e] = term(xf[j + e], dv[e], lo);
        lv[e] = lo;
        sp.offer(dv[e], j + e, flo, fhi);
      }
      __builtin_nontemporal_store(dv, reinterpret_cast<v4f*>(row + j));
      if (COMP) *reinterpret_cast<v4f*>(rlo + j) = lv;
    }
    if (end + lane < n) one(end + lane);
    sp.write_ranked(seeds + i * 128);
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
  // slots of the candidate blocks that hold live candidates (the searches
  // return before reading a block past the last live candidate)
  const int cw = (nc + 63) / 64 * 64 < cpad ? (nc + 63) / 64 * 64 : cpad;
  // rows across blocks, candidate slots across threads (no index division)
  for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
    const float* __restrict__ xr = X + j * d;
    float* __restrict__ out = Xc + j * cpad;
    for (int c = threadIdx.x; c < cw; c += blockDim.x) out[c] = c < nc ? xr[cand[c]] : 0.0f;
  }
}

// ------------------------------------------------------------------ losses

// Sequential float sum over q = 0..n-1 (or n-1..0) of per-point terms, in
// exactly that order, one wave: lanes compute 64 terms at a time, which are
// added one by one (readlane) so the rounding matches a scalar loop.  The
// next 256 terms are loaded while the current 256 are added.
template <bool REV, class Term>
__device__ float seq_sum(int64_t n, Term term) {
  const int lane = threadIdx.x & 63;
  constexpr int U = 4;  // 64*U terms per batch
  auto fetch = [&](int64_t k0, float (&t)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + u * 64 + lane;  // k-th term in summation order
      t[u] = k < n ? term(REV ? n - 1 - k : k) : 0.0f;
    }
  };
  float s = 0.0f;
  float cur[U], nxt[U];
  fetch(0, nxt);
  for (int64_t k0 = 0; k0 < n; k0 += 64 * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) cur[u] = nxt[u];
    if (k0 + 64 * U < n) fetch(k0 + 64 * U, nxt);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t kb = k0 + u * 64;
      if (kb >= n) break;
      const int cnt = (int)((n - kb) < 64 ? (n - kb) : 64);
      for (int i = 0; i < cnt; ++i) s = s + lane_value(cur[u], i);
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

// loss_functions.c:41-50
__device__ float loss_rmse(const float* __restrict__ y, const float* __restrict__ p, int64_t n) {
  const float s = seq_sum<false>(n, [&](int64_t q) {
    const float t = y[q] - p[q];
    return t * t;
  });
  return (float)sqrt((double)(s / (float)n));
}

// loss_functions.c:52-62 (summed from the last point down)
__device__ float loss_rmse_norm(const float* __restrict__ y, const float* __restrict__ p,
                                int64_t n) {
  const float s = seq_sum<true>(n, [&](int64_t q) {
    const float tv = y[q];
    const float t = (float)((double)(tv - p[q]) / (1.0 + (double)tv));
    return t * t;
  });
  return (float)sqrt((double)(s / (float)n));
}

// The same sums with one candidate per lane: lane c adds its own terms in the
// reference's order, so one VALU add serves 64 candidates and no term goes
// through readlane.  The 64 candidates' rows are processed in tiles of 64
// points: kLossLoaders loader waves read the tile's rows coalesced (256 bytes
// per row), evaluate the terms and store them transposed in LDS; the last
// wave of the block only adds, lane c along row c of the tile, while the
// loaders stage the next tile (double-buffered, one barrier per tile) and
// fetch the one kLossAhead - 1 tiles further on.
// term(y_q, p_q) gives the addend of point q from its label and prediction;
// both are loaded kLossAhead - 1 tiles ahead.
constexpr int kLossTile = 64;
constexpr int kLossLoaders = 8;
constexpr int kLossRows = 64 / kLossLoaders;   // rows per loader wave
constexpr int kLossStride = kLossTile + 4;     // 16-byte aligned rows, distinct banks
constexpr int kLossThreads = (kLossLoaders + 1) * 64;
constexpr int kLossAhead = 4;                  // tiles a loader holds in registers
constexpr int kLossLdsFloats = 2 * 64 * kLossStride;

template <bool REV, class Term>
__device__ __forceinline__ float lane_seq_sum(const float* __restrict__ P0,
                                              const float* __restrict__ y, int64_t n,
                                              float* __restrict__ tiles, Term term) {
  // wave-uniform, so the loader / adder split is a scalar branch
  const int wave = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t ntile = (n + kLossTile - 1) / kLossTile;
  float s = 0.0f;
  // the adder's chain of dependent adds is the critical path: its wave wins
  // issue over the loaders that share its SIMD
  if (wave == kLossLoaders) __builtin_amdgcn_s_setprio(3);
  // loaders keep tile u in register slot u % kLossAhead: the loads of a tile
  // are issued kLossAhead - 1 tiles before it is staged, which hides the
  // memory latency behind the adder's work on the tiles in between
  float raw[kLossAhead][kLossRows];
  float ys[kLossAhead];  // the label of the lane's point, loaded with its rows
  // Past the last point a lane loads point n - 1 (stage() writes 0 there):
  // the loads stay unconditional, so none of them sits in a branch whose end
  // would wait for every load in flight.
  auto fetch = [&](int64_t t, float (&r_)[kLossRows], float& yv) {  // loader: its rows of tile t
    const int64_t k = t * kLossTile + lane;  // k-th point in summation order
    const int64_t kc = k < n ? k : n - 1;
    const int64_t q = REV ? n - 1 - kc : kc;
    const int lw = wave < kLossLoaders ? wave : 0;  // the adder loads loader 0's rows (unused)
    const float* base = P0 + (int64_t)(lw * kLossRows) * n + q;
#pragma unroll
    for (int r = 0; r < kLossRows; ++r) r_[r] = base[(int64_t)r * n];
    yv = y[q];
  };
  auto stage = [&](int64_t t, const float (&r_)[kLossRows], float yv, float* buf) {
    const int64_t k = t * kLossTile + lane;
#pragma unroll
    for (int r = 0; r < kLossRows; ++r)
      buf[(wave * kLossRows + r) * kLossStride + lane] = k < n ? term(yv, r_[r]) : 0.0f;
  };
  if (wave < kLossLoaders) {
#pragma unroll
    for (int u = 0; u < kLossAhead; ++u)
      if (u < ntile) fetch(u, raw[u], ys[u]);
    stage(0, raw[0], ys[0], tiles);
  }
  __syncthreads();
  // the adder's work on one full tile: all of the row's reads first, then the
  // dependent adds, so the LDS latency is paid once per tile instead of once
  // per four terms
  auto add_full = [&](const float* row) {
    float4 v[kLossTile / 4];
#pragma unroll
    for (int i = 0; i < kLossTile / 4; ++i) v[i] = *reinterpret_cast<const float4*>(row + 4 * i);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < kLossTile / 4; ++i) {
      s = s + v[i].x;
      s = s + v[i].y;
      s = s + v[i].z;
      s = s + v[i].w;
    }
  };
  // one tile: slot U = t % kLossAhead is a compile-time index of raw.  FULL:
  // tiles t .. t + kLossAhead exist and t is whole, so the step has no
  // branch but the loader / adder split (a branch around the loads would make
  // the compiler wait for every load in flight, or copy the slots)
  auto tile_step = [&](int64_t t, auto slot, auto full) {
    constexpr int U = decltype(slot)::value;
    constexpr bool FULL = decltype(full)::value;
    if (!FULL && t >= ntile) return;  // the same for every wave of the block
    const float* cur = tiles + (t & 1) * 64 * kLossStride;
    if (wave == kLossLoaders) {
      const float* row = cur + lane * kLossStride;
      if (FULL || n - t * kLossTile >= kLossTile) {
        add_full(row);
      } else {
        const int cnt = (int)(n - t * kLossTile);
        for (int i = 0; i < cnt; ++i) s = s + row[i];
      }
    } else if (FULL || t + 1 < ntile) {
      stage(t + 1, raw[(U + 1) % kLossAhead], ys[(U + 1) % kLossAhead],
            tiles + ((t + 1) & 1) * 64 * kLossStride);
    }
    if (FULL) {
      // tile t's slot is free.  Every wave loads (the adder's loads are never
      // used), so the slot registers are written outside the loader branch
      // and the compiler neither copies them nor waits on them at its end.
      fetch(t + kLossAhead, raw[U], ys[U]);
    } else if (wave != kLossLoaders && t + 1 < ntile) {
      // past the last tile the loads repeat it (never staged)
      fetch(t + kLossAhead < ntile ? t + kLossAhead : ntile - 1, raw[U], ys[U]);
    }
    __syncthreads();  // tile t consumed, tile t + 1 staged
  };
  static_assert(kLossAhead == 4, "tile_step is unrolled for four slots");
  using Yes = std::integral_constant<bool, true>;
  using No = std::integral_constant<bool, false>;
  int64_t t0 = 0;
  for (; t0 + 2 * kLossAhead <= ntile; t0 += kLossAhead) {
    tile_step(t0, std::integral_constant<int, 0>{}, Yes{});
    tile_step(t0 + 1, std::integral_constant<int, 1>{}, Yes{});
    tile_step(t0 + 2, std::integral_constant<int, 2>{}, Yes{});
    tile_step(t0 + 3, std::integral_constant<int, 3>{}, Yes{});
  }
  for (; t0 < ntile; t0 += kLossAhead) {
    tile_step(t0, std::integral_constant<int, 0>{}, No{});
    tile_step(t0 + 1, std::integral_constant<int, 1>{}, No{});
    tile_step(t0 + 2, std::integral_constant<int, 2>{}, No{});
    tile_step(t0 + 3, std::integral_constant<int, 3>{}, No{});
  }
  if (wave == kLossLoaders) __builtin_amdgcn_s_setprio(0);
  return s;
}

__host__ __device__ inline bool lane_loss(int code) {
  return code == kRmse || code == kRmseNorm || code == kStd || code == kRmseStd;
}

// loss_functions.c:41-50, 52-62, 129-145 and :33-34, one candidate per lane:
// lane r's predictions are row r of P0.  Called by the whole block; the value
// is that of the adding wave (wave kLossLoaders).
__device__ float lane_loss_value(int code, const float* __restrict__ y,
                                 const float* __restrict__ P0, int64_t n,
                                 float* __restrict__ tile) {
  const float* __restrict__ p = P0 + (int64_t)(threadIdx.x & 63) * n;
  auto rmse = [&]() {
    const float s = lane_seq_sum<false>(P0, y, n, tile, [&](float yq, float pq) {
      const float t = yq - pq;
      return t * t;
    });
    return (float)sqrt((double)(s / (float)n));
  };
  auto stdv = [&]() {
    float mean = lane_seq_sum<false>(P0, y, n, tile, [&](float yq, float pq) { return yq - pq; });
    mean = mean / (float)n;
    const float t = (y[n - 1] - p[n - 1]) - mean;
    const float s = (t * t) / (float)(n - 1);
    return (float)sqrt((double)s);
  };
  switch (code) {
    case kRmse: return rmse();
    case kRmseNorm: {
      const float s = lane_seq_sum<true>(P0, y, n, tile, [&](float tv, float pq) {
        const float t = (float)((double)(tv - pq) / (1.0 + (double)tv));
        return t * t;
      });
      return (float)sqrt((double)(s / (float)n));
    }
    case kStd: return stdv();
    case kRmseStd: return rmse() + stdv();
    default: return __int_as_float(0x7fc00000);
  }
}

// loss_functions.c:129-145, keeping the assignment at :141 (only the last
// squared deviation survives)
__device__ float loss_std(const float* __restrict__ y, const float* __restrict__ p, int64_t n) {
  float mean = seq_sum<false>(n, [&](int64_t q) { return y[q] - p[q]; });
  mean = mean / (float)n;
  const float t = (y[n - 1] - p[n - 1]) - mean;
  const float s = (t * t) / (float)(n - 1);
  return (float)sqrt((double)s);
}

// Order statistic `kth` (0-based) of value(i), i < n, one wave: radix select
// over a total order of the float bits, 8 bits per pass, histograms in LDS.
// The sorted position the reference reads (qsort, loss_functions.c:173-183)
// holds this value; only +0 and -0 could tie without being equal bits, and
// both callers are indifferent to the sign of a zero median.
template <class Value>
__device__ float wave_select(int64_t n, int64_t kth, unsigned int* hist, Value value) {
  const int lane = threadIdx.x & 63;
  auto key = [](float v) {
    const unsigned int b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  };
  unsigned int prefix = 0, mask = 0;
  int64_t rank = kth;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int b = lane; b < 256; b += 64) hist[b] = 0;
    __syncthreads();
    for (int64_t i = lane; i < n; i += 64) {
      const unsigned int k = key(value(i));
      if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    int64_t cum = 0;
    unsigned int sel = 255;
    for (int b = 0; b < 256; ++b) {  // every lane scans the same LDS histogram
      const int64_t h = hist[b];
      if (cum + h > rank) {
        sel = (unsigned int)b;
        break;
      }
      cum += h;
    }
    rank -= cum;
    prefix |= sel << shift;
    mask |= 255u << shift;
    __syncthreads();
  }
  const unsigned int b = (prefix & 0x80000000u) ? (prefix & 0x7fffffffu) : ~prefix;
  return __uint_as_float(b);
}

// Element n/2 of buf[0, n) in the order the reference's qsort leaves it
// (qsort_order.h), tmp of n floats; the block is one wave.  Lanes sort the 64
// subtrees at depth 6 of the merge recursion, one each, then merge level by
// level; the barriers also make each level's global-memory writes visible to
// the next.
__device__ float reference_median(float* buf, float* tmp, int64_t n) {
  constexpr int kDepth = 6;
  const int lane = threadIdx.x & 63;
  sort_subtree(buf, tmp, n, kDepth, lane);
  __syncthreads();
  for (int t = kDepth - 1; t >= 0; --t) {
    if (lane < (1 << t)) merge_level_node(buf, tmp, n, t, lane);
    __syncthreads();
  }
  const float m = buf[n / 2];
  __syncthreads();
  return m;
}

// Median (element n/2 of the reference's sorted copy) of value(i), i < n:
// radix select unless some value is NaN, else the reference's sort in
// scratch (2n floats).
template <class Value>
__device__ float median_of(int64_t n, unsigned int* hist, float* scratch, Value value) {
  const int lane = threadIdx.x & 63;
  bool nan = false;
  for (int64_t i = lane; i < n; i += 64) {
    const float v = value(i);
    nan |= v != v;
  }
  if (!__any(nan) || !scratch) return wave_select(n, n / 2, hist, value);
  float* buf = scratch;
  for (int64_t i = lane; i < n; i += 64) buf[i] = value(i);
  __syncthreads();
  return reference_median(buf, scratch + n, n);
}

// loss_functions.c:67-111: median of |e - median(e)|, e = y - p (mad) or
// (y - p) / (1 + y) rounded to float (mad_norm).  scratch: 2n floats.
__device__ float loss_mad(const float* __restrict__ y, const float* __restrict__ p, int64_t n,
                          bool norm, unsigned int* hist, float* scratch) {
  auto err = [&](int64_t q) {
    const float tv = y[q];
    return norm ? (float)((double)(tv - p[q]) / (1.0 + (double)tv)) : tv - p[q];
  };
  const float m1 = median_of(n, hist, scratch, err);
  return median_of(n, hist, scratch, [&](int64_t q) { return fabsf(err(q) - m1); });
}

__device__ float loss_value(int code, const float* __restrict__ y, const float* __restrict__ p,
                            int64_t n, unsigned int* hist, float* scratch) {
  const int lane = threadIdx.x & 63;
  switch (code) {
    case kZeroOne: {  // an integer count: any order; eight loads of each row in flight
      int cnt = 0;
      int64_t q = lane;
      for (; q + 7 * 64 < n; q += 8 * 64) {
        float yv[8], pv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          yv[u] = y[q + u * 64];
          pv[u] = p[q + u * 64];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) cnt += zero_one_err(yv[u], pv[u]);
      }
      for (; q < n; q += 64) cnt += zero_one_err(y[q], p[q]);
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
      return (float)cnt / (float)n;
    }
    case kRmse: return loss_rmse(y, p, n);
    case kRmseNorm: return loss_rmse_norm(y, p, n);
    case kMad: return loss_mad(y, p, n, false, hist, scratch);
    case kMadNorm: return loss_mad(y, p, n, true, hist, scratch);
    case kStd: return loss_std(y, p, n);
    case kRmseNormMadNormStd: {  // loss_functions.c:147-152, summed in double
      const float a = loss_rmse_norm(y, p, n), b = loss_mad(y, p, n, true, hist, scratch);
      const float c = loss_std(y, p, n);
      return (float)((double)a / 0.19 + (double)b / 0.041 + (double)c / 0.16);
    }
    case kRmseNormMadNormGalaxies: {  // loss_functions.c:154-159
      const float a = loss_rmse_norm(y, p, n), b = loss_mad(y, p, n, true, hist, scratch);
      return (float)((double)a / 0.019 + (double)b / 0.009);
    }
    case kRmseStd: return loss_rmse(y, p, n) + loss_std(y, p, n);  // loss_functions.c:33-34
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
                                             int nl, float* __restrict__ scratch) {
  __shared__ unsigned int hist[256];
  if (st->stopped) return;
  const int c = blockIdx.x, m = blockIdx.y;
  if (c >= st->ncand) return;
  const int64_t slot = (int64_t)m * cpad + c;
  const float* p = P + slot * n;
  float* scr = scratch ? scratch + slot * 2 * n : nullptr;  // median losses only
  for (int l = 0; l < nl; ++l) {
    const int code = losses[l];
    if (lane_loss(code)) continue;  // k_loss_lanes
    const float e = loss_value(code, y, p, n, hist, scr);
    if (threadIdx.x == 0) err[((int64_t)l * nk + m) * cpad + c] = e;
  }
}

// The sequential-sum losses (rmse, rmse_norm, std, rmse_std): one candidate
// per lane of the adding wave, 64 candidates per block, blockIdx.y = k value.
__global__ __launch_bounds__(kLossThreads) void k_loss_lanes(const float* __restrict__ P,
                                                   const float* __restrict__ y,
                                                   const int* __restrict__ losses,
                                                   const SelState* __restrict__ st,
                                                   float* __restrict__ err, int64_t n, int cpad,
                                                   int nk, int nl) {
  __shared__ __align__(16) float tile[kLossLdsFloats];
  if (st->stopped) return;
  const int ncand = st->ncand;
  if ((int)blockIdx.x * 64 >= ncand) return;
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), m = blockIdx.y;
  const bool adder = (threadIdx.x >> 6) == kLossLoaders;
  const float* P0 = P + ((int64_t)m * cpad + blockIdx.x * 64) * n;  // 64 rows, all < cpad
  for (int l = 0; l < nl; ++l) {
    const int code = losses[l];
    if (!lane_loss(code)) continue;
    const float e = lane_loss_value(code, y, P0, n, tile);
    if (adder && c < ncand) err[((int64_t)l * nk + m) * cpad + c] = e;
  }
}

// Loss of one prediction vector (get_loss_all); scratch: 2n floats.
__global__ __launch_bounds__(64) void k_loss_single(const float* __restrict__ y,
                                                    const float* __restrict__ p, int64_t n,
                                                    int code, float* __restrict__ out,
                                                    float* __restrict__ scratch) {
  __shared__ unsigned int hist[256];
  const float e = loss_value(code, y, p, n, hist, scratch);
  if (threadIdx.x == 0) *out = e;
}

// ------------------------------------------------------------- arg-min step
//
// One workgroup of 256 threads: best k per candidate (first k value, in
// k_vals order, with the strictly smallest first-loss error;
// speedynn/lib/cpu.c:110-120), arg-min over candidates in ascending feature
// order with strict '<' (nearest_neighbors_feature_selection.c:318-329
// forward, :155-167 backward), records (:336-342 / :174-181), auto-stop
// (:446-465) and the next step's candidate list.  The reference's scan keeps
// the first strictly smaller error, starting from FLT_MAX: the winner is the
// least (error, feature) pair among errors below FLT_MAX, which a tree
// reduction finds in any order.
__global__ __launch_bounds__(256) void k_select(SelState* __restrict__ st, const Problem pb,
                                                const int* __restrict__ kvals,
                                                const float* __restrict__ err,
                                                int* __restrict__ cand, int* __restrict__ sel,
                                                int* __restrict__ ordering,
                                                float* __restrict__ errors, float* __restrict__ val,
                                                int* __restrict__ bestk,
                                                float* __restrict__ cand_err,
                                                const int* __restrict__ perm) {
  __shared__ float s_v[256];
  __shared__ int s_f[256], s_c[256];
  __shared__ int s_count[4];
  __shared__ int s_stop;
  if (st->stopped) return;
  const int t = threadIdx.x;
  const int nc = st->ncand;
  const int cpad = pb.cpad, nk = pb.nk, nl = pb.nl, d = pb.d;
  const int step = st->step;
  const float qnan = __int_as_float(0x7fc00000);
  float bv = kMaxF;
  int bf = 0x7fffffff, bc = -1;
  for (int c = t; c < nc; c += 256) {
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
    const float v = bm >= 0 ? err[(int64_t)bm * cpad + c] : qnan;
    const int f = cand[c];
    if (cand_err) cand_err[(int64_t)step * d + f] = v;
    if (v < kMaxF && (v < bv || (v == bv && f < bf))) {
      bv = v;
      bf = f;
      bc = c;
    }
  }
  s_v[t] = bv;
  s_f[t] = bf;
  s_c[t] = bc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      const float v2 = s_v[t + o];
      const int f2 = s_f[t + o];
      if (s_c[t + o] >= 0 && (s_c[t] < 0 || v2 < s_v[t] || (v2 == s_v[t] && f2 < s_f[t]))) {
        s_v[t] = v2;
        s_f[t] = f2;
        s_c[t] = s_c[t + o];
      }
    }
    __syncthreads();
  }
  const int best = s_c[0];
  if (best < 0) {
    if (t == 0) {
      st->error = 1;
      st->stopped = 2;
    }
    return;
  }
  const bool fwd = pb.method == 1;
  if (t == 0) {
    const int f = cand[best];
    const int pos = st->pos;
    sel[f] = fwd ? 1 : 0;
    ordering[pos] = f;
    for (int l = 0; l < nl; ++l) errors[l * d + pos] = val[l * cpad + best];
    st->best_k = bestk[best];
    st->fstar = f;
    st->fslot = best;
    st->pending = 1;
    st->num_final += fwd ? 1 : -1;
    st->step = step + 1;
    bool stop = false;
    if (pb.auto_feature) {
      if (fwd) stop = pos > 0 && errors[pos - 1] < errors[pos];
      else stop = pos > 0 && pos < d - 1 && errors[pos + 1] < errors[pos];
    }
    const int next = fwd ? pos + 1 : pos - 1;
    if (fwd ? !(next < pb.num_desired) : !(next >= pb.num_desired)) stop = true;
    st->pos = next;
    s_stop = stop ? 1 : 0;
  }
  __syncthreads();  // sel[f] written; every read of cand[] above is done
  // next candidate list: the unselected (forward) or selected (backward)
  // features in slot order, compacted with a ballot per wave
  const int check = fwd ? 0 : 1;
  const int wave = t >> 6;
  int nc2 = 0;
  for (int base = 0; base < d; base += 256) {
    const int gi = base + t;
    const int g = gi < d ? perm[gi] : 0;
    const bool keep = gi < d && sel[g] == check;
    const unsigned long long m = __ballot(keep);
    if ((t & 63) == 0) s_count[wave] = (int)__popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      off += w < wave ? s_count[w] : 0;
      tot += s_count[w];
    }
    if (keep)
      cand[nc2 + off +
           (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = g;
    nc2 += tot;
    __syncthreads();
  }
  if (t == 0) {
    st->ncand = nc2;
    st->stopped = (s_stop || nc2 == 0) ? 1 : 0;
  }
}

// Candidate slot order.  The pruned search bounds a block of 64 candidates by
// the largest bound among them, so features of similar scale are grouped:
// slots follow the column's sum of squared deviations (float64, fixed
// reduction order, hence deterministic) in descending order, ties by index.
// Descending, the partly filled last block holds the smallest features, and
// every block's largest member is as small as 64-slot blocks allow.  key[f]
// per block.
__global__ __launch_bounds__(256) void k_feature_key(const float* __restrict__ XT, int64_t n,
                                                     double* __restrict__ key) {
  __shared__ double s1[256], s2[256];
  const float* x = XT + (int64_t)blockIdx.x * n;
  double a = 0.0, b = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    const double v = x[j];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) key[blockIdx.x] = s2[0] - s1[0] * s1[0] / (double)n;
}

// perm[rank(f)] = f and iperm[f] = rank(f) (iperm may be null), rank by
// (descending key, index); identity when key is null.  A NaN key (a column
// holding NaN or infinities) ranks as +inf, so the ranks stay a permutation.
__global__ __launch_bounds__(256) void k_feature_rank(const double* __restrict__ key, int d,
                                                      int* __restrict__ perm,
                                                      int* __restrict__ iperm) {
  const double kInf = __longlong_as_double(0x7ff0000000000000LL);
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    int r = f;
    if (key) {
      r = 0;
      const double kf = key[f] != key[f] ? kInf : key[f];
      for (int g = 0; g < d; ++g) {
        const double kg = key[g] != key[g] ? kInf : key[g];
        r += (kg > kf || (kg == kf && g < f)) ? 1 : 0;
      }
    }
    perm[r] = f;
    if (iperm) iperm[f] = r;
  }
}

// frange = {min, max} of the column of the feature the last step chose (the
// forward warm bound's largest square); one block.
__global__ __launch_bounds__(1024) void k_col_range(const float* __restrict__ XT,
                                                    const SelState* __restrict__ st, int64_t n,
                                                    float* __restrict__ frange) {
  __shared__ float s_lo[16], s_hi[16];
  const int g = st->fstar;
  if (st->stopped || g < 0) return;
  const float* __restrict__ col = XT + (int64_t)g * n;
  float lo = __int_as_float(0x7f800000), hi = -lo;
  for (int64_t j = threadIdx.x; j < n; j += 1024) {
    const float v = col[j];
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, o));
    hi = fmaxf(hi, __shfl_xor(hi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) {
      lo = fminf(lo, s_lo[w]);
      hi = fmaxf(hi, s_hi[w]);
    }
    frange[0] = fminf(lo, s_lo[0]);
    frange[1] = fmaxf(hi, s_hi[0]);
  }
}

// Initial state: all features unselected (forward) or selected (backward).
__global__ void k_init_state(SelState* __restrict__ st, const Problem pb, int* __restrict__ cand,
                             int* __restrict__ sel, int* __restrict__ ordering,
                             float* __restrict__ errors, const int* __restrict__ perm) {
  const int t = threadIdx.x;
  const bool fwd = pb.method == 1;
  for (int g = t; g < pb.d; g += blockDim.x) {
    sel[g] = fwd ? 0 : 1;
    ordering[g] = 0;
    cand[g] = perm[g];
  }
  for (int g = t; g < pb.nl * pb.d; g += blockDim.x) errors[g] = 0.0f;
  if (t == 0) {
    st->step = 0;
    st->fstar = -1;
    st->nonfinite = 0;
    st->fslot = 0;
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

hipError_t launch_fill_i32(int* p, int64_t count, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_1d(count, 256, 1024)), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_1d(count, 256, 1024)), dim3(256), 0, s, p, count);
  return hipGetLastError();
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

hipError_t launch_gram_exact(const float* XT, float* D, float* Dlo, float* M, int64_t n, int d,
                             hipStream_t s) {
  // the compensated Gram keeps the whole grid; otherwise the upper triangle
  // of T x T tiles, each mirrored
  const int64_t T = (n + 63) / 64;
  dim3 g((unsigned)T, (unsigned)T), gt((unsigned)(T * (T + 1) / 2));
  dim3 b(256);
  if (Dlo && M) hipLaunchKernelGGL((k_gram_exact<true, true, false>), g, b, 0, s, XT, D, Dlo, M, n, d);
  else if (Dlo) hipLaunchKernelGGL((k_gram_exact<true, false, false>), g, b, 0, s, XT, D, Dlo, M, n, d);
  else if (M) hipLaunchKernelGGL((k_gram_exact<false, true, true>), gt, b, 0, s, XT, D, Dlo, M, n, d);
  else hipLaunchKernelGGL((k_gram_exact<false, false, true>), gt, b, 0, s, XT, D, Dlo, M, n, d);
  return hipGetLastError();
}

hipError_t launch_quantize_m(const float* M, unsigned char* Mq, int* Mexp, int64_t n,
                             hipStream_t s) {
  if (n % 4) return hipErrorInvalidValue;
  const int64_t rows = (n + 3) / 4;
  hipLaunchKernelGGL(k_quantize_m, dim3((unsigned)(rows < 8192 ? rows : 8192)), dim3(256), 0, s, M,
                     Mq, Mexp, n);
  return hipGetLastError();
}

hipError_t launch_max_square(const float* Xc, const SelState* st, float* M, int64_t n, int cpad,
                             hipStream_t s) {
  const int64_t T = (n + 63) / 64;  // upper triangle of T x T tiles
  dim3 g((unsigned)(T * (T + 1) / 2));
  hipLaunchKernelGGL(k_max_square, g, dim3(256), 0, s, Xc, st, M, n, cpad);
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

hipError_t launch_update_seeds(float* D, float* Dlo, const float* XT, int64_t n,
                               SelState* st, int mult, int* seeds, int64_t npf,
                               const RowList& rl, const float* tauw, const float* frange, int d,
                               const int* cand, const int* iperm, const unsigned char* Mq,
                               const int* Mexp, hipStream_t s, int defer) {
  const int64_t rows = (n + 3) / 4;
  const int g = (int)(rows < 8192 ? rows : 8192);
  if (n % 4) {  // rows off 16-byte boundaries: the seeds only (no lists, no deferred writes)
    if (rl.idx || defer) return hipErrorInvalidValue;
    if (Dlo)
      hipLaunchKernelGGL((k_update_seeds_any<true>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st,
                         mult, seeds, npf);
    else
      hipLaunchKernelGGL((k_update_seeds_any<false>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n, st,
                         mult, seeds, npf);
    return hipGetLastError();
  }
  // backward lists with warm bounds need no new seeds (k_update_seeds)
  const bool no_seeds = mult < 0 && rl.idx && tauw && !Dlo;
  if (defer < 0 || defer > 2 || (defer && !no_seeds)) return hipErrorInvalidValue;
  if (Dlo)
    hipLaunchKernelGGL((k_update_seeds<true, 0, true>), dim3(g), dim3(256), 0, s, D, Dlo, XT, n,
                       st, mult, seeds, npf, rl, tauw, frange, d, cand, iperm, Mq, Mexp);
  else if (defer == 1)
    hipLaunchKernelGGL((k_update_seeds<false, 1, false>), dim3(g), dim3(256), 0, s, D, Dlo, XT,
                       n, st, mult, seeds, npf, rl, tauw, frange, d, cand, iperm, Mq, Mexp);
  else if (defer == 2)
    hipLaunchKernelGGL((k_update_seeds<false, 2, false>), dim3(g), dim3(256), 0, s, D, Dlo, XT,
                       n, st, mult, seeds, npf, rl, tauw, frange, d, cand, iperm, Mq, Mexp);
  else if (no_seeds)
    hipLaunchKernelGGL((k_update_seeds<false, 0, false>), dim3(g), dim3(256), 0, s, D, Dlo, XT,
                       n, st, mult, seeds, npf, rl, tauw, frange, d, cand, iperm, Mq, Mexp);
  else
    hipLaunchKernelGGL((k_update_seeds<false, 0, true>), dim3(g), dim3(256), 0, s, D, Dlo, XT,
                       n, st, mult, seeds, npf, rl, tauw, frange, d, cand, iperm, Mq, Mexp);
  return hipGetLastError();
}

hipError_t launch_compact(const float* X, const int* cand, SelState* st, float* Xc,
                          int64_t n, int d, int cpad, hipStream_t s) {
  hipLaunchKernelGGL(k_compact, dim3(grid_1d(n, 1, 16384)), dim3(256), 0, s, X, cand, st,
                     Xc, n, d, cpad);
  return hipGetLastError();
}

int loss_paths(const int* codes, int count) {
  int paths = 0;
  for (int l = 0; l < count; ++l) paths |= lane_loss(codes[l]) ? 2 : 1;
  return paths;
}

hipError_t launch_loss(const float* P, const float* y, const int* losses, const SelState* st,
                       float* err, int64_t n, int cpad, int nk, int nl, int max_cand,
                       float* scratch, int paths, hipStream_t s) {
  if (paths & 1) {
    dim3 g((unsigned)max_cand, (unsigned)nk);
    hipLaunchKernelGGL(k_loss, g, dim3(64), 0, s, P, y, losses, st, err, n, cpad, nk, nl,
                       scratch);
  }
  if (paths & 2) {
    dim3 g((unsigned)((max_cand + 63) / 64), (unsigned)nk);
    hipLaunchKernelGGL(k_loss_lanes, g, dim3(kLossThreads), 0, s, P, y, losses, st, err, n, cpad, nk, nl);
  }
  return hipGetLastError();
}

hipError_t launch_loss_single(const float* y, const float* p, int64_t n, int code, float* out,
                              float* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_single, dim3(1), dim3(64), 0, s, y, p, n, code, out, scratch);
  return hipGetLastError();
}

bool needs_median_scratch(const int* codes, int count) {
  for (int l = 0; l < count; ++l)
    if (codes[l] == kMad || codes[l] == kMadNorm || codes[l] == kRmseNormMadNormStd ||
        codes[l] == kRmseNormMadNormGalaxies)
      return true;
  return false;
}

hipError_t launch_select(SelState* st, const Problem& pb, const int* kvals, const float* err,
                         int* cand, int* sel, int* ordering, float* errors, float* val,
                         int* bestk, float* cand_err, const int* perm, hipStream_t s) {
  hipLaunchKernelGGL(k_select, dim3(1), dim3(256), 0, s, st, pb, kvals, err, cand, sel, ordering,
                     errors, val, bestk, cand_err, perm);
  return hipGetLastError();
}

hipError_t launch_feature_order(const float* XT, int64_t n, int d, double* key, int* perm,
                                int* iperm, bool by_scale, hipStream_t s) {
  if (by_scale) hipLaunchKernelGGL(k_feature_key, dim3(d), dim3(256), 0, s, XT, n, key);
  hipLaunchKernelGGL(k_feature_rank, dim3(1), dim3(256), 0, s, by_scale ? key : nullptr, d,
                     perm, iperm);
  return hipGetLastError();
}

hipError_t launch_col_range(const float* XT, const SelState* st, int64_t n, float* frange,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_col_range, dim3(1), dim3(1024), 0, s, XT, st, n, frange);
  return hipGetLastError();
}

hipError_t launch_init_state(SelState* st, const Problem& pb, int* cand, int* sel, int* ordering,
                             float* errors, const int* perm, hipStream_t s) {
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(256), 0, s, st, pb, cand, sel, ordering, errors,
                     perm);
  return hipGetLastError();
}

}  // namespace snn
