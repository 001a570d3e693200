// The order in which the reference's median sort leaves an array that holds
// NaN.  The reference sorts with qsort and the comparator of
// speedynn/lib/loss_functions.c:163-169 (-1 if a - b < 0, +1 if a - b > 0,
// else 0), which calls NaN equal to everything, so the sorted array depends on
// the sort algorithm.  The reference links glibc 2.35, whose qsort is a
// top-down merge sort: the first half holds n/2 elements, and a merge takes
// the left run's element whenever the comparator gives <= 0, i.e. unless
// a - b > 0.  These functions are that textbook merge sort; without NaN every
// correct sort gives the same values.
//
// Plain C++ usable on the device (kernels.hip) and on the host (the CPU test
// that checks them against the C library's qsort).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SNN_HD __host__ __device__
#else
#define SNN_HD
#endif

namespace snn {

// Range of node idx at depth `depth` of the recursion: the bits of idx, most
// significant first, choose the right half.
SNN_HD inline void sort_node(int64_t n, int depth, int idx, int64_t& lo, int64_t& len) {
  lo = 0;
  len = n;
  for (int t = depth - 1; t >= 0; --t) {
    const int64_t n1 = len / 2;
    if ((idx >> t) & 1) {
      lo += n1;
      len -= n1;
    } else {
      len = n1;
    }
  }
}

// Merge the sorted runs b[0, n1) and b[n1, n1 + n2) in place through tmp.
SNN_HD inline void merge_runs(float* b, float* tmp, int64_t n1, int64_t n2) {
  int64_t i = 0, j = n1, o = 0;
  const int64_t e = n1 + n2;
  while (i < n1 && j < e) {
    const float l = b[i], r = b[j];
    if (!(l - r > 0.0f)) {
      tmp[o++] = l;
      ++i;
    } else {
      tmp[o++] = r;
      ++j;
    }
  }
  while (i < n1) tmp[o++] = b[i++];
  for (int64_t t = 0; t < o; ++t) b[t] = tmp[t];
}

// The whole recursion on b[0, n) (post-order, explicit stack; depth <= 33
// for n < 2^31).
SNN_HD inline void merge_sort_ref(float* b, float* tmp, int64_t n) {
  int64_t flo[40], fn[40];
  int fstate[40];
  int sp = 0;
  flo[0] = 0;
  fn[0] = n;
  fstate[0] = 0;
  while (sp >= 0) {
    const int64_t lo = flo[sp], len = fn[sp];
    if (len <= 1) {
      --sp;
      continue;
    }
    const int64_t n1 = len / 2;
    if (fstate[sp] == 0) {
      fstate[sp] = 1;
      ++sp;
      flo[sp] = lo;
      fn[sp] = n1;
      fstate[sp] = 0;
    } else if (fstate[sp] == 1) {
      fstate[sp] = 2;
      ++sp;
      flo[sp] = lo + n1;
      fn[sp] = len - n1;
      fstate[sp] = 0;
    } else {
      merge_runs(b + lo, tmp + lo, n1, len - n1);
      --sp;
    }
  }
}

// The same sort split over 2^depth workers: worker w sorts subtree w at that
// depth (sort_subtree), then for t = depth-1 ... 0, workers w < 2^t merge the
// two children of node w at depth t (merge_level_node).
SNN_HD inline void sort_subtree(float* b, float* tmp, int64_t n, int depth, int w) {
  int64_t lo, len;
  sort_node(n, depth, w, lo, len);
  merge_sort_ref(b + lo, tmp + lo, len);
}

SNN_HD inline void merge_level_node(float* b, float* tmp, int64_t n, int t, int w) {
  int64_t lo, len;
  sort_node(n, t, w, lo, len);
  merge_runs(b + lo, tmp + lo, len / 2, len - len / 2);
}

}  // namespace snn
