// Host check of speedynn_amd/csrc/qsort_order.h: sorting with the reference's
// comparator (speedynn/lib/loss_functions.c:163-169) through the C library's
// qsort must leave every array, NaN included, exactly as merge_sort_ref and
// its 64-worker split do.  Prints "ok <cases>" or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../speedynn_amd/csrc/qsort_order.h"

static int cmp_ref(const void* a, const void* b) {
  if (*(const float*)a - *(const float*)b < 0) return -1;
  if (*(const float*)a - *(const float*)b > 0) return +1;
  return 0;
}

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next() {
  state += 0x9E3779B97F4A7C15ull;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static float draw(int mix) {
  const uint64_t r = next();
  const int pick = (int)(r % 16);
  const float nan_pos = std::nanf("");
  const float nan_neg = -std::nanf("");
  if (mix & 1) {
    if (pick == 0) return nan_pos;
    if (pick == 1) return nan_neg;
  }
  if (mix & 2) {
    if (pick == 2) return INFINITY;
    if (pick == 3) return -INFINITY;
    if (pick == 4) return 0.0f;
    if (pick == 5) return -0.0f;
  }
  if (mix & 4) return (float)(r % 5);  // many exact ties
  return (float)((double)(r >> 11) / 9007199254740992.0 * 2.0 - 1.0);
}

int main() {
  int cases = 0;
  const int sizes[] = {1, 2, 3, 5, 7, 8, 31, 63, 64, 65, 100, 127, 128, 129, 300, 1000, 4097, 20000};
  for (int n : sizes) {
    for (int mix = 0; mix < 8; ++mix) {
      for (int rep = 0; rep < 3; ++rep) {
        std::vector<float> a(n), b, c, tmp(n);
        for (int i = 0; i < n; ++i) a[i] = draw(mix);
        b = a;
        c = a;
        std::qsort(a.data(), n, sizeof(float), cmp_ref);
        snn::merge_sort_ref(b.data(), tmp.data(), n);
        for (int w = 0; w < 64; ++w) snn::sort_subtree(c.data(), tmp.data(), n, 6, w);
        for (int t = 5; t >= 0; --t)
          for (int w = 0; w < (1 << t); ++w) snn::merge_level_node(c.data(), tmp.data(), n, t, w);
        if (std::memcmp(a.data(), b.data(), sizeof(float) * n) != 0) {
          std::printf("mismatch merge_sort_ref n=%d mix=%d rep=%d\n", n, mix, rep);
          return 1;
        }
        if (std::memcmp(a.data(), c.data(), sizeof(float) * n) != 0) {
          std::printf("mismatch split n=%d mix=%d rep=%d\n", n, mix, rep);
          return 1;
        }
        ++cases;
      }
    }
  }
  std::printf("ok %d\n", cases);
  return 0;
}
