// Per-feature sorted columns for the first step of a forward selection.
//
// With D = 0 the distance of candidate f is d_qj = fl((x_qf - x_jf)^2): a
// one-dimensional kNN per feature.  Each column of XT is sorted once
// (rocPRIM segmented radix sort of (value, point index) pairs, stable, so
// equal values keep ascending index), and the search walks outward from each
// query's position (k_search_sorted in search.hip).  The walk relies on d
// being monotone along the sorted order on each side of the query, which
// holds for finite values; k_check_finite flags problems whose X holds
// NaN or infinity, and those keep the dense first step.
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "common.h"
#include "kernels.h"

namespace snn {

namespace {

struct PositionInColumn {  // value of element i of the flattened [d][n] array: i mod n
  unsigned int n;
  __host__ __device__ int operator()(unsigned int i) const { return (int)(i % n); }
};

struct ColumnStart {  // offset of column f
  unsigned int n;
  __host__ __device__ unsigned int operator()(unsigned int f) const { return f * n; }
};

auto index_iter(unsigned int n) {
  return rocprim::make_transform_iterator(rocprim::make_counting_iterator<unsigned int>(0u),
                                          PositionInColumn{n});
}

auto offset_iter(unsigned int n) {
  return rocprim::make_transform_iterator(rocprim::make_counting_iterator<unsigned int>(0u),
                                          ColumnStart{n});
}

}  // namespace

// st->nonfinite = 1 when some x is NaN or infinite.
__global__ __launch_bounds__(256) void k_check_finite(const float* __restrict__ XT, int64_t count,
                                                      SelState* __restrict__ st) {
  bool bad = false;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    bad |= !isfinite(XT[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&st->nonfinite, 1);
}

size_t sort_columns_temp_bytes(int64_t n, int d) {
  size_t bytes = 0;
  const unsigned int un = (unsigned int)n;
  auto offs = offset_iter(un);
  if (rocprim::segmented_radix_sort_pairs(nullptr, bytes, (const float*)nullptr, (float*)nullptr,
                                          index_iter(un), (int*)nullptr,
                                          (unsigned int)(n * d), (unsigned int)d, offs, offs + 1) !=
      hipSuccess)
    return 0;
  return bytes;
}

hipError_t launch_check_finite(const float* XT, int64_t count, SelState* st, hipStream_t s) {
  const int grid = (int)((count + 255) / 256 < 4096 ? (count + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_check_finite, dim3(grid), dim3(256), 0, s, XT, count, st);
  return hipGetLastError();
}

hipError_t launch_sort_columns(const float* XT, int64_t n, int d, SelState* st, float* skey,
                               int* sidx, void* temp, size_t temp_bytes, hipStream_t s) {
  const int64_t count = n * d;
  hipError_t e = launch_check_finite(XT, count, st, s);
  if (e != hipSuccess) return e;
  const unsigned int un = (unsigned int)n;
  auto offs = offset_iter(un);
  size_t bytes = temp_bytes;
  return rocprim::segmented_radix_sort_pairs(temp, bytes, XT, skey, index_iter(un), sidx,
                                             (unsigned int)count, (unsigned int)d, offs, offs + 1,
                                             0, 32, s);
}

}  // namespace snn
