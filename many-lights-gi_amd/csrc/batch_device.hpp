// Device helpers of the persistent batched kernels.
#pragma once

#include "kernels.hpp"

namespace mlgi {

// Frame of the batch whose range of one stage's items holds `item`: the last
// frame with first <= item (frames are sorted by that stage's first item).
template <uint32_t BatchFrame::*kFirst>
__device__ __forceinline__ int batch_find(const BatchFrame* __restrict__ f, int n, uint32_t item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (f[mid].*kFirst <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

constexpr int kBatchXcds = 8;

// Workgroups are dispatched to the XCDs round robin. A persistent workgroup
// b (of a grid that is a multiple of 8) serves XCD b % 8, and each XCD walks
// one contiguous eighth of the item space, so neighbouring items (e.g. the
// interleave classes of one screen tile, or one frame's ISMs) share an L2.
struct XcdRange {
  uint32_t begin, end, start, stride;
};
__device__ __forceinline__ XcdRange xcd_range(uint32_t total, uint32_t slot, uint32_t slots_per_xcd) {
  const uint32_t per_xcd = (total + kBatchXcds - 1) / kBatchXcds;
  const uint32_t xcd = blockIdx.x % kBatchXcds;
  const uint32_t begin = xcd * per_xcd;
  const uint32_t end = min(total, begin + per_xcd);
  return XcdRange{begin, end, begin + slot, slots_per_xcd};
}

// Persistent grid size: `per_cu` workgroups per CU on 256 CUs, at most one
// per item, rounded up to a multiple of the XCD count.
inline unsigned persistent_grid(uint32_t items, unsigned per_cu) {
  uint64_t g = 256ull * per_cu;
  if (items < g) g = items;
  if (g == 0) g = 1;
  return (unsigned)((g + kBatchXcds - 1) / kBatchXcds * kBatchXcds);
}

}  // namespace mlgi
