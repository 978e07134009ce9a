// Device helpers of the batched kernels (one launch per stage over all frames
// of a batch).
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

// Loads and stores through global-address-space pointers. Pointers read from
// a frame table (or any struct in memory) are generic, and generic accesses
// compile to flat instructions, which count against both the vector-memory
// and the LDS wait counters: every LDS wait then also waits for outstanding
// global loads. These helpers make such accesses global_load / global_store.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// 16-byte loads: through the native vector type (the HIP vector class would
// be copied through a generic reference, which is flat again).
__device__ __forceinline__ float4 gload4(const float4* p) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f v = *(const __attribute__((address_space(1))) v4f*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gload4(const uint4* p) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  const v4u v = *(const __attribute__((address_space(1))) v4u*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore4(uint4* p, uint4 v) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  v4u w = {v.x, v.y, v.z, v.w};
  *(__attribute__((address_space(1))) v4u*)p = w;
}
// Relaxed, agent-scope atomics on global memory (no return value needed
// except for the add).
__device__ __forceinline__ uint32_t gatomic_add(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add((__attribute__((address_space(1))) uint32_t*)p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gatomic_min(uint32_t* p, uint32_t v) {
  (void)__hip_atomic_fetch_min((__attribute__((address_space(1))) uint32_t*)p, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gatomic_or(uint32_t* p, uint32_t v) {
  (void)__hip_atomic_fetch_or((__attribute__((address_space(1))) uint32_t*)p, v, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void gstore(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
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

// Item of workgroup blockIdx.x in a launch of one workgroup per item over a
// grid of xcd_items_grid(items) workgroups: XCD b % 8 takes one contiguous
// eighth of the items, as xcd_range does for a persistent grid, and the
// hardware dispatches each item's workgroup as others retire. Items at or
// past `items` have no work.
__device__ __forceinline__ uint32_t xcd_item(uint32_t items) {
  const uint32_t per_xcd = (items + kBatchXcds - 1) / kBatchXcds;
  return (blockIdx.x % kBatchXcds) * per_xcd + blockIdx.x / kBatchXcds;
}
inline unsigned xcd_items_grid(uint32_t items) {
  return (unsigned)((items + kBatchXcds - 1) / kBatchXcds * kBatchXcds);
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
