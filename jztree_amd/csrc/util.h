// Device-wide scan / fill / compaction helpers shared by the jztree_amd ops.
// All launches are stream-ordered, allocate nothing and are graph-capturable (fills are kernels,
// never hipMemsetAsync).
#pragma once
#include "common.h"

namespace jzt {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

// number of int32 temp words a scan of n elements needs
inline int64_t scan_tmp_words(int64_t n) { return div_up(n, kScanTile) + 1; }

// out[i] = sum_{j<i} in[j] (exclusive) or sum_{j<=i} (inclusive); in may alias out.
// tmp must hold scan_tmp_words(n) int32.
int scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, bool inclusive, int32_t* tmp);

// process-wide tuning knobs (set through jzt_set_tuning; the values the kernels were written for
// are the defaults)
enum TuneKnob { kTuneLeafWaves = 0, kTuneSortKpt = 1, kTuneXcd = 2, kTuneCount = 3 };
int tune_get(int knob);

// byte offsets of the regions inside one scratch buffer are kept 256-byte aligned
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// fill n 32-bit words with v
int fill_u32(hipStream_t s, void* p, uint32_t v, int64_t n);

// Stable compaction: out[j] = index of the j-th nonzero flag (j < size), remaining out = fill
// (or *fill_dev when given);
// *count_out (device) = number of nonzero flags (may exceed size).  tmp: scan_tmp_words(n) + n words.
int where_flags_i8(hipStream_t s, const int8_t* flags, int64_t n, int32_t* out, int64_t size, int32_t fill,
                   const int32_t* fill_dev, int32_t* count_out, int32_t* tmp);

// wave64 inclusive prefix sum of an int (all 64 lanes active)
__device__ __forceinline__ int wave_incl_scan(int v) {
  int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    v += (lane >= o) ? t : 0;
  }
  return v;
}

// block-wide exclusive scan for blockDim.x == NT (multiple of 64); returns exclusive prefix and
// writes the block total to *total.  `sbuf` needs NT/64 ints of shared memory.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int* sbuf, int* total) {
  constexpr int NW = NT / 64;
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) sbuf[w] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) {
    int x = sbuf[i];
    off += (i < w) ? x : 0;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Grid-stride kernels over the segments of a split array spl[0..n] (non-decreasing, as every
// split array of the tree and its lists is, with padded slots repeating the final value): the
// segments from the first index where spl reaches its final value on are empty, so the loop stops
// there instead of striding through a padded allocation.  Found by a binary search at kernel start
// (the same value in every lane).
__device__ __forceinline__ int64_t active_items(const int32_t* __restrict__ spl, int64_t n) {
  const int32_t last = spl[n];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (spl[mid] < last) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// XCD-aware order of a grid-stride loop over n items.  The chip deals workgroups to its 8 XCDs
// round-robin (workgroup b runs on XCD b % 8) and every XCD has its own L2, so the plain order
// (block b takes items b, b + grid, ...) spreads neighbouring items over all eight L2s.  With `on`
// (and a grid that is a multiple of 8), XCD x takes the x-th eighth of the items and its blocks
// stride through that range: the items an XCD has in flight are neighbours, which in z-order share
// most of their sources.  The item set is the same either way.
struct ItemRange {
  int64_t first, end, step;
};
__device__ __forceinline__ ItemRange xcd_items(int64_t n, bool on) {
  const int64_t g = gridDim.x;
  if (!on || (g & 7) != 0) return {(int64_t)blockIdx.x, n, g};
  const int64_t x = blockIdx.x & 7, per = (n + 7) >> 3;
  return {x * per + (int64_t)(blockIdx.x >> 3), n < (x + 1) * per ? n : (x + 1) * per, g >> 3};
}
// Dynamic form of xcd_items for one-wave blocks: the blocks of an XCD take the items of its eighth
// one at a time from a shared counter (counters[x], zeroed before the launch), so the XCD's blocks
// finish together however uneven the items are; the item set and the XCD-contiguous order are
// those of xcd_items.  It pays where items are expensive and uneven (kNN query leaves: -5% kernel
// time at 2^26); for the cheap, even FoF leaves every fetch order tried was slower than the static
// one.  queue_next is called by all 64 lanes (wave-uniform result).
struct ItemQueue {
  int64_t begin, end;
  uint32_t* ctr;
};
__device__ __forceinline__ ItemQueue xcd_queue(int64_t n, bool on, uint32_t* counters) {
  const int64_t g = gridDim.x;
  if (!on || (g & 7) != 0) return {0, n, counters};
  const int64_t x = blockIdx.x & 7, per = (n + 7) >> 3;
  return {x * per, n < (x + 1) * per ? n : (x + 1) * per, counters + x};
}
__device__ __forceinline__ int64_t queue_next(const ItemQueue& q) {
  uint32_t v = 0;
  if (threadIdx.x == 0) v = atomicAdd(q.ctr, 1u);
  return q.begin + (int64_t)__builtin_amdgcn_readfirstlane(v);
}
// grid for xcd_items: at most kMaxGrid blocks, a multiple of 8
inline unsigned grid_for8(int64_t n) {
  int64_t g = n <= 0 ? 8 : (n < kMaxGrid ? n : kMaxGrid);
  return (unsigned)((g + 7) & ~(int64_t)7);
}

}  // namespace jzt
