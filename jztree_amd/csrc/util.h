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
enum TuneKnob { kTuneLeafWaves = 0, kTuneSortKpt = 1, kTuneCount = 2 };
int tune_get(int knob);

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

}  // namespace jzt
