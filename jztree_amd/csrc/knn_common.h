// Per-lane exact top-k shared by the kNN leaf kernels (knn.hip, knn_fused.hip).
//
// Each lane owns one query and keeps its k best candidates as u64 keys (r2 bits << 32 | id),
// ascending, in registers.  Keys order candidates by (canonical FP32 r2, reported id): the
// reference's SortedNearestK (knn.cuh:38-106) keeps (r2, index) pairs sorted by r2 with its
// insertion order deciding ties; here ties are broken by id so the result is unique.
#pragma once
#include "util.h"

namespace jzt {

__device__ __forceinline__ float key_r2(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }

// Insert `key` into the ascending list best[0..K) (caller guarantees key < best[K-1]; keys are
// unique).  Top-down: slot i takes best[i-1] where key < best[i-1], takes key where
// best[i-1] <= key < best[i], and keeps its value otherwise; each comparison serves two
// neighbouring slots.  Every 8 slots the wave leaves the (fully unrolled, statically indexed)
// network once no active lane still shifts.
template <int K>
__device__ __forceinline__ void topk_insert(uint64_t (&best)[K], uint64_t key) {
  if constexpr (K == 1) {
    best[0] = key;
  } else {
    bool hi = true;                     // key < best[i]
    bool lo_next = key < best[K - 2];   // compare issued one step ahead of its selects
#pragma unroll
    for (int i = K - 1; i >= 1; i--) {
      const bool lo = lo_next;          // key < best[i-1]
      if (i >= 2) lo_next = key < best[i - 2];
      const uint64_t t = hi ? key : best[i];
      best[i] = lo ? best[i - 1] : t;
      hi = lo;
      // materialise the slot before the next exit test: otherwise the selects sink below the
      // exits and the compare masks of several blocks stay live in (spilled) SGPRs
      asm volatile("" : "+v"(best[i]));
      if ((i & 7) == 0 && !wave_any(lo)) return;
    }
    best[0] = hi ? key : best[0];
  }
}

constexpr uint64_t kEmptyKey = (0x7f800000ull << 32) | 0xffffffffull;  // (r2 = +inf, id = -1)

__device__ __forceinline__ int rdlane(int v, int i) { return __builtin_amdgcn_readlane(v, i); }
__device__ __forceinline__ float rdlane(float v, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}
__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// One source particle as staged in LDS: position and reported id in one 16-byte record.
struct alignas(16) SrcRec {
  float p[3];
  uint32_t id;
};

// ---- buffered top-k ------------------------------------------------------------------------
// A candidate that beats a lane's current k-th key is not inserted at once: it is appended to the
// lane's row of an LDS buffer (kBufCap(K) entries).  When some lane's row is full (or the query is
// done) the wave flushes: each lane sorts its row with a bitonic network and merges it into its
// sorted top-k (half-cleaner against the reversed row, then a bitonic merge).  A flush costs
// about as much as three single insertions but absorbs up to kBufCap of them per lane, and the
// network no longer runs whenever any one lane accepts a candidate.  Rows are kBufStride u64 apart
// (odd stride: lanes reading the same slot hit different banks).

#ifndef JZT_TOPK_BUF
#define JZT_TOPK_BUF 16
#endif
template <int K> constexpr int kBufCap = K < JZT_TOPK_BUF ? K : JZT_TOPK_BUF;
constexpr int kBufStride = 17;
constexpr int kBufWords = 64 * kBufStride;  // u64 per wave

// Keys compared as doubles: the high word of a key is the bit pattern of a non-negative float r2
// (at most +inf, 0x7f800000), so read as an IEEE double every key is non-negative, never NaN or
// infinite, and the double order equals the unsigned integer order (denormal doubles included,
// which the FP64 mode keeps).  A compare-exchange is then one v_min_f64 and one v_max_f64
// instead of a 64-bit compare and four 32-bit selects.
__device__ __forceinline__ double key_as_f64(uint64_t k) { return __longlong_as_double((long long)k); }
__device__ __forceinline__ uint64_t f64_as_key(double d) { return (uint64_t)__double_as_longlong(d); }

__device__ __forceinline__ void ce_asc(uint64_t& a, uint64_t& b) {
  const double x = key_as_f64(a), y = key_as_f64(b);
  a = f64_as_key(__builtin_fmin(x, y));
  b = f64_as_key(__builtin_fmax(x, y));
}

// ascending sort of a register array by Batcher's odd-even merge sort (N a power of two; 63
// compare-exchanges for N = 16 against 80 for the bitonic network): for each merge width 2p, merge
// steps of distance k = p, p/2, ..., 1 compare i+j with i+j+k inside one 2p-block.
template <int N>
__device__ __forceinline__ void oddeven_sort_regs(uint64_t (&v)[N]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j <= N - 1 - k; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < N; i++) {
          if (i <= k - 1 && i <= N - j - k - 1 && (i + j) / (2 * p) == (i + j + k) / (2 * p)) ce_asc(v[i + j], v[i + j + k]);
        }
      }
    }
  }
}

// Merge the lane's buffered keys row[0..cnt) into the ascending best[0..K): best becomes the K
// smallest of the union, ascending.  cnt is reset.  The row is read whole (slots past cnt hold
// stale keys, replaced by empty ones), so the reads need no per-slot branch.
template <int K>
__device__ __forceinline__ void topk_flush(uint64_t (&best)[K], const uint64_t* row, int& cnt) {
  constexpr int M = kBufCap<K>;
  uint64_t b[M];
#pragma unroll
  for (int i = 0; i < M; i++) b[i] = row[i];
#pragma unroll
  for (int i = 0; i < M; i++) b[i] = i < cnt ? b[i] : kEmptyKey;
  cnt = 0;
  oddeven_sort_regs<M>(b);
  // half-cleaner of best ++ reverse(b padded to K): the low half holds the K smallest, bitonic
#pragma unroll
  for (int i = 0; i < M; i++) {
    best[K - M + i] = f64_as_key(__builtin_fmin(key_as_f64(best[K - M + i]), key_as_f64(b[M - 1 - i])));
  }
  // bitonic merge to ascending order
#pragma unroll
  for (int stride = K / 2; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < K; i++) {
      const int j = i ^ stride;
      if (j > i) ce_asc(best[i], best[j]);
    }
  }
}

// Stream n staged source particles into the lane's buffer (and the top-k on flushes).  `wrap`
// (wave-uniform) applies the periodic wrap; the caller clears it only when no difference of the
// chunk needs it.  row = this lane's buffer row, cnt = its fill (at most kBufCap on entry and on
// return).  Sources go JZT_CONSUME_UNROLL at a time: their LDS reads are issued together and the
// wave flushes before a group whenever some lane has less room than a whole group.  Returns
// whether best[] changed (a flush happened).
//
// A candidate is accepted when its r2 is <= the r2 of the lane's k-th key: a superset of "key <
// best[K-1]" that only adds equal-r2 candidates with larger ids, which the flush (an exact merge of
// keys) drops again.  The candidate is written to the row's next free slot whether it is accepted
// or not, and the fill advances by the accept flag: no divergent branch per candidate.  A group
// starts with at most CAP - U entries, so every such write lands inside the row.
#ifndef JZT_CONSUME_UNROLL
#define JZT_CONSUME_UNROLL 4
#endif
template <int K, bool PER, bool FLOOR, int D>
__device__ __forceinline__ bool consume_chunk(uint64_t (&best)[K], const Vec<D, float>& q, const SrcRec* s_rec,
                                              int n, bool wrap, float boxsize, float half, uint64_t floor_key,
                                              uint64_t* row, int& cnt) {
  constexpr int U = JZT_CONSUME_UNROLL;
  constexpr int CAP = kBufCap<K>;
  static_assert(U >= 1 && U <= CAP, "unroll must fit a buffer row");
  bool flushed = false;
  float kr = key_r2(best[K - 1]);
  uint32_t* const row32 = reinterpret_cast<uint32_t*>(row);
  for (int j = 0; j < n; j += U) {  // one flush site: the network is inlined once
    if (wave_any(cnt > CAP - U)) {
      topk_flush<K>(best, row, cnt);
      flushed = true;
      kr = key_r2(best[K - 1]);
    }
    SrcRec c[U];
#pragma unroll
    for (int u = 0; u < U; u++) c[u] = s_rec[j + u < n ? j + u : j];
    float r2[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      Vec<D, float> p;
#pragma unroll
      for (int d = 0; d < D; d++) p[d] = c[u].p[d];
      if (PER && wrap) r2[u] = dist2P<true, D, float>(q, p, boxsize, half);
      else r2[u] = dist2P<false, D, float>(q, p, boxsize, half);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool acc = j + u < n && r2[u] <= kr;
      if (FLOOR) acc = acc && (((uint64_t)__float_as_uint(r2[u]) << 32) | c[u].id) > floor_key;
      row32[2 * cnt] = c[u].id;
      row32[2 * cnt + 1] = __float_as_uint(r2[u]);
      cnt += acc ? 1 : 0;
    }
  }
  return flushed;
}

// Write a lane's k results (ascending keys) to its output row: radius = sqrt(r2), id = low word.
// Rows of whole float4/int4 groups that are 16-byte aligned go out as 16-byte stores (a quarter of
// the store instructions; with rows in input order every lane writes a different row, and the
// address processing of 64 scattered rows per instruction dominates), the rest element-wise.
template <int K>
__device__ __forceinline__ void write_row(const uint64_t (&best)[K], float* __restrict__ rad_out,
                                          int32_t* __restrict__ id_out, int64_t o, int nk) {
  if constexpr (K % 4 == 0) {
    if (nk >= K && ((reinterpret_cast<uintptr_t>(rad_out + o) | reinterpret_cast<uintptr_t>(id_out + o)) & 15) == 0) {
#pragma unroll
      for (int i = 0; i < K; i += 4) {
        float4 r;
        int4 d;
        r.x = sqrtf(key_r2(best[i]));
        r.y = sqrtf(key_r2(best[i + 1]));
        r.z = sqrtf(key_r2(best[i + 2]));
        r.w = sqrtf(key_r2(best[i + 3]));
        d.x = (int32_t)(uint32_t)best[i];
        d.y = (int32_t)(uint32_t)best[i + 1];
        d.z = (int32_t)(uint32_t)best[i + 2];
        d.w = (int32_t)(uint32_t)best[i + 3];
        *reinterpret_cast<float4*>(rad_out + o + i) = r;
        *reinterpret_cast<int4*>(id_out + o + i) = d;
      }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < K; i++) {
    if (i < nk) {
      rad_out[o + i] = sqrtf(key_r2(best[i]));
      id_out[o + i] = (int32_t)(uint32_t)best[i];
    }
  }
}

}  // namespace jzt
