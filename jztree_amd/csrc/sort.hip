// Z-order sort of positions (replaces jztree's PosZorderSort, sort.cuh:40-76).
//
// The reference sorts the float vectors with a CUB merge sort and the comparator z_pos_less
// (math.cuh:381-402).  Here the order is computed with an LSD radix sort of a 64-bit Morton key:
//
//   1. bounds: max exponent of the finite coordinates, any negative, any inf  (read 12 B/pt)
//   2. keygen: 63-bit interleaved key + input index; fused 8x256 digit histograms (r12 w12 B/pt)
//   3. digit offsets: exclusive scan of the 8 histograms (1 workgroup)
//   4. 8 onesweep passes (8-bit digits): dynamic tile ids, LDS ranking with wave64 ballots,
//      decoupled look-back on per-tile digit counts, LDS-staged coalesced scatter (r12 w12 B/pt)
//   5. fix-up: runs of equal keys are re-ordered with the exact comparator (rare)
//   6. gather: pos_out = pos_in[idx], id_out = idx                           (r4+12 w12+4 B/pt)
//
// Key construction.  For coordinates whose largest finite exponent is Lt, each dim contributes a
// B-bit code (B = 63/dim, i.e. 21 for 3D): the fixed-point bits of |x| at levels Lt-B+1 .. Lt.
// If any coordinate is negative or infinite, the top 2 bits of each code become a digit ranking
// {-inf, negative, non-negative, +inf} (digit-interleaved above all magnitude bits) and the
// magnitude bits of negatives are complemented.  Interleaving these codes (x most significant
// within a level) gives a key whose order is consistent with the exact Morton order: K(a) < K(b)
// implies a precedes b.  Equal keys are resolved by the exact comparator, then by input index.
// NaN rows get the key 2^64-1 and sort last.  The exact comparator is z_pos_less with two quirks
// of the reference removed (exactly equal coordinates get level -inf instead of e-24, and -0.0 is
// treated as +0.0); on every pair where the reference's comparator is a strict weak order the two
// orders agree.
#include <mutex>

#include "util.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr int kSortThreads = 256;  // 4 waves
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kKPTMin = 8;                        // keys per thread: 8, 12 (default) or 16, tuned at run
                                                  // time; look-back state is sized for the smallest tile
constexpr int kRadix = 256;
constexpr int kPasses = 8;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagInc = 2u << 30, kCountMask = (1u << 30) - 1;

// ---------------------------------------------------------------------------------------------
// bit spreading (interleave with x most significant within a level)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t spread3(uint64_t x) {  // 21 bits -> every 3rd bit
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ uint64_t spread2(uint64_t x) {  // 31 bits -> every 2nd bit
  x &= 0x7fffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

template <typename T> struct FBits;
template <> struct FBits<float> {
  using U = uint32_t;
  static constexpr int kMant = 23, kBias = 127, kExpMax = 255;
  static __device__ __forceinline__ U bits(float x) { return __float_as_uint(x); }
};
template <> struct FBits<double> {
  using U = uint64_t;
  static constexpr int kMant = 52, kBias = 1023, kExpMax = 2047;
  static __device__ __forceinline__ U bits(double x) { return (uint64_t)__double_as_longlong(x); }
};

// floor(|x| / 2^(Lt-nb+1)) for finite x with exponent <= Lt: the nb fixed-point bits below Lt
template <typename T> __device__ __forceinline__ uint64_t mag_code(T x, int Lt, int nb) {
  using FB = FBits<T>;
  using U = typename FB::U;
  U u = FB::bits(x) & ~((U)1 << (sizeof(U) * 8 - 1));
  int be = (int)(u >> FB::kMant);
  U mant = u & (((U)1 << FB::kMant) - 1);
  U M = be ? (mant | ((U)1 << FB::kMant)) : mant;
  int E = (be ? be : 1) - FB::kBias;
  int s = Lt - E + FB::kMant + 1 - nb;
  if (s >= (int)(sizeof(U) * 8)) return 0;
  if (s < 0) return (uint64_t)M << (-s);  // more code bits than mantissa bits (2D float32)
  return (uint64_t)(M >> s);
}

// top digit: -inf 0, negative finite 1, non-negative finite (incl. -0.0) 2, +inf 3
template <typename T> __device__ __forceinline__ int top_digit(T x) {
  if (__builtin_isinf(x)) return x < T(0) ? 0 : 3;
  return x < T(0) ? 1 : 2;
}

template <int D, typename T>
__device__ __forceinline__ uint64_t morton_key(const T* x, int Lt, bool general) {
  constexpr int B = 63 / D;
  uint64_t c[D];
  uint64_t top = 0;
  if constexpr (std::is_integral<T>::value) {
    // integers (math.cuh:293-307): the order of the two's-complement bits with the sign flipped;
    // Lt = shift that brings the highest bit any two coordinates differ in to the top of B bits
    using U = typename IntOf<T>::utype;
    constexpr U kSign = (U)1 << (sizeof(U) * 8 - 1);
#pragma unroll
    for (int d = 0; d < D; d++) c[d] = ((uint64_t)((U)x[d] ^ kSign) >> Lt) & ((1ull << B) - 1ull);
  } else {
  bool isnan_row = false;
#pragma unroll
  for (int d = 0; d < D; d++) isnan_row |= __builtin_isnan(x[d]);
  if (isnan_row) return ~0ull;
  if (!general) {
#pragma unroll
    for (int d = 0; d < D; d++) c[d] = mag_code<T>(x[d], Lt, B);
  } else {
    constexpr int nb = B - 2;
#pragma unroll
    for (int d = 0; d < D; d++) {
      int t = top_digit<T>(x[d]);
      uint64_t m = 0;
      if (t == 1) m = ((1ull << nb) - 1) - mag_code<T>(x[d], Lt, nb);
      if (t == 2) m = mag_code<T>(x[d], Lt, nb);
      c[d] = m;
      top = (top << 2) | (uint64_t)t;
    }
  }
  }
  uint64_t key = 0;
  if constexpr (D == 3) {
    key = (spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]);
  } else if constexpr (D == 2) {
    key = (spread2(c[0]) << 1) | spread2(c[1]);
  } else {
    for (int b = B - 1; b >= 0; b--)
      for (int d = 0; d < D; d++) key = (key << 1) | ((c[d] >> b) & 1ull);
  }
  if (general) key |= top << (D * (B - 2));
  return key;
}

// ---------------------------------------------------------------------------------------------
// exact comparator (device copy of the rule described above); -1/0/+1
// ---------------------------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ int exact_level(T a, T b) {
  using FB = FBits<T>;
  using U = typename FB::U;
  int ta = top_digit<T>(a), tb = top_digit<T>(b);
  if (ta != tb) return 1 << 20;
  U ab = FB::bits(a) & ~((U)1 << (sizeof(U) * 8 - 1));
  U bb = FB::bits(b) & ~((U)1 << (sizeof(U) * 8 - 1));
  if (ab == bb) return -(1 << 20);
  int ae = (int)(ab >> FB::kMant) - FB::kBias, be = (int)(bb >> FB::kMant) - FB::kBias;
  if (ae == be) {
    int lz = (sizeof(U) == 4) ? __clz((uint32_t)(ab ^ bb)) : __clzll((uint64_t)(ab ^ bb));
    return max(ae, 1 - FB::kBias) + ((int)(sizeof(U) * 8) - FB::kMant - 1 - lz);
  }
  return max(ae, be);
}

template <int D, typename T> __device__ int exact_cmp(const T* p1, const T* p2) {
  if constexpr (std::is_integral<T>::value) {
    // math.cuh:293-307 and :381-402 for integers: the dimension with the highest differing bit
    // (the first of equal ones) decides, by signed comparison
    using U = typename IntOf<T>::utype;
    int md = -1, m = -1;
    for (int d = 0; d < D; d++) {
      const U x = (U)p1[d] ^ (U)p2[d];
      const int l = x == 0 ? -1 : (int)(sizeof(U) * 8) - 1 - (sizeof(U) == 4 ? __clz((uint32_t)x) : __clzll((uint64_t)x));
      if (l > m) {
        m = l;
        md = d;
      }
    }
    if (md < 0) return 0;
    return p1[md] < p2[md] ? -1 : 1;
  } else {
  bool n1 = false, n2 = false;
  for (int d = 0; d < D; d++) {
    n1 |= __builtin_isnan(p1[d]);
    n2 |= __builtin_isnan(p2[d]);
  }
  if (n1 || n2) return (int)n1 - (int)n2;
  int md = 0, m = exact_level<T>(p1[0], p2[0]);
  for (int d = 1; d < D; d++) {
    int l = exact_level<T>(p1[d], p2[d]);
    if (l > m) {
      m = l;
      md = d;
    }
  }
  if (m == -(1 << 20)) return 0;
  T a = p1[md], b = p2[md];
  if (m == (1 << 20)) return top_digit<T>(a) < top_digit<T>(b) ? -1 : 1;
  return a < b ? -1 : 1;
  }
}

// ---------------------------------------------------------------------------------------------
// 1. bounds
// ---------------------------------------------------------------------------------------------
// meta[0] = max biased exponent of finite nonzero coordinates, meta[1] = any negative, meta[2] = any inf
template <typename T>
__device__ __forceinline__ void bounds_value(T x, uint32_t& emax, uint32_t& neg, uint32_t& inf) {
  using FB = FBits<T>;
  using U = typename FB::U;
  const U u = FB::bits(x) & ~((U)1 << (sizeof(U) * 8 - 1));
  const uint32_t be = (uint32_t)(u >> FB::kMant);
  const bool fin = be != (uint32_t)FB::kExpMax;
  if (fin && u != 0) emax = max(emax, be ? be : 1u);
  neg |= (x < T(0)) ? 1u : 0u;
  inf |= (!fin && (u & (((U)1 << FB::kMant) - 1)) == 0) ? 1u : 0u;
}

// The coordinates are read as 16-byte vectors, four per thread in flight (a single 4-byte load
// per iteration left the kernel waiting on memory latency: 0.36 ms for 2^26 3-D points); an
// unaligned array or the tail past the last whole vector is read value by value.
template <typename T>
__global__ __launch_bounds__(256) void k_bounds(const T* __restrict__ pos, int64_t nval, uint32_t* meta) {
  constexpr int V = 16 / (int)sizeof(T);  // values per vector
  constexpr int UNR = 4;
  using Vec16 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  uint32_t emax = 0, neg = 0, inf = 0;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  const int64_t nvec = (reinterpret_cast<uintptr_t>(pos) & 15) == 0 ? nval / V : 0;
  const Vec16* pv = reinterpret_cast<const Vec16*>(pos);
  int64_t i = tid;
  for (; i + (UNR - 1) * st < nvec; i += UNR * st) {
    Vec16 a[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) a[u] = pv[i + u * st];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const T* e = reinterpret_cast<const T*>(&a[u]);
#pragma unroll
      for (int c = 0; c < V; c++) bounds_value<T>(e[c], emax, neg, inf);
    }
  }
  for (; i < nvec; i += st) {
    const Vec16 a = pv[i];
    const T* e = reinterpret_cast<const T*>(&a);
#pragma unroll
    for (int c = 0; c < V; c++) bounds_value<T>(e[c], emax, neg, inf);
  }
  for (int64_t j = nvec * V + tid; j < nval; j += st) bounds_value<T>(pos[j], emax, neg, inf);
  // wave reductions then one atomic per wave
  for (int o = 32; o > 0; o >>= 1) {
    emax = max(emax, (uint32_t)__shfl_xor((int)emax, o, 64));
    neg |= (uint32_t)__shfl_xor((int)neg, o, 64);
    inf |= (uint32_t)__shfl_xor((int)inf, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&meta[0], emax);
    if (neg) atomicOr(&meta[1], 1u);
    if (inf) atomicOr(&meta[2], 1u);
  }
}

// integers: meta[0] / meta[3] = low / high word of the OR over all coordinates of (x XOR the first
// coordinate), whose top bit is the highest bit in which any two coordinates differ
template <typename T>
__global__ __launch_bounds__(256) void k_bounds_int(const T* __restrict__ pos, int64_t nval, uint32_t* meta) {
  using U = typename IntOf<T>::utype;
  const uint64_t u0 = (uint64_t)(U)pos[0];
  uint64_t acc = 0;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < nval; i += st) acc |= (uint64_t)(U)pos[i] ^ u0;
  uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
  for (int o = 32; o > 0; o >>= 1) {
    lo |= (uint32_t)__shfl_xor((int)lo, o, 64);
    hi |= (uint32_t)__shfl_xor((int)hi, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (lo) atomicOr(&meta[0], lo);
    if (hi) atomicOr(&meta[3], hi);
  }
}

template <typename T> __device__ __forceinline__ int top_level_from_meta(const uint32_t* meta) {
  using FB = FBits<T>;
  uint32_t be = meta[0];
  return (int)(be ? be : 1u) - FB::kBias;
}

// ---------------------------------------------------------------------------------------------
// 2. keygen + digit histograms
// ---------------------------------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_keygen(const T* __restrict__ pos, int64_t n, const uint32_t* __restrict__ meta,
                                                uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[kPasses][kRadix];
  for (int i = threadIdx.x; i < kPasses * kRadix; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  int Lt;
  bool general;
  if constexpr (std::is_integral<T>::value) {
    const uint64_t v = ((uint64_t)meta[3] << 32) | meta[0];
    const int hb = v ? 63 - __clzll(v) : -1;
    Lt = max(0, hb + 1 - 63 / D);
    general = false;
  } else {
    Lt = top_level_from_meta<T>(meta);
    general = (meta[1] | meta[2]) != 0;
  }
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    T x[D];
#pragma unroll
    for (int d = 0; d < D; d++) x[d] = pos[i * D + d];
    uint64_t k = morton_key<D, T>(x, Lt, general);
    keys[i] = k;
    vals[i] = (uint32_t)i;
#pragma unroll
    for (int p = 0; p < kPasses; p++) atomicAdd(&sh[p][(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) {
    uint32_t c = (&sh[0][0])[j];
    if (c) atomicAdd(&hist[j], c);
  }
}

// ---------------------------------------------------------------------------------------------
// 3. digit offsets
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hist_scan(uint32_t* hist) {
  __shared__ int sbuf[4];
  for (int p = 0; p < kPasses; p++) {
    int v = (int)hist[p * kRadix + threadIdx.x];
    int tot;
    int ex = block_excl_scan<256>(v, sbuf, &tot);
    hist[p * kRadix + threadIdx.x] = (uint32_t)ex;
  }
}

// ---------------------------------------------------------------------------------------------
// 4. onesweep pass
// ---------------------------------------------------------------------------------------------
template <int KPT>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const uint64_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           int64_t n, int shift,
                                                           const uint32_t* __restrict__ bucket_off,
                                                           uint32_t* lookback, uint32_t* tile_counter) {
  __shared__ uint32_t s_wcnt[kSortWaves][kRadix];
  __shared__ uint32_t s_lstart[kRadix];
  __shared__ uint32_t s_gbase[kRadix];
  __shared__ uint64_t s_keys[(kSortThreads * KPT)];
  __shared__ int s_scan[kSortThreads / 64];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < kSortWaves * kRadix; i += kSortThreads) (&s_wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t base = (int64_t)tile * (kSortThreads * KPT);
  const int64_t wbase = base + (int64_t)w * 64 * KPT;

  uint64_t key[KPT];
  uint32_t val[KPT];  // loaded with the keys, so the tile waits for global memory once
  uint32_t dig[KPT];
  uint32_t rank[KPT];
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    int64_t idx = wbase + r * 64 + lane;
    key[r] = idx < n ? kin[idx] : ~0ull;
    val[r] = idx < n ? vin[idx] : 0u;
    dig[r] = (uint32_t)(key[r] >> shift) & 255u;
  }
  const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    uint32_t d = dig[r];
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1u;
      uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    uint32_t pre = s_wcnt[w][d];
    uint32_t below = (uint32_t)__popcll(m & lt_mask);
    rank[r] = pre + below;
    if (below == 0) s_wcnt[w][d] = pre + (uint32_t)__popcll(m);
  }
  __syncthreads();

  // per-digit tile totals and per-wave exclusive offsets (thread tid owns digit tid)
  const int d = tid;
  uint32_t total = 0;
#pragma unroll
  for (int ww = 0; ww < kSortWaves; ww++) {
    uint32_t c = s_wcnt[ww][d];
    s_wcnt[ww][d] = total;
    total += c;
  }
  // decoupled look-back across tiles, one digit per thread
  uint32_t* lb = lookback + (int64_t)tile * kRadix;
  uint32_t excl = 0;
  if (tile == 0) {
    st_rel(&lb[d], kFlagInc | total);
  } else {
    st_rel(&lb[d], kFlagAgg | total);
    int64_t j = (int64_t)tile - 1;
    while (true) {
      uint32_t v = ld_acq(&lookback[j * kRadix + d]);
      uint32_t f = v & ~kCountMask;
      if (f == 0) continue;
      excl += v & kCountMask;
      if (f == kFlagInc) break;
      j--;
    }
    st_rel(&lb[d], kFlagInc | (excl + total));
  }
  int tot;
  uint32_t lstart = (uint32_t)block_excl_scan<kSortThreads>((int)total, s_scan, &tot);
  s_lstart[d] = lstart;
  s_gbase[d] = bucket_off[d] + excl - lstart;
  __syncthreads();

  // scatter keys into LDS in tile-local digit order (each key's local slot kept for its value)
  uint32_t slot[KPT];
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    uint32_t dd = dig[r];
    slot[r] = s_lstart[dd] + s_wcnt[w][dd] + rank[r];
    s_keys[slot[r]] = key[r];
  }
  __syncthreads();
  // write out in local order; a slot's digit comes from its key, its global index is kept for
  // the value pass (slot i = tid + r * kSortThreads)
  const int tile_n = (int)((n - base) < (kSortThreads * KPT) ? (n - base) : (kSortThreads * KPT));
  uint32_t gidx[KPT];
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    const int i = tid + r * kSortThreads;
    gidx[r] = 0xffffffffu;
    if (i < tile_n) {
      const uint64_t kk = s_keys[i];
      gidx[r] = s_gbase[(uint32_t)(kk >> shift) & 255u] + (uint32_t)i;
      if (gidx[r] < (uint64_t)n) kout[gidx[r]] = kk;
    }
  }
  __syncthreads();
  uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
  for (int r = 0; r < KPT; r++) s_vals[slot[r]] = val[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    const int i = tid + r * kSortThreads;
    if (i < tile_n && gidx[r] < (uint64_t)n) vout[gidx[r]] = s_vals[i];
  }
}

// ---------------------------------------------------------------------------------------------
// 5. fix-up of equal-key runs (exact comparator, then index)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_run_flags(const uint64_t* __restrict__ keys, int64_t n,
                                                   int8_t* __restrict__ flags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  flags[i] = (i + 1 < n && keys[i + 1] == k && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

template <int D, typename T>
__device__ __forceinline__ bool idx_less(const T* __restrict__ pos, uint32_t a, uint32_t b) {
  int c = exact_cmp<D, T>(pos + (int64_t)a * D, pos + (int64_t)b * D);
  return (c < 0) || (c == 0 && a < b);
}

// One workgroup per run of equal keys (grid-stride over the run list).  The run's end is found by
// a strided parallel search, sortedness by (exact order, index) is checked pairwise, and unsorted
// runs are put in order by a bitonic network (ascending comparators only, so any length works)
// applied in place to the run's indices in global memory.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_fix_runs(const uint64_t* __restrict__ keys, uint32_t* vals, int64_t n,
                                                  const T* __restrict__ pos, const int32_t* __restrict__ runs,
                                                  const int32_t* __restrict__ nruns_p, int64_t max_runs,
                                                  const int32_t* __restrict__ setk) {
  __shared__ unsigned long long s_end;
  const int64_t nruns = min((int64_t)nruns_p[0], max_runs);
  for (int64_t r = blockIdx.x; r < nruns; r += gridDim.x) {
    const int64_t start = runs[r];
    const uint64_t k = keys[start];
    if (threadIdx.x == 0) s_end = (unsigned long long)n;
    __syncthreads();
    const int32_t sk = setk ? setk[start] : 0;
    for (int64_t j = start + 1 + threadIdx.x; j < n; j += blockDim.x) {
      if (keys[j] != k || (setk && setk[j] != sk)) {
        atomicMin(&s_end, (unsigned long long)j);
        break;
      }
    }
    __syncthreads();
    const int64_t end = (int64_t)s_end;
    bool bad = false;
    for (int64_t j = start + threadIdx.x; j + 1 < end; j += blockDim.x)
      bad |= !idx_less<D, T>(pos, vals[j], vals[j + 1]);
    if (__syncthreads_or(bad)) {
      const int64_t len = end - start;
      uint32_t* a = vals + start;
      int64_t np2 = 1;
      while (np2 < len) np2 <<= 1;
      for (int64_t size = 2; size <= np2; size <<= 1) {
        for (int64_t t = threadIdx.x; t < np2 / 2; t += blockDim.x) {
          int64_t blk = t / (size / 2), off = t % (size / 2);
          int64_t i = blk * size + off, j = blk * size + size - 1 - off;
          if (j < len) {
            uint32_t x = a[i], y = a[j];
            if (idx_less<D, T>(pos, y, x)) {
              a[i] = y;
              a[j] = x;
            }
          }
        }
        __syncthreads();
        for (int64_t stride = size / 4; stride > 0; stride >>= 1) {
          for (int64_t t = threadIdx.x; t < np2 / 2; t += blockDim.x) {
            int64_t blk = t / stride, off = t % stride;
            int64_t i = blk * 2 * stride + off, j = i + stride;
            if (j < len) {
              uint32_t x = a[i], y = a[j];
              if (idx_less<D, T>(pos, y, x)) {
                a[i] = y;
                a[j] = x;
              }
            }
          }
          __syncthreads();
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// 6. gather
// ---------------------------------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ pos, const uint32_t* __restrict__ vals, int64_t n,
                                                T* __restrict__ pos_out, typename IntOf<T>::type* __restrict__ id_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v = vals[i];
  if (v >= (uint64_t)n) v = 0;  // unreachable for a valid permutation; keeps a logic error in bounds
#pragma unroll
  for (int d = 0; d < D; d++) pos_out[i * D + d] = pos[(int64_t)v * D + d];
  id_out[i] = (typename IntOf<T>::type)v;
}

// one onesweep pass with the tuned tile (keys per thread); look-back slots are `stride` tiles apart
int onesweep(hipStream_t s, const uint64_t* ki, const uint32_t* vi, uint64_t* ko, uint32_t* vo, int64_t n, int shift,
             const uint32_t* bucket, uint32_t* lookback, uint32_t* counter) {
  const int kpt = tune_get(kTuneSortKpt);
  if (kpt == 8)
    hipLaunchKernelGGL(k_onesweep<8>, dim3((unsigned)div_up(n, kSortThreads * 8)), dim3(kSortThreads), 0, s, ki, vi, ko,
                       vo, n, shift, bucket, lookback, counter);
  else if (kpt == 16)
    hipLaunchKernelGGL(k_onesweep<16>, dim3((unsigned)div_up(n, kSortThreads * 16)), dim3(kSortThreads), 0, s, ki, vi,
                       ko, vo, n, shift, bucket, lookback, counter);
  else
    hipLaunchKernelGGL(k_onesweep<12>, dim3((unsigned)div_up(n, kSortThreads * 12)), dim3(kSortThreads), 0, s, ki, vi,
                       ko, vo, n, shift, bucket, lookback, counter);
  return 0;
}

struct SortLayout {
  size_t keys0, vals0, keys1, vals1, hist, meta, lookback, poscopy, runflags, runtmp, runs, nruns, total;
  int64_t ntiles;
};

SortLayout sort_layout(int64_t n, int dim, int dtype) {
  SortLayout L;
  size_t o = 0;
  L.ntiles = div_up(n > 0 ? n : 1, kSortThreads * kKPTMin);  // look-back stride of every pass
  L.keys0 = o; o = align256(o + 8 * (size_t)n);
  L.vals0 = o; o = align256(o + 4 * (size_t)n);
  L.keys1 = o; o = align256(o + 8 * (size_t)n);
  L.vals1 = o; o = align256(o + 4 * (size_t)n);
  L.hist = o; o += 4 * (size_t)kPasses * kRadix;
  L.meta = o; o += 4 * 16;
  L.lookback = o; o = align256(o + 4 * (size_t)kPasses * L.ntiles * kRadix);
  L.poscopy = o; o = align256(o + (size_t)n * dim * ((dtype == JZT_F64 || dtype == JZT_I64) ? 8 : 4));
  L.runflags = o; o = align256(o + (size_t)n);
  L.runtmp = o; o = align256(o + 4 * (size_t)(scan_tmp_words(n) + n));
  L.runs = o; o = align256(o + 4 * (size_t)(n / 2 + 1));
  L.nruns = o; o = align256(o + 4);
  L.total = o;
  return L;
}

template <int D, typename T>
int zsort_impl(hipStream_t s, const T* pos_in, T* pos_out, typename IntOf<T>::type* id_out, char* tmp, int64_t n) {
  SortLayout L = sort_layout(n, D, sizeof(T) == 8 ? JZT_F64 : JZT_F32);
  uint64_t* k0 = (uint64_t*)(tmp + L.keys0);
  uint64_t* k1 = (uint64_t*)(tmp + L.keys1);
  uint32_t* v0 = (uint32_t*)(tmp + L.vals0);
  uint32_t* v1 = (uint32_t*)(tmp + L.vals1);
  uint32_t* hist = (uint32_t*)(tmp + L.hist);
  uint32_t* meta = (uint32_t*)(tmp + L.meta);
  uint32_t* lookback = (uint32_t*)(tmp + L.lookback);
  // zero histograms, meta (incl. tile counters) and look-back state in one fill
  if (fill_u32(s, hist, 0u, (int64_t)((L.poscopy - L.hist) / 4))) return 1;
  const T* src = pos_in;
  if ((const void*)pos_in == (const void*)pos_out) {
    T* cp = (T*)(tmp + L.poscopy);
    if (hipMemcpyAsync(cp, pos_in, (size_t)n * D * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail("zsort: copy of aliased input failed");
    src = cp;
  }
  int64_t blocks = div_up(n * D, 256);
  if (blocks > 4096) blocks = 4096;
  if constexpr (std::is_integral<T>::value)
    hipLaunchKernelGGL((k_bounds_int<T>), dim3((unsigned)blocks), dim3(256), 0, s, src, n * D, meta);
  else
    hipLaunchKernelGGL((k_bounds<T>), dim3((unsigned)blocks), dim3(256), 0, s, src, n * D, meta);
  blocks = div_up(n, 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL((k_keygen<D, T>), dim3((unsigned)blocks), dim3(256), 0, s, src, n, meta, k0, v0, hist);
  hipLaunchKernelGGL(k_hist_scan, dim3(1), dim3(256), 0, s, hist);
  if (check_launch("zsort keygen")) return 1;
  for (int p = 0; p < kPasses; p++) {
    const uint64_t* ki = (p & 1) ? k1 : k0;
    const uint32_t* vi = (p & 1) ? v1 : v0;
    uint64_t* ko = (p & 1) ? k0 : k1;
    uint32_t* vo = (p & 1) ? v0 : v1;
    onesweep(s, ki, vi, ko, vo, n, 8 * p,
                       hist + p * kRadix, lookback + (int64_t)p * L.ntiles * kRadix, meta + 8 + p);
  }
  if (check_launch("zsort onesweep")) return 1;
  // fix-up of equal-key runs
  int8_t* rflags = (int8_t*)(tmp + L.runflags);
  int32_t* runs = (int32_t*)(tmp + L.runs);
  int32_t* nruns = (int32_t*)(tmp + L.nruns);
  const int64_t max_runs = n / 2 + 1;
  hipLaunchKernelGGL(k_run_flags, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, k0, n, rflags);
  if (where_flags_i8(s, rflags, n, runs, max_runs, 0, nullptr, nruns, (int32_t*)(tmp + L.runtmp))) return 1;
  hipLaunchKernelGGL((k_fix_runs<D, T>), dim3(1024), dim3(256), 0, s, k0, v0, n, src, runs, nruns, max_runs,
                     (const int32_t*)nullptr);
  hipLaunchKernelGGL((k_gather<D, T>), dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, src, v0, n, pos_out, id_out);
  return check_launch("zsort gather");
}

// ---------------------------------------------------------------------------------------------
// 7. segmented sort (many independent point sets, one pipeline)
// ---------------------------------------------------------------------------------------------
// Points of set s occupy [off[s], off[s+1]) of the input; the output keeps every set in its range,
// z-ordered inside it exactly as a single-set sort of that set would order it (the Morton window
// of each set comes from that set's own bounds).  After the 8 Morton passes a stable pass (two if
// there are more than 256 sets) on the set id restores the set ranges without disturbing the
// order inside them; equal-key runs are fixed per set.

constexpr int kSetLds = 4096;  // set offsets staged in LDS up to this many sets

__device__ __forceinline__ int32_t set_of(const int64_t* off, int64_t nsets, int64_t p) {
  int64_t lo = 0, hi = nsets - 1;  // last s with off[s] <= p
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return (int32_t)lo;
}

__global__ __launch_bounds__(256) void k_set_ids(const int64_t* __restrict__ off, int64_t nsets, int64_t n,
                                                 int32_t* __restrict__ setid) {
  __shared__ int64_t s_off[kSetLds];
  const bool lds = nsets <= kSetLds;
  if (lds)
    for (int64_t i = threadIdx.x; i < nsets; i += blockDim.x) s_off[i] = off[i];
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    setid[p] = lds ? set_of(s_off, nsets, p) : set_of(off, nsets, p);
}

// per-set bounds: metas[3*s + {0,1,2}] as meta[] of k_bounds for set s.  Each wave scans one
// contiguous range of points, 64 at a time (coalesced), so it meets only the few sets its range
// touches: lanes accumulate while the whole wave stays in the current set, and the wave reduces
// and flushes once per set (one atomic per set and wave).  A step that straddles a set boundary
// is flushed lane by lane.
__device__ __forceinline__ void bounds_flush(uint32_t* metas, int32_t set, uint32_t emax, uint32_t neg, uint32_t inf) {
  atomicMax(&metas[3 * set], emax);
  if (neg) atomicOr(&metas[3 * set + 1], 1u);
  if (inf) atomicOr(&metas[3 * set + 2], 1u);
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_bounds_seg(const T* __restrict__ pos, int64_t n,
                                                    const int32_t* __restrict__ setid, uint32_t* metas) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int64_t per = ((n + nwaves - 1) / nwaves + 63) & ~(int64_t)63;
  const int64_t r0 = wave * per, r1 = min(n, r0 + per);
  uint32_t emax = 0, neg = 0, inf = 0;  // this lane's share of the current set
  int32_t cur = -1;                     // wave-uniform
  for (int64_t b = r0; b < r1; b += 64) {
    const int64_t i = b + lane;
    const bool valid = i < r1;
    const int32_t sid = valid ? setid[i] : -1;
    const int32_t s0 = __builtin_amdgcn_readfirstlane(sid);
    const bool uniform = __ballot(valid && sid != s0) == 0ull;
    if (!uniform || s0 != cur) {  // leave the current set: flush it once per wave
      if (cur >= 0) {
        for (int o = 32; o > 0; o >>= 1) {
          emax = max(emax, (uint32_t)__shfl_xor((int)emax, o, 64));
          neg |= (uint32_t)__shfl_xor((int)neg, o, 64);
          inf |= (uint32_t)__shfl_xor((int)inf, o, 64);
        }
        if (lane == 0) bounds_flush(metas, cur, emax, neg, inf);
      }
      emax = neg = inf = 0;
      cur = uniform ? s0 : -1;
    }
    if (valid) {
      uint32_t e = 0, ng = 0, nf = 0;
#pragma unroll
      for (int d = 0; d < D; d++) bounds_value<T>(pos[i * D + d], e, ng, nf);
      if (uniform) {
        emax = max(emax, e);
        neg |= ng;
        inf |= nf;
      } else {
        bounds_flush(metas, sid, e, ng, nf);  // a step across a set boundary
      }
    }
  }
  if (cur >= 0) {
    for (int o = 32; o > 0; o >>= 1) {
      emax = max(emax, (uint32_t)__shfl_xor((int)emax, o, 64));
      neg |= (uint32_t)__shfl_xor((int)neg, o, 64);
      inf |= (uint32_t)__shfl_xor((int)inf, o, 64);
    }
    if (lane == 0) bounds_flush(metas, cur, emax, neg, inf);
  }
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_keygen_seg(const T* __restrict__ pos, int64_t n,
                                                    const int32_t* __restrict__ setid,
                                                    const uint32_t* __restrict__ metas, uint64_t* __restrict__ keys,
                                                    uint64_t* __restrict__ keys_orig, uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[kPasses][kRadix];
  for (int i = threadIdx.x; i < kPasses * kRadix; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t* meta = metas + 3 * setid[i];
    const int Lt = top_level_from_meta<T>(meta);
    const bool general = (meta[1] | meta[2]) != 0;
    T x[D];
#pragma unroll
    for (int d = 0; d < D; d++) x[d] = pos[i * D + d];
    const uint64_t k = morton_key<D, T>(x, Lt, general);
    keys[i] = k;
    keys_orig[i] = k;
    vals[i] = (uint32_t)i;
#pragma unroll
    for (int p = 0; p < kPasses; p++) atomicAdd(&sh[p][(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kPasses * kRadix; j += blockDim.x) {
    const uint32_t c = (&sh[0][0])[j];
    if (c) atomicAdd(&hist[j], c);
  }
}

// digit histograms of the set ids (pass 0: bits 0-7, pass 1: bits 8-15), from the set sizes
__global__ void k_set_hist(const int64_t* __restrict__ off, int64_t nsets, uint32_t* __restrict__ hist) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nsets; s += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)(off[s + 1] - off[s]);
    if (c) {
      atomicAdd(&hist[s & 255], c);
      atomicAdd(&hist[kRadix + ((s >> 8) & 255)], c);
    }
  }
}

// after the Morton passes: set key and Morton key of every sorted position
__global__ void k_set_keys(const uint32_t* __restrict__ vals, const int32_t* __restrict__ setid, int64_t n,
                           uint64_t* __restrict__ setkey) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) setkey[i] = (uint64_t)(uint32_t)setid[vals[i] < (uint64_t)n ? vals[i] : 0];
}
__global__ void k_final_keys(const uint32_t* __restrict__ vals, const uint64_t* __restrict__ keys_orig,
                             const int32_t* __restrict__ setid, int64_t n, uint64_t* __restrict__ kf,
                             int32_t* __restrict__ setk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = vals[i] < (uint64_t)n ? vals[i] : 0u;
  kf[i] = keys_orig[v];
  setk[i] = setid[v];
}
__global__ __launch_bounds__(256) void k_run_flags_seg(const uint64_t* __restrict__ keys,
                                                       const int32_t* __restrict__ setk, int64_t n,
                                                       int8_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int32_t sk = setk[i];
  const bool next = i + 1 < n && keys[i + 1] == k && setk[i + 1] == sk;
  const bool prev = i > 0 && keys[i - 1] == k && setk[i - 1] == sk;
  flags[i] = (next && !prev) ? 1 : 0;
}

struct SegLayout {
  SortLayout base;
  size_t setid, setk, korig, metas, total;
};

SegLayout seg_layout(int64_t n, int64_t nsets, int dim, int dtype) {
  SegLayout L;
  L.base = sort_layout(n, dim, dtype);
  size_t o = L.base.total;
  L.setid = o; o = align256(o + 4 * (size_t)n);
  L.setk = o; o = align256(o + 4 * (size_t)n);
  L.korig = o; o = align256(o + 8 * (size_t)n);
  L.metas = o; o = align256(o + 4 * 3 * (size_t)(nsets > 0 ? nsets : 1));
  L.total = o;
  return L;
}

template <int D, typename T>
int zsort_seg_impl(hipStream_t s, const T* pos_in, T* pos_out, typename IntOf<T>::type* id_out, const int64_t* off,
                   int64_t nsets, char* tmp, int64_t n) {
  const SegLayout SL = seg_layout(n, nsets, D, std::is_same<T, double>::value ? JZT_F64 : JZT_F32);
  const SortLayout& L = SL.base;
  uint64_t* k0 = (uint64_t*)(tmp + L.keys0);
  uint64_t* k1 = (uint64_t*)(tmp + L.keys1);
  uint32_t* v0 = (uint32_t*)(tmp + L.vals0);
  uint32_t* v1 = (uint32_t*)(tmp + L.vals1);
  uint32_t* hist = (uint32_t*)(tmp + L.hist);
  uint32_t* meta = (uint32_t*)(tmp + L.meta);
  uint32_t* lookback = (uint32_t*)(tmp + L.lookback);
  int32_t* setid = (int32_t*)(tmp + SL.setid);
  uint64_t* korig = (uint64_t*)(tmp + SL.korig);
  uint32_t* metas = (uint32_t*)(tmp + SL.metas);
  if (fill_u32(s, hist, 0u, (int64_t)((L.poscopy - L.hist) / 4))) return 1;
  if (fill_u32(s, metas, 0u, 3 * nsets)) return 1;
  const T* src = pos_in;
  if ((const void*)pos_in == (const void*)pos_out) {
    T* cp = (T*)(tmp + L.poscopy);
    if (hipMemcpyAsync(cp, pos_in, (size_t)n * D * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail("zsort_segmented: copy of aliased input failed");
    src = cp;
  }
  const unsigned gb = (unsigned)std::min(div_up(n, 256), (int64_t)2048);
  hipLaunchKernelGGL(k_set_ids, dim3(gb), dim3(256), 0, s, off, nsets, n, setid);
  hipLaunchKernelGGL((k_bounds_seg<D, T>), dim3(gb), dim3(256), 0, s, src, n, setid, metas);
  hipLaunchKernelGGL((k_keygen_seg<D, T>), dim3(gb), dim3(256), 0, s, src, n, setid, metas, k0, korig, v0, hist);
  hipLaunchKernelGGL(k_hist_scan, dim3(1), dim3(256), 0, s, hist);
  if (check_launch("zsort_segmented keygen")) return 1;
  for (int p = 0; p < kPasses; p++) {
    const uint64_t* ki = (p & 1) ? k1 : k0;
    const uint32_t* vi = (p & 1) ? v1 : v0;
    uint64_t* ko = (p & 1) ? k0 : k1;
    uint32_t* vo = (p & 1) ? v0 : v1;
    onesweep(s, ki, vi, ko, vo, n, 8 * p,
                       hist + p * kRadix, lookback + (int64_t)p * L.ntiles * kRadix, meta + 8 + p);
  }
  if (check_launch("zsort_segmented onesweep")) return 1;
  // stable set passes: (set key, value) from (k1, v0); histograms, look-back and tile counters of
  // Morton passes 0 and 1 are reused after clearing
  const int npass_set = nsets > 256 ? 2 : 1;
  if (fill_u32(s, hist, 0u, 2 * kRadix)) return 1;
  if (fill_u32(s, lookback, 0u, 2 * L.ntiles * kRadix)) return 1;
  if (fill_u32(s, meta + 8, 0u, 2)) return 1;
  hipLaunchKernelGGL(k_set_hist, dim3((unsigned)std::min(div_up(nsets, 256), (int64_t)1024)), dim3(256), 0, s, off, nsets,
                     hist);
  hipLaunchKernelGGL(k_hist_scan, dim3(1), dim3(256), 0, s, hist);  // scans all 8 rows; rows 0-1 matter
  hipLaunchKernelGGL(k_set_keys, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, v0, setid, n, k1);
  const uint32_t* vfinal = v0;
  {
    onesweep(s, k1, v0, k0, v1, n, 0, hist,
                       lookback, meta + 8);
    vfinal = v1;
    if (npass_set == 2) {
      onesweep(s, k0, v1, k1, v0, n, 8,
                         hist + kRadix, lookback + (int64_t)L.ntiles * kRadix, meta + 9);
      vfinal = v0;
    }
  }
  if (check_launch("zsort_segmented set passes")) return 1;
  uint32_t* vf = const_cast<uint32_t*>(vfinal);
  uint64_t* kf = (vf == v1) ? k1 : k0;  // the key buffer not holding set keys any more
  int32_t* setk = (int32_t*)(tmp + SL.setk);
  hipLaunchKernelGGL(k_final_keys, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, vf, korig, setid, n, kf, setk);
  int8_t* rflags = (int8_t*)(tmp + L.runflags);
  int32_t* runs = (int32_t*)(tmp + L.runs);
  int32_t* nruns = (int32_t*)(tmp + L.nruns);
  const int64_t max_runs = n / 2 + 1;
  hipLaunchKernelGGL(k_run_flags_seg, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, kf, setk, n, rflags);
  if (where_flags_i8(s, rflags, n, runs, max_runs, 0, nullptr, nruns, (int32_t*)(tmp + L.runtmp))) return 1;
  hipLaunchKernelGGL((k_fix_runs<D, T>), dim3(1024), dim3(256), 0, s, kf, vf, n, src, runs, nruns, max_runs, setk);
  hipLaunchKernelGGL((k_gather<D, T>), dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, src, vf, n, pos_out, id_out);
  return check_launch("zsort_segmented gather");
}

// ---------------------------------------------------------------------------------------------
// 8. group order (FoF catalogue): stable sort of the particles by their group label
// ---------------------------------------------------------------------------------------------
// _fof_order (fof.py:510-519): igr[i] = igroup[i] for the filled particles (i < npart), n for the
// rest; isort = stable argsort of igr; counts[i] = number of filled particles labelled isort[i]
// (the group size at each group's first member, whose index is the label; 0 elsewhere; labels
// out of [0, n) are not counted, as the reference's scatter-add drops them).  Labels have
// ceil(bits(n) / 8) radix digits, sorted by that many of the onesweep passes of the z-order sort;
// the label counts come from wave-aggregated atomics (equal labels in consecutive lanes are added
// once).

__global__ __launch_bounds__(256) void k_label_keys(const int32_t* __restrict__ igroup,
                                                    const int32_t* __restrict__ npart_dev, int64_t n, int npass,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ hist, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t sh[4][kRadix];
  for (int i = threadIdx.x; i < 4 * kRadix; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  const int64_t np = npart_dev ? min(max((int64_t)npart_dev[0], (int64_t)0), n) : n;
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // whole waves step together (the aggregation below needs every lane)
  const int64_t nw = (n + 63) & ~(int64_t)63;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i - lane < nw; i += stride) {
    const bool in = i < n;
    const int64_t lab64 = (in && i < np) ? (int64_t)igroup[i] : n;
    const bool counted = in && i < np && lab64 >= 0 && lab64 < n;
    const uint32_t key = in ? (uint32_t)(lab64 >= 0 && lab64 <= n ? lab64 : n) : 0u;
    if (in) {
      keys[i] = key;
      vals[i] = (uint32_t)i;
      for (int p = 0; p < npass; p++) atomicAdd(&sh[p][(key >> (8 * p)) & 255u], 1u);
    }
    // runs of equal counted labels in consecutive lanes: the first lane of a run adds its length
    const int prev = __shfl_up((int)key, 1, 64);
    const bool prev_counted = __shfl_up((int)counted, 1, 64) != 0;
    const bool head = counted && !(lane > 0 && prev_counted && (uint32_t)prev == key);
    const uint64_t heads = __ballot(head), cmask = __ballot(counted);
    if (head) {
      const uint64_t above = lane == 63 ? 0ull : ~((2ull << lane) - 1ull);
      const uint64_t stop = (heads | ~cmask) & above;
      const int end = stop ? __builtin_ctzll(stop) : 64;
      atomicAdd(&cnt[key], (uint32_t)(end - lane));
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < npass * kRadix; j += blockDim.x) {
    const uint32_t c = (&sh[0][0])[j];
    if (c) atomicAdd(&hist[j], c);
  }
}

__global__ void k_label_counts(const uint32_t* __restrict__ isort, const uint32_t* __restrict__ cnt, int64_t n,
                               int32_t* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counts[i] = (int32_t)cnt[isort[i] < (uint64_t)n ? isort[i] : 0];
}

struct OrderLayout {
  size_t keys0, keys1, vals0, vals1, cnt, hist, meta, lookback, total;
  int64_t ntiles;
};

OrderLayout order_layout(int64_t n) {
  OrderLayout L;
  size_t o = 0;
  L.ntiles = div_up(n > 0 ? n : 1, kSortThreads * kKPTMin);
  L.keys0 = o; o = align256(o + 8 * (size_t)n);
  L.keys1 = o; o = align256(o + 8 * (size_t)n);
  L.vals0 = o; o = align256(o + 4 * (size_t)n);
  L.vals1 = o; o = align256(o + 4 * (size_t)n);
  // cnt, hist, meta and the look-back state are zeroed together
  L.cnt = o; o += 4 * (size_t)(n + 1);
  L.hist = o; o += 4 * (size_t)kPasses * kRadix;
  L.meta = o; o += 4 * 16;
  L.lookback = o; o = align256(o + 4 * (size_t)4 * L.ntiles * kRadix);
  L.total = o;
  return L;
}

int fof_order_impl(hipStream_t s, const int32_t* igroup, const int32_t* npart, int64_t n, int32_t* isort,
                   int32_t* counts, char* tmp) {
  const OrderLayout L = order_layout(n);
  uint64_t* k0 = (uint64_t*)(tmp + L.keys0);
  uint64_t* k1 = (uint64_t*)(tmp + L.keys1);
  uint32_t* v0 = (uint32_t*)(tmp + L.vals0);
  uint32_t* v1 = (uint32_t*)(tmp + L.vals1);
  uint32_t* cnt = (uint32_t*)(tmp + L.cnt);
  uint32_t* hist = (uint32_t*)(tmp + L.hist);
  uint32_t* meta = (uint32_t*)(tmp + L.meta);
  uint32_t* lookback = (uint32_t*)(tmp + L.lookback);
  if (fill_u32(s, cnt, 0u, (int64_t)((L.total - L.cnt) / 4))) return 1;
  int bits = 0;
  while (bits < 32 && ((uint64_t)n >> bits) != 0) bits++;  // keys are <= n
  const int npass = std::max(1, (bits + 7) / 8);
  const unsigned gb = (unsigned)std::min(div_up(n, 256), (int64_t)2048);
  hipLaunchKernelGGL(k_label_keys, dim3(gb), dim3(256), 0, s, igroup, npart, n, npass, k0, v0, hist, cnt);
  hipLaunchKernelGGL(k_hist_scan, dim3(1), dim3(256), 0, s, hist);
  if (check_launch("FofOrder keys")) return 1;
  for (int p = 0; p < npass; p++) {
    const uint64_t* ki = (p & 1) ? k1 : k0;
    const uint32_t* vi = (p & 1) ? v1 : v0;
    uint64_t* ko = (p & 1) ? k0 : k1;
    // the last pass writes the permutation straight into isort
    uint32_t* vo = (p == npass - 1) ? (uint32_t*)isort : ((p & 1) ? v0 : v1);
    onesweep(s, ki, vi, ko, vo, n, 8 * p, hist + p * kRadix, lookback + (int64_t)p * L.ntiles * kRadix, meta + 8 + p);
  }
  hipLaunchKernelGGL(k_label_counts, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, (const uint32_t*)isort, cnt, n,
                     counts);
  return check_launch("FofOrder");
}

// ---------------------------------------------------------------------------------------------
// 9. SearchSortedZ (sort.cuh:82-132): insertion indices into a z-sorted array
// ---------------------------------------------------------------------------------------------
// For each query x: the index i with have[i-1] < x <= have[i] in the exact z-order of this sort
// (binary search with the exact comparator), i.e. the left insertion point; n_have past the end.
// leaf_search: `have` holds one point per leaf (e.g. the leaf centres); the index of the leaf x
// belongs to, the one of its two z-order neighbours with the smaller level difference (the lower
// one on ties), as the reference decides it.
template <int D, typename T>
__global__ void k_search_sorted_z(const T* __restrict__ have, const T* __restrict__ query, int64_t n_have,
                                  int64_t n_query, bool leaf_search, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_query) return;
  T x[D];
#pragma unroll
  for (int d = 0; d < D; d++) x[d] = query[i * D + d];
  int64_t lo = 0, hi = n_have;  // first index with !(have[idx] < x)
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (exact_cmp<D, T>(have + mid * D, x) < 0) lo = mid + 1;
    else hi = mid;
  }
  int64_t r = lo;
  if constexpr (!std::is_integral<T>::value) if (leaf_search) {
    // leaves lo-1 and lo are the z-order neighbours of x
    const int64_t a = lo > 0 ? lo - 1 : 0, b = lo < n_have ? lo : n_have - 1;
    const int lv1 = msb_diff_level<D, T>(load_vec<D, T>(have, a), load_vec<D, T>(x, 0));
    const int lv2 = b > a ? msb_diff_level<D, T>(load_vec<D, T>(have, b), load_vec<D, T>(x, 0)) : 1 << 20;
    r = lv1 <= lv2 ? a : b;
  }
  out[i] = (int32_t)r;
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_search_sorted_z(jzt_stream_t stream, const void* posz_have, const void* posz_query, int64_t n_have,
                                   int64_t n_query, int dim, int dtype, int leaf_search, int32_t* indices) {
  hipStream_t s = (hipStream_t)stream;
  if (n_query <= 0) return 0;
  if (n_have <= 0) return fail("jzt_search_sorted_z: empty array to search");
  if (n_have >= (int64_t)INT32_MAX) return fail("jzt_search_sorted_z: n_have must be < 2^31");
  const unsigned g = (unsigned)div_up(n_query, 256);
  const bool leaf = leaf_search != 0;
#define JZT_SSZ(D, T)                                                                                            \
  hipLaunchKernelGGL((k_search_sorted_z<D, T>), dim3(g), dim3(256), 0, s, (const T*)posz_have, (const T*)posz_query, \
                     n_have, n_query, leaf, indices)
  if ((dtype == JZT_I32 || dtype == JZT_I64) && leaf) return fail("jzt_search_sorted_z: leaf search needs float positions");
  switch ((dtype == JZT_F64 ? 100 : dtype == JZT_F32 ? 0 : dtype == JZT_I32 ? 200 : dtype == JZT_I64 ? 300 : 1000) + dim) {
    case 2: JZT_SSZ(2, float); break;
    case 3: JZT_SSZ(3, float); break;
    case 4: JZT_SSZ(4, float); break;
    case 5: JZT_SSZ(5, float); break;
    case 6: JZT_SSZ(6, float); break;
    case 7: JZT_SSZ(7, float); break;
    case 8: JZT_SSZ(8, float); break;
    case 9: JZT_SSZ(9, float); break;
    case 10: JZT_SSZ(10, float); break;
    case 102: JZT_SSZ(2, double); break;
    case 103: JZT_SSZ(3, double); break;
    case 104: JZT_SSZ(4, double); break;
    case 105: JZT_SSZ(5, double); break;
    case 106: JZT_SSZ(6, double); break;
    case 107: JZT_SSZ(7, double); break;
    case 108: JZT_SSZ(8, double); break;
    case 109: JZT_SSZ(9, double); break;
    case 110: JZT_SSZ(10, double); break;
    case 202: JZT_SSZ(2, int32_t); break;
    case 203: JZT_SSZ(3, int32_t); break;
    case 204: JZT_SSZ(4, int32_t); break;
    case 205: JZT_SSZ(5, int32_t); break;
    case 206: JZT_SSZ(6, int32_t); break;
    case 207: JZT_SSZ(7, int32_t); break;
    case 208: JZT_SSZ(8, int32_t); break;
    case 209: JZT_SSZ(9, int32_t); break;
    case 210: JZT_SSZ(10, int32_t); break;
    case 302: JZT_SSZ(2, int64_t); break;
    case 303: JZT_SSZ(3, int64_t); break;
    case 304: JZT_SSZ(4, int64_t); break;
    case 305: JZT_SSZ(5, int64_t); break;
    case 306: JZT_SSZ(6, int64_t); break;
    case 307: JZT_SSZ(7, int64_t); break;
    case 308: JZT_SSZ(8, int64_t); break;
    case 309: JZT_SSZ(9, int64_t); break;
    case 310: JZT_SSZ(10, int64_t); break;
    default: return fail("jzt_search_sorted_z: unsupported dim/dtype");
  }
#undef JZT_SSZ
  return check_launch("SearchSortedZ");
}

extern "C" size_t jzt_fof_order_tmp_bytes(int64_t n) { return order_layout(n).total; }

extern "C" int jzt_fof_order(jzt_stream_t stream, const int32_t* igroup, const int32_t* npart, int64_t n,
                             int32_t* isort, int32_t* counts, void* tmp, size_t tmp_bytes) {
  if (n < 0) return fail("jzt_fof_order: negative n");
  if (n == 0) return 0;
  if (n >= (int64_t)kCountMask) return fail("jzt_fof_order: n must be < 2^30");
  if (tmp_bytes < order_layout(n).total) return fail("jzt_fof_order: tmp buffer too small");
  return fof_order_impl((hipStream_t)stream, igroup, npart, n, isort, counts, (char*)tmp);
}

extern "C" size_t jzt_zsort_tmp_bytes(int64_t n, int dim, int dtype) { return sort_layout(n, dim, dtype).total; }

extern "C" int jzt_pos_zorder_sort(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out, void* tmp,
                                   size_t tmp_bytes, int64_t n, int dim, int dtype) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0) return fail("jzt_pos_zorder_sort: negative n");
  if (n == 0) return 0;
  if (n >= (int64_t)kCountMask) return fail("jzt_pos_zorder_sort: n must be < 2^30");
  size_t need = sort_layout(n, dim, dtype).total;
  if (tmp_bytes < need)
    return fail("jzt_pos_zorder_sort: tmp buffer too small (have " + std::to_string(tmp_bytes) + ", need " +
                std::to_string(need) + ")");
  char* t = (char*)tmp;
  // dims 2-10, as the reference instantiates PosZorderSort (_generate_ffi.py:9-21)
#define JZT_ZS(D)                                                                                                \
  case D:                                                                                                        \
    switch (dtype) {                                                                                             \
      case JZT_F32: return zsort_impl<D, float>(s, (const float*)pos_in, (float*)pos_out, (int32_t*)id_out, t, n); \
      case JZT_F64:                                                                                              \
        return zsort_impl<D, double>(s, (const double*)pos_in, (double*)pos_out, (int64_t*)id_out, t, n);        \
      case JZT_I32:                                                                                              \
        return zsort_impl<D, int32_t>(s, (const int32_t*)pos_in, (int32_t*)pos_out, (int32_t*)id_out, t, n);     \
      default:                                                                                                   \
        return zsort_impl<D, int64_t>(s, (const int64_t*)pos_in, (int64_t*)pos_out, (int64_t*)id_out, t, n);     \
    }
  if (dtype == JZT_F32 || dtype == JZT_F64 || dtype == JZT_I32 || dtype == JZT_I64) {
    switch (dim) {
      JZT_ZS(2) JZT_ZS(3) JZT_ZS(4) JZT_ZS(5) JZT_ZS(6) JZT_ZS(7) JZT_ZS(8) JZT_ZS(9) JZT_ZS(10)
      default: break;
    }
  }
#undef JZT_ZS
  return fail("jzt_pos_zorder_sort: unsupported dim/dtype (dim " + std::to_string(dim) + ")");
}

// ---------------------------------------------------------------------------------------------
// PosZorderSort without a caller workspace: the sort's scratch lives in a fixed device array.
// ---------------------------------------------------------------------------------------------
// Each device that the library runs on holds its own copy of `scratch` until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only a stand-in for the
// symbol); it uses the device address that hipGetSymbolAddress returns for the current device,
// looked up once per device on its first call and kept.  Calls that overlap on one device share
// the array and race; every region the sort needs zeroed is cleared by the sort itself.
namespace jzt {
namespace {
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;
}  // namespace
}  // namespace jzt

static __device__ __align__(256) char scratch[jzt::kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[jzt::kMaxDevices] = {};

// device address of this device's scratch array (nullptr after a failure, reported with fail())
void* scratch_for_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= jzt::kMaxDevices) {
    jzt::fail("jzt_pos_zorder_sort_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      jzt::fail(std::string("jzt_pos_zorder_sort_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" int jzt_pos_zorder_sort_noalloc(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out,
                                           int64_t n, int dim, int dtype) {
  if (n < 0) return fail("jzt_pos_zorder_sort_noalloc: negative n");
  if (n == 0) return 0;
  if (dim < 2 || dim > 10 || (dtype != JZT_F32 && dtype != JZT_F64))
    return fail("jzt_pos_zorder_sort_noalloc: unsupported dim/dtype (dim " + std::to_string(dim) + ")");
  if (n >= (int64_t)kCountMask) return fail("jzt_pos_zorder_sort_noalloc: n must be < 2^30");
  const size_t need = sort_layout(n, dim, dtype).total;
  if (need > kNoAllocScratchBytes)
    return fail("jzt_pos_zorder_sort_noalloc: " + std::to_string(n) + " points need " + std::to_string(need) +
                " bytes of scratch, more than the fixed " + std::to_string(kNoAllocScratchBytes) +
                " bytes (use jzt_pos_zorder_sort with a caller workspace)");
  void* tmp = scratch_for_current_device();
  if (!tmp) return 1;
  return jzt_pos_zorder_sort(stream, pos_in, pos_out, id_out, tmp, need, n, dim, dtype);
}

extern "C" size_t jzt_zsort_segmented_tmp_bytes(int64_t n, int64_t nsets, int dim, int dtype) {
  return seg_layout(n, nsets, dim, dtype).total;
}

extern "C" int jzt_pos_zorder_sort_segmented(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out,
                                             const int64_t* set_offsets, int64_t nsets, void* tmp, size_t tmp_bytes,
                                             int64_t n, int dim, int dtype) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || nsets < 1) return fail("jzt_pos_zorder_sort_segmented: need n >= 0 and nsets >= 1");
  if (nsets > 65536) return fail("jzt_pos_zorder_sort_segmented: at most 65536 sets");
  if (n == 0) return 0;
  if (n >= (int64_t)kCountMask) return fail("jzt_pos_zorder_sort_segmented: n must be < 2^30");
  const size_t need = seg_layout(n, nsets, dim, dtype).total;
  if (tmp_bytes < need)
    return fail("jzt_pos_zorder_sort_segmented: tmp buffer too small (have " + std::to_string(tmp_bytes) +
                ", need " + std::to_string(need) + ")");
  char* t = (char*)tmp;
  if (dtype == JZT_F32) {
    if (dim == 3)
      return zsort_seg_impl<3, float>(s, (const float*)pos_in, (float*)pos_out, (int32_t*)id_out, set_offsets, nsets, t, n);
    if (dim == 2)
      return zsort_seg_impl<2, float>(s, (const float*)pos_in, (float*)pos_out, (int32_t*)id_out, set_offsets, nsets, t, n);
  } else if (dtype == JZT_F64) {
    if (dim == 3)
      return zsort_seg_impl<3, double>(s, (const double*)pos_in, (double*)pos_out, (int64_t*)id_out, set_offsets, nsets,
                                       t, n);
    if (dim == 2)
      return zsort_seg_impl<2, double>(s, (const double*)pos_in, (double*)pos_out, (int64_t*)id_out, set_offsets, nsets,
                                       t, n);
  }
  return fail("jzt_pos_zorder_sort_segmented: unsupported dim/dtype");
}
