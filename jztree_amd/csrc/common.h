// Shared device helpers for the jztree_amd HIP kernels (gfx950, wave64).
//
// Z-order level math restates src/jztree_cuda/common/math.cuh of jztree (msb_xor / msb_diff_level
// :241-327, lvl_vec :524-541, round_*_pow2_cent :404-465, wrap_dx :471-488, distance_squared
// :511-522, mindist2/maxdist2 :628-661).  All of it is integer/bit exact; the only floating-point
// arithmetic is the canonical distance (an FMA chain, matching nvcc's default contraction of the
// reference's `res += dx*dx`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <type_traits>

namespace jzt {

constexpr int kWave = 64;

// ------------------------------------------------------------------------------------------------
// error reporting (host)
// ------------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(const std::string& msg);
int check_launch(const char* what);

template <typename T> struct IntOf;
template <> struct IntOf<float> { using type = int32_t; using utype = uint32_t; };
template <> struct IntOf<double> { using type = int64_t; using utype = uint64_t; };
template <> struct IntOf<int32_t> { using type = int32_t; using utype = uint32_t; };
template <> struct IntOf<int64_t> { using type = int64_t; using utype = uint64_t; };

// ------------------------------------------------------------------------------------------------
// small vector type, AoS layout (dim, T) exactly like the reference's Vec<dim,tvec>
// ------------------------------------------------------------------------------------------------
template <int D, typename T> struct Vec {
  T v[D];
  __device__ __forceinline__ T& operator[](int i) { return v[i]; }
  __device__ __forceinline__ const T& operator[](int i) const { return v[i]; }
};

template <int D, typename T>
__device__ __forceinline__ Vec<D, T> load_vec(const T* __restrict__ p, int64_t i) {
  Vec<D, T> r;
#pragma unroll
  for (int d = 0; d < D; d++) r[d] = p[i * D + d];
  return r;
}
template <int D, typename T>
__device__ __forceinline__ void store_vec(T* __restrict__ p, int64_t i, const Vec<D, T>& x) {
#pragma unroll
  for (int d = 0; d < D; d++) p[i * D + d] = x[d];
}

// ------------------------------------------------------------------------------------------------
// msb_xor: level (power of two) of the highest differing bit of |a| and |b| (math.cuh:241-311)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int msb_xor(float a, float b) {
  if (__builtin_isnan(a) || __builtin_isnan(b) || (__builtin_signbit(a) != __builtin_signbit(b))) return 128;
  uint32_t ab = __float_as_uint(a) & 0x7fffffffu, bb = __float_as_uint(b) & 0x7fffffffu;
  int ae = (int)(ab >> 23) - 127, be = (int)(bb >> 23) - 127;
  if (ae == be) return max(ae, -126) + (8 - __clz(ab ^ bb));
  return max(ae, be);
}
__device__ __forceinline__ int msb_xor(double a, double b) {
  if (__builtin_signbit(a) != __builtin_signbit(b)) return 1024;
  uint64_t ab = (uint64_t)__double_as_longlong(a) & 0x7fffffffffffffffull;
  uint64_t bb = (uint64_t)__double_as_longlong(b) & 0x7fffffffffffffffull;
  int ae = (int)(ab >> 52) - 1023, be = (int)(bb >> 52) - 1023;
  if (ae == be) return max(ae, -1022) + (11 - __clzll(ab ^ bb));
  return max(ae, be);
}

template <int D, typename T>
__device__ __forceinline__ int msb_diff_level(const Vec<D, T>& a, const Vec<D, T>& b) {
  int mdim = 0, m = msb_xor(a[0], b[0]);
#pragma unroll
  for (int i = 1; i < D; i++) {
    int n = msb_xor(a[i], b[i]);
    mdim = n > m ? i : mdim;
    m = n > m ? n : m;
  }
  return (m + 1) * D - mdim;
}

template <typename T> __host__ __device__ constexpr int max_level(int dim) {
  return std::is_same<T, float>::value ? dim * 129 : dim * 1025;
}
// level given to the boundary between two points of different sets in a batch: above every Morton
// level, so no cell, leaf or node ever spans two sets
constexpr int kLvlSet = 1 << 16;

template <typename T> __host__ __device__ constexpr int min_level(int dim) {
  return std::is_same<T, float>::value ? dim * -150 : dim * -1075;
}

// python-style floor division of a binary level into per-dim levels (math.cuh:524-541)
template <int D> __device__ __forceinline__ void lvl_vec(int level, int* l) {
  int o = (level + 2000 * D) / D - 2000;
  int r = level - o * D;
#pragma unroll
  for (int i = 0; i < D; i++) l[i] = o + (r >= (D - i) ? 1 : 0);
}

// exact power of two 2^e (as ldexp(1, e)); e within the normal/denormal range or flushed to 0/inf
__device__ __forceinline__ float pow2f(int e) { return ldexpf(1.0f, e); }
__device__ __forceinline__ double pow2d(int e) { return ldexp(1.0, e); }
template <typename T> __device__ __forceinline__ T pow2(T v, int e);
template <> __device__ __forceinline__ float pow2<float>(float v, int e) { return ldexpf(v, e); }
template <> __device__ __forceinline__ double pow2<double>(double v, int e) { return ldexp(v, e); }

// centre of the cell of size 2^level containing x (math.cuh:404-465)
__device__ __forceinline__ float round_pow2_cent(float x, int level) {
  int32_t xb = (int32_t)(__float_as_uint(x) & 0x7fffffffu);
  int32_t xe = (xb >> 23) - 127;
  if (level >= 128) return 0.f;
  if (level > xe) return ldexpf(__builtin_signbit(x) ? -0.5f : 0.5f, level);
  int32_t keep = max(xe, -126) - level;
  if (keep >= 23) return x;
  uint32_t nb = ((uint32_t)xb & (0xFFFFFFFFu << (23 - keep))) | (1u << (22 - keep));
  float r = __uint_as_float(nb);
  return __builtin_signbit(x) ? -r : r;
}
__device__ __forceinline__ double round_pow2_cent(double x, int level) {
  int64_t xb = (int64_t)((uint64_t)__double_as_longlong(x) & 0x7fffffffffffffffull);
  int32_t xe = (int32_t)(xb >> 52) - 1023;
  if (level >= 1024) return 0.0;
  if (level > xe) return ldexp(__builtin_signbit(x) ? -0.5 : 0.5, level);
  int32_t keep = max(xe, -1022) - level;
  if (keep >= 52) return x;
  uint64_t nb = ((uint64_t)xb & (0xFFFFFFFFFFFFFFFFull << (52 - keep))) | (1ull << (51 - keep));
  double r = __longlong_as_double((long long)nb);
  return __builtin_signbit(x) ? -r : r;
}

// ------------------------------------------------------------------------------------------------
// canonical distances
// ------------------------------------------------------------------------------------------------
// single-rounding fused multiply-add in the operand precision
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
// wrap into [-L/2, L/2) by at most one add/subtract (math.cuh:471-488)
template <typename T> __device__ __forceinline__ T wrap_dx(T dx, T boxsize) {
  if (boxsize > T(0)) {
    T bh = (T)(float)(T(0.5) * boxsize);  // the reference keeps the half box in float
    dx = dx < -bh ? dx + boxsize : dx;
    dx = dx >= bh ? dx - boxsize : dx;
  }
  return dx;
}

// r2 = fma(dz,dz, fma(dy,dy, dx*dx)) -- the canonical FP32 r^2 of this library
template <int D, typename T>
__device__ __forceinline__ T dist2(const Vec<D, T>& a, const Vec<D, T>& b, T boxsize) {
  T d = wrap_dx<T>(a[0] - b[0], boxsize);
  T r = d * d;
#pragma unroll
  for (int i = 1; i < D; i++) {
    T e = wrap_dx<T>(a[i] - b[i], boxsize);
    r = fmaT(e, e, r);
  }
  return r;
}
// same value; PERIODIC chosen at compile time (boxsize > 0 iff PERIODIC)
template <bool PERIODIC, typename T> __device__ __forceinline__ T wrapP(T dx, T boxsize, T half) {
  if constexpr (PERIODIC) {
    dx = dx < -half ? dx + boxsize : dx;
    dx = dx >= half ? dx - boxsize : dx;
  }
  return dx;
}
template <bool PERIODIC, int D, typename T>
__device__ __forceinline__ T dist2P(const Vec<D, T>& a, const Vec<D, T>& b, T boxsize, T half) {
  T d = wrapP<PERIODIC, T>(a[0] - b[0], boxsize, half);
  T r = d * d;
#pragma unroll
  for (int i = 1; i < D; i++) {
    T e = wrapP<PERIODIC, T>(a[i] - b[i], boxsize, half);
    r = fmaT(e, e, r);
  }
  return r;
}

template <int D, typename T>
__device__ __forceinline__ T mindist2(const Vec<D, T>& x1, const Vec<D, T>& x2, const Vec<D, T>& wh, T boxsize) {
  T r2 = T(0);
#pragma unroll
  for (int i = 0; i < D; i++) {
    T dc = wrap_dx<T>(x1[i] - x2[i], boxsize);
    T dx = max(fabs(dc) - wh[i], T(0));
    r2 = fmaT(dx, dx, r2);
  }
  return r2;
}
template <int D, typename T>
__device__ __forceinline__ T maxdist2(const Vec<D, T>& x1, const Vec<D, T>& x2, const Vec<D, T>& wh, T boxsize) {
  T r2 = T(0);
#pragma unroll
  for (int i = 0; i < D; i++) {
    T dc = wrap_dx<T>(x1[i] - x2[i], boxsize);
    T dx = max(fabs(dc) + wh[i], T(0));
    r2 = fmaT(dx, dx, r2);
  }
  return r2;
}

// half extents of a node of binary level `level` (math.cuh:583-594)
template <int D, typename T> __device__ __forceinline__ Vec<D, T> lvl_half_ext(int level) {
  int l[D];
  lvl_vec<D>(level, l);
  Vec<D, T> e;
#pragma unroll
  for (int i = 0; i < D; i++) e[i] = pow2<T>(T(1), l[i] - 1);
  return e;
}
template <int D, typename T> __device__ __forceinline__ Vec<D, T> lvl_ext(int level) {
  int l[D];
  lvl_vec<D>(level, l);
  Vec<D, T> e;
#pragma unroll
  for (int i = 0; i < D; i++) e[i] = pow2<T>(T(1), l[i]);
  return e;
}

// Relative slack applied to every node-level distance bound used for pruning.  The bounds and
// the particle r2 are each computed with a few FP32 roundings (relative error < 2^-21); a slack
// of 2^-16 keeps every pruning decision conservative, so kNN lists and FoF groups are exactly
// those of the particle-level predicate.
constexpr float kBoundSlack = 1.52587890625e-05f;  // 2^-16

// ------------------------------------------------------------------------------------------------
// wave helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

template <typename T> __device__ __forceinline__ T uniform(T x) {
  return (T)__builtin_amdgcn_readfirstlane((int)x);
}

// coherent (L2) relaxed atomic load/store, agent scope: vector memory instructions only
__device__ __forceinline__ int ld_acq(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acq(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }
// Grid of a kernel whose blocks stride over `n` work items: at most kMaxGrid blocks, so launching
// over a padded allocation costs no block dispatch per empty slot.
constexpr int64_t kMaxGrid = 65536;
inline unsigned grid_for(int64_t n) { return (unsigned)(n <= 0 ? 1 : (n < kMaxGrid ? n : kMaxGrid)); }

}  // namespace jzt
