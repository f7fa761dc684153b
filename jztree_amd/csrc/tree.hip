// Leaf and node construction from z-sorted positions (replaces jztree's tree.cuh kernels).
//
//  FlagLeafBoundaries  tree.cuh:13-98    boundary b (between particles b-1 and b) is a leaf split if
//                                        the binary Morton cell it separates holds > max_size
//                                        particles (counted in an LDS window of scan_size boundaries)
//  FindNodeBoundaries  tree.cuh:104-199  per boundary of the leaf representatives: its level and the
//                                        [lbound, rbound) extent of the node it splits (binary search)
//  GetNodeGeometry     tree.cuh:312-372  level, centre and extent of the smallest cell holding a range
//
// All integer/bit exact.  Launch: 256 threads (4 waves) per workgroup, one node per lane (four
// boundaries per lane in FlagLeafBoundaries).
#include "util.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr int kTreeThreads = 256;
constexpr int kFlagPer = 4;  // boundaries per thread in k_flag_leaf: the LDS window's halo is shared by more of them
constexpr int kFlagSpan = kTreeThreads * kFlagPer;
constexpr int kFlagChunk = 8;  // levels per block in k_flag_leaf's outward scans
constexpr int kFlagPad = 8 * kFlagChunk + kFlagChunk;  // slots past either end of its LDS window

template <int D, typename T>
__global__ __launch_bounds__(kTreeThreads) void k_flag_leaf(const T* __restrict__ posz, const uint8_t* __restrict__ ptype,
                                                            const int32_t* __restrict__ lvl_bound,
                                                            const int32_t* __restrict__ npart_p,
                                                            int8_t* __restrict__ split_flags, int32_t* __restrict__ lvl,
                                                            int64_t size_part, int max_size, int scan_size,
                                                            int num_types, const int32_t* __restrict__ seg) {
  extern __shared__ int32_t s_mem[];
  const int nload = kFlagSpan + 2 * scan_size + 1;
  const int nslot = nload + 2 * kFlagPad;
  // levels, and the maximum of each run of kFlagChunk levels ending at a slot; kFlagPad slots on
  // either side hold INT_MIN (the scans may read past the window: those reads never hit)
  int32_t* const s_level = s_mem + kFlagPad;
  int32_t* const s_max8 = s_mem + nslot + kFlagPad;
  uint8_t* s_types = reinterpret_cast<uint8_t*>(s_mem + 2 * nslot);
  const int64_t nump = npart_p[0];
  const int lb_l = lvl_bound[0], lb_r = lvl_bound[1];
  const int64_t ioff = (int64_t)blockIdx.x * kFlagSpan - scan_size - 1;

  // s_level[i] = level of boundary ioff+i+1, i.e. between particles ioff+i and ioff+i+1
  for (int i = threadIdx.x; i < nload - 1; i += kTreeThreads) {
    int64_t ip = ioff + i;
    int L;
    uint8_t ty = 0;
    if (ip < 0) {
      L = lb_l;
    } else if (ip + 1 >= nump) {
      L = lb_r;
    } else {
      L = (seg && seg[ip] != seg[ip + 1]) ? kLvlSet
                                          : msb_diff_level<D, T>(load_vec<D, T>(posz, ip), load_vec<D, T>(posz, ip + 1));
      if (num_types > 1) ty = ptype[ip];
    }
    s_level[i] = L;
    s_types[i] = ty;
  }
  for (int i = threadIdx.x; i < 2 * kFlagPad + 1; i += kTreeThreads) {
    const int j = i < kFlagPad ? i - kFlagPad : nload - 1 + (i - kFlagPad);
    s_level[j] = INT32_MIN;
    s_max8[j] = INT32_MIN;
  }
  __syncthreads();
  if (num_types == 1) {
    for (int i = threadIdx.x; i < nload - 1; i += kTreeThreads) {
      int m = s_level[i];
#pragma unroll
      for (int u = 1; u < kFlagChunk; u++) m = max(m, s_level[i - u]);
      s_max8[i] = m;
    }
    __syncthreads();
  }

  for (int rep = 0; rep < kFlagPer; rep++) {
    const int64_t node_idx = (int64_t)blockIdx.x * kFlagSpan + rep * kTreeThreads + threadIdx.x;
    const int idx = rep * kTreeThreads + threadIdx.x + scan_size;
    const int my = s_level[idx];
    bool is_split = false;
    if (num_types == 1) {
      // one type: lsize = idx - (the nearest i < idx in the window with s_level[i] >= my), or the whole
      // window if there is none -- what the reset-and-count loop below leaves -- i.e. the distance
      // to the nearest level >= my on the left, capped at scan_size (likewise rsize to the right).
      // The scan goes by blocks of kFlagChunk levels: eight block maxima per step (one LDS round
      // trip covers 64 levels), then the levels of the first block that holds a hit.  Reads past
      // the window land in the INT_MIN pads or beyond scan_size, where the cap hides them.
      const int32_t* const here = s_level + idx;
      const int32_t* const here8 = s_max8 + idx;
      auto nearest = [&](auto right_c) {
        constexpr bool R = decltype(right_c)::value;
        for (int d0 = 1; d0 <= scan_size; d0 += 8 * kFlagChunk) {
          int mb[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int d = d0 + u * kFlagChunk;  // block of distances [d, d + kFlagChunk - 1]
            mb[u] = R ? here8[d + kFlagChunk - 1] : here8[-d];
          }
          int blk = -1;
#pragma unroll
          for (int u = 7; u >= 0; u--) blk = mb[u] >= my ? u : blk;
          if (blk >= 0) {
            const int d = d0 + blk * kFlagChunk;
            int lv[kFlagChunk];
#pragma unroll
            for (int v = 0; v < kFlagChunk; v++) lv[v] = R ? here[d + v] : here[-(d + v)];
            int hit = d + kFlagChunk - 1;  // the block max is a hit, so some v matches
#pragma unroll
            for (int v = kFlagChunk - 1; v >= 0; v--) hit = lv[v] >= my ? d + v : hit;
            return min(hit, scan_size);
          }
        }
        return scan_size;
      };
      const int lsize = nearest(std::false_type());
      const int rsize = nearest(std::true_type());
      is_split |= (lsize + rsize > max_size);
      is_split |= (my > lb_l) && (node_idx - lsize <= 0);
      is_split |= (my > lb_r) && (node_idx + rsize >= nump);
    } else {
      for (int type = 0; type < num_types; type++) {
        int lsize = 0;
        for (int i = idx - scan_size; i < idx; i++) lsize = (s_level[i] < my ? lsize : 0) + (s_types[i] == type ? 1 : 0);
        int rsize = 0;
        for (int i = idx + scan_size; i > idx; i--) rsize = (s_level[i] < my ? rsize : 0) + (s_types[i] == type ? 1 : 0);
        is_split |= (lsize + rsize > max_size);
        is_split |= (my > lb_l) && (node_idx - lsize <= 0);
        is_split |= (my > lb_r) && (node_idx + rsize >= nump);
      }
    }
    is_split |= my >= kLvlSet;  // set boundary of a batch
    is_split &= node_idx <= nump;
    is_split |= node_idx == 0;
    is_split |= node_idx == nump;
    if (node_idx < size_part + 1) {
      split_flags[node_idx] = (int8_t)is_split;
      lvl[node_idx] = my;
    }
  }
}

template <int D, typename T>
__global__ __launch_bounds__(kTreeThreads) void k_find_node_bounds(const T* __restrict__ pos_in,
                                                                   const T* __restrict__ pos_boundary,
                                                                   const int32_t* __restrict__ nleaves_p,
                                                                   int32_t* __restrict__ levels, int32_t* __restrict__ lbound,
                                                                   int32_t* __restrict__ rbound, int64_t size_nodes,
                                                                   int lvl_invalid, const int32_t* __restrict__ seg) {
  int64_t idx = (int64_t)blockIdx.x * kTreeThreads + threadIdx.x;
  const int n = nleaves_p[0];
  const int nnodes = n + 1;
  if (idx >= size_nodes) return;
  if (idx >= nnodes) {
    levels[idx] = lvl_invalid;
    lbound[idx] = n;
    rbound[idx] = n;
    return;
  }
  Vec<D, T> p1 = (idx == 0) ? load_vec<D, T>(pos_boundary, 0) : load_vec<D, T>(pos_in, idx - 1);
  Vec<D, T> p2 = (idx == nnodes - 1) ? load_vec<D, T>(pos_boundary, 1) : load_vec<D, T>(pos_in, idx);
  if (n <= 0) {  // no leaves: the single boundary spans nothing (the reference would read pos_in[-1])
    levels[idx] = msb_diff_level<D, T>(p1, p2);
    lbound[idx] = 0;
    rbound[idx] = 0;
    return;
  }
  // batch sets: points of different sets are separated at kLvlSet; the array ends take the set of
  // their neighbour.  Sets are contiguous, so levels stay monotone along the searches below.
  const int64_t ia = idx > 0 ? idx - 1 : 0, ib = idx < n ? idx : n - 1;
  auto level = [&](int64_t i, const Vec<D, T>& a, int64_t j, const Vec<D, T>& b) {
    return (seg && seg[i] != seg[j]) ? kLvlSet : msb_diff_level<D, T>(a, b);
  };
  const int target = level(ia, p1, ib, p2);
  int lb, rb;
  if (level(0, load_vec<D, T>(pos_in, 0), ib, p2) <= target) {
    lb = 0;
  } else {
    int ibefore = 0, iinside = (int)idx;
    while (ibefore + 1 < iinside) {
      int it = (ibefore + iinside) / 2;
      if (level(it, load_vec<D, T>(pos_in, it), ib, p2) > target) ibefore = it;
      else iinside = it;
    }
    lb = iinside;
  }
  if (level(ia, p1, n - 1, load_vec<D, T>(pos_in, n - 1)) <= target) {
    rb = n;
  } else {
    int iinside = (int)idx - 1, iafter = n - 1;
    while (iinside + 1 < iafter) {
      int it = (iinside + iafter) / 2;
      if (level(ia, p1, it, load_vec<D, T>(pos_in, it)) > target) iafter = it;
      else iinside = it;
    }
    rb = iafter;
  }
  levels[idx] = target;
  lbound[idx] = lb;
  rbound[idx] = rb;
}

template <typename T> __device__ __forceinline__ T quiet_nan();
template <> __device__ __forceinline__ float quiet_nan<float>() { return __uint_as_float(0x7fc00000u); }
template <> __device__ __forceinline__ double quiet_nan<double>() { return __longlong_as_double(0x7ff8000000000000ll); }

template <int D, typename T>
__global__ __launch_bounds__(kTreeThreads) void k_node_geometry(const T* __restrict__ pos, const int32_t* __restrict__ lbound,
                                                                const int32_t* __restrict__ rbound,
                                                                const int32_t* __restrict__ nnodes_p,
                                                                int32_t* __restrict__ level, T* __restrict__ center,
                                                                T* __restrict__ extent, int64_t size_nodes,
                                                                int64_t size_part, int lvl_invalid, uint32_t mode,
                                                                int upper) {
  int64_t idx = (int64_t)blockIdx.x * kTreeThreads + threadIdx.x;
  if (idx >= size_nodes) return;
  if (idx >= nnodes_p[0]) {
    if (mode & 1u) level[idx] = lvl_invalid;
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (mode & 2u) center[idx * D + d] = quiet_nan<T>();
      if (mode & 4u) extent[idx * D + d] = T(0);
    }
    return;
  }
  auto clip = [&](int64_t a) { return a < 0 ? (int64_t)0 : (a > size_part - 1 ? size_part - 1 : a); };
  const int64_t lb = lbound[idx], rb = rbound[idx];
  Vec<D, T> x0b = load_vec<D, T>(pos, clip(lb));
  Vec<D, T> x1a = load_vec<D, T>(pos, clip(rb - 1));
  int lvl;
  if (!upper) {
    lvl = msb_diff_level<D, T>(x0b, x1a);
  } else {
    Vec<D, T> x0a = load_vec<D, T>(pos, clip(lb - 1));
    Vec<D, T> x1b = load_vec<D, T>(pos, clip(rb));
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (lb - 1 < 0) x0a[d] = -INFINITY;
      if (rb >= size_part) x1b[d] = INFINITY;
    }
    lvl = min(msb_diff_level<D, T>(x0a, x0b), msb_diff_level<D, T>(x1a, x1b)) - 1;
  }
  int l[D];
  lvl_vec<D>(lvl, l);
  if (mode & 1u) level[idx] = lvl;
#pragma unroll
  for (int d = 0; d < D; d++) {
    if (mode & 2u) center[idx * D + d] = round_pow2_cent(x0b[d], l[d]);
    if (mode & 4u) extent[idx * D + d] = pow2<T>(T(1), l[d]);
  }
}

template <int D, typename T>
int flag_leaf_impl(hipStream_t s, const void* posz, const uint8_t* ptype, const int32_t* lvl_bound, const int32_t* npart,
                   int8_t* flags, int32_t* lvl, int64_t size_part, int max_size, int scan_size, int num_types,
                   const int32_t* seg) {
  int64_t nb = size_part + 1;
  int nload = kFlagSpan + 2 * scan_size + 1;
  size_t shmem = (size_t)(2 * (nload + 2 * kFlagPad)) * sizeof(int32_t) + (size_t)nload;
  if (shmem > 60000) return fail("FlagLeafBoundaries: scan_size too large");
  hipLaunchKernelGGL((k_flag_leaf<D, T>), dim3((unsigned)div_up(nb, kFlagSpan)), dim3(kTreeThreads), shmem, s,
                     (const T*)posz, ptype, lvl_bound, npart, flags, lvl, size_part, max_size, scan_size, num_types,
                     seg);
  return check_launch("FlagLeafBoundaries");
}

template <int D, typename T>
int find_node_impl(hipStream_t s, const void* pos_in, const void* pos_boundary, const int32_t* nleaves, int32_t* lv,
                   int32_t* lb, int32_t* rb, int64_t size_nodes, int lvl_invalid, const int32_t* seg) {
  hipLaunchKernelGGL((k_find_node_bounds<D, T>), dim3((unsigned)div_up(size_nodes, kTreeThreads)), dim3(kTreeThreads),
                     0, s, (const T*)pos_in, (const T*)pos_boundary, nleaves, lv, lb, rb, size_nodes, lvl_invalid, seg);
  return check_launch("FindNodeBoundaries");
}

template <int D, typename T>
int geometry_impl(hipStream_t s, const void* pos, const int32_t* lbound, const int32_t* rbound, const int32_t* nnodes,
                  int32_t* level, void* center, void* extent, int64_t size_nodes, int64_t size_part, int lvl_invalid,
                  uint32_t mode, int upper) {
  hipLaunchKernelGGL((k_node_geometry<D, T>), dim3((unsigned)div_up(size_nodes, kTreeThreads)), dim3(kTreeThreads), 0,
                     s, (const T*)pos, lbound, rbound, nnodes, level, (T*)center, (T*)extent, size_nodes, size_part,
                     lvl_invalid, mode, upper);
  return check_launch("GetNodeGeometry");
}

// ------------------------------------------------------------------------------------------------
// regularization level (tree.py:55-68 _level_histogram, :160-187 find_regularization_level)
// ------------------------------------------------------------------------------------------------
// Block-private histogram in LDS (levels cluster on a few bins: global atomics alone would
// serialise on them), merged with one global atomic per nonzero bin and block.  Level ranges past
// kRegLdsLevels bins (float64 positions in 4 or more dims: 2099 levels per dim) count straight
// into the global histogram.
constexpr int kRegLdsLevels = 8192;  // 64 KB of u64 bins
__global__ __launch_bounds__(256) void k_reg_hist(const int32_t* __restrict__ lvl, const int32_t* __restrict__ w,
                                                  int64_t n, int min_lvl, int nlvl,
                                                  unsigned long long* __restrict__ hist) {
  extern __shared__ unsigned long long s_hist[];
  const bool lds = nlvl <= kRegLdsLevels;
  unsigned long long* h = lds ? s_hist : hist;
  if (lds) {
    for (int j = threadIdx.x; j < nlvl; j += blockDim.x) s_hist[j] = 0ull;
    __syncthreads();
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int b = lvl[i] - min_lvl;
    b = b < 0 ? 0 : b;  // jnp.bincount clips below and drops above
    const int64_t wi = w ? (int64_t)w[i] : 1;
    if (b < nlvl && wi != 0) atomicAdd(&h[b], (unsigned long long)wi);
  }
  if (lds) {
    __syncthreads();
    for (int j = threadIdx.x; j < nlvl; j += blockDim.x)
      if (s_hist[j]) atomicAdd(&hist[j], s_hist[j]);
  }
}

// FP32 evaluation in ascending level order (the reference reduces under XLA's order; only
// borderline ties can differ).  The running sums are formed by one thread in that order; the two
// "first level whose running sum reaches ..." searches then test every level in parallel on those
// same values (a block-wide minimum of the matching indices), instead of a serial loop with an
// early exit.
constexpr int kRegThreads = 256;
__global__ __launch_bounds__(kRegThreads) void k_reg_level(const unsigned long long* __restrict__ hist, int nlvl,
                                                           int min_lvl, int max_lvl, float fac, float frac,
                                                           float* __restrict__ gbuf, int32_t* out) {
  extern __shared__ float s_f[];  // [0, nlvl) counts, [nlvl, 2 nlvl) running sums
  __shared__ int s_first[2];
  __shared__ float s_lim;
  // past kRegLdsLevels levels the two arrays live in global memory (gbuf, 2 nlvl floats); the
  // barriers below order the workgroup's global accesses as they do its LDS accesses
  float* s_cnt = nlvl <= kRegLdsLevels ? s_f : gbuf;
  float* s_cum = s_cnt + nlvl;
  const int tid = threadIdx.x;
  for (int b = tid; b < nlvl; b += kRegThreads) s_cnt[b] = (float)(int64_t)hist[b];
  if (tid == 0) s_first[0] = s_first[1] = INT_MAX;
  __syncthreads();
  if (tid == 0) {
    float c = 0.f;
    for (int b = 0; b < nlvl; b++) {
      c += s_cnt[b];
      s_cum[b] = c;
    }
  }
  __syncthreads();
  const float total = nlvl > 0 ? s_cum[nlvl - 1] : 0.f;
  const float thresh = frac * total;
  for (int b = tid; b < nlvl; b += kRegThreads)
    if (s_cum[b] >= thresh) atomicMin(&s_first[0], b);
  __syncthreads();
  const int ref = s_first[0] == INT_MAX ? max_lvl : s_first[0] + min_lvl;
  if (tid == 0) {
    float vol_sum = 0.f, cts_sum = 0.f;
    for (int b = 0; b < nlvl && b + min_lvl <= ref; b++) {
      float c = s_cnt[b];
      float v = ldexpf(1.0f, b + min_lvl - ref);
      vol_sum += v * c;
      cts_sum += c;
    }
    const float avg = vol_sum / cts_sum;
    s_lim = fac * avg;
  }
  __syncthreads();
  const float lim = s_lim;
  for (int b = tid; b < nlvl; b += kRegThreads) {
    float v = ldexpf(1.0f, b + min_lvl - ref);
    if (s_cum[b] >= thresh && v >= lim) atomicMin(&s_first[1], b);
  }
  __syncthreads();
  if (tid == 0) out[0] = s_first[1] == INT_MAX ? max_lvl : s_first[1] + min_lvl;
}

// Node geometry of a float64 tree for float32 distance kernels on the float32-rounded points.
// The node's Morton cell (c, level) holds its float64 points.  It is replaced by the smallest
// enclosing Morton cell whose per-dim sides are at least 2^(E_i - 23), where every coordinate of
// the cell has magnitude <= 2^E_i: the boundaries of that cell (multiples of its side) and its
// centre (multiples of half a side) are then float32 numbers, so rounding to float32, which is
// monotone and leaves those boundaries fixed, keeps every point inside it.  The float32 cell is
// therefore a conservative bound for the rounded points, as the float64 cell is for the originals.
// Levels past the float32 range (a node spanning signs) become the float32 "whole space" level.
template <int D>
__global__ void k_geometry_to_f32(const double* __restrict__ cent, const int32_t* __restrict__ lvl, int64_t n,
                                  int lvl_invalid, float* __restrict__ cent_out, int32_t* __restrict__ lvl_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int level = lvl[i];
  double c[D];
  bool finite = true;
#pragma unroll
  for (int d = 0; d < D; d++) {
    c[d] = cent[i * D + d];
    finite &= __builtin_isfinite(c[d]);
  }
  const int top32 = 128 * D;  // float32: cells spanning signs or infinities
  if (level == lvl_invalid || !finite || level >= top32) {
#pragma unroll
    for (int d = 0; d < D; d++) cent_out[i * D + d] = (float)c[d];
    lvl_out[i] = (level == lvl_invalid || !finite) ? level : top32;
    return;
  }
  int L = level;
  for (int guard = 0; guard < 64 * D; guard++) {
    int l[D];
    lvl_vec<D>(L, l);
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; d++) {
      // |x| <= |centre| + side for x in the cell; E = exponent with that bound <= 2^E
      const double m = fabs(round_pow2_cent(c[d], l[d])) + ldexp(1.0, l[d]);
      int e;
      frexp(m, &e);  // m < 2^e
      ok &= l[d] >= max(e - 23, -149);
    }
    if (ok) break;
    L++;
  }
  if (L >= top32) L = top32;
  int l[D];
  lvl_vec<D>(L, l);
#pragma unroll
  for (int d = 0; d < D; d++) cent_out[i * D + d] = (float)round_pow2_cent(c[d], l[d]);
  lvl_out[i] = L;
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_geometry_to_f32(jzt_stream_t stream, const double* cent, const int32_t* lvl, int64_t n, int dim,
                                   int lvl_invalid, float* cent_out, int32_t* lvl_out) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0) return 0;
  const unsigned g = (unsigned)div_up(n, 256);
#define JZT_G2F(D)                                                                                                 \
  case D:                                                                                                          \
    hipLaunchKernelGGL(k_geometry_to_f32<D>, dim3(g), dim3(256), 0, s, cent, lvl, n, lvl_invalid, cent_out, lvl_out); \
    break;
  switch (dim) {
    JZT_G2F(2) JZT_G2F(3) JZT_G2F(4) JZT_G2F(5) JZT_G2F(6) JZT_G2F(7) JZT_G2F(8) JZT_G2F(9) JZT_G2F(10)
    default:
      return fail("GeometryToF32: unsupported dim");
  }
#undef JZT_G2F
  return check_launch("GeometryToF32");
}

extern "C" size_t jzt_regularization_tmp_bytes(int min_lvl, int max_lvl) {
  // the u64 histogram, plus the level pass's float arrays when they do not fit in LDS
  const size_t nlvl = (size_t)(max_lvl - min_lvl + 1);
  return nlvl * 8 + (nlvl > (size_t)kRegLdsLevels ? nlvl * 8 : 0);
}

extern "C" int jzt_regularization_level(jzt_stream_t stream, const int32_t* lvl, const int32_t* weights, int64_t n,
                                        int32_t* out_level, int min_lvl, int max_lvl, float max_volume_fac,
                                        float min_percentile, void* tmp, size_t tmp_bytes) {
  hipStream_t s = (hipStream_t)stream;
  const int nlvl = max_lvl - min_lvl + 1;
  if (tmp_bytes < jzt_regularization_tmp_bytes(min_lvl, max_lvl)) return fail("regularization: tmp too small");
  if (nlvl < 1 || nlvl > (1 << 24)) return fail("regularization: level range out of bounds");
  const bool lds = nlvl <= kRegLdsLevels;
  if (fill_u32(s, tmp, 0u, (int64_t)nlvl * 2)) return 1;
  if (n > 0)
    hipLaunchKernelGGL(k_reg_hist, dim3((unsigned)min(div_up(n, 256), (int64_t)512)), dim3(256),
                       lds ? (size_t)nlvl * sizeof(unsigned long long) : 0, s, lvl, weights, n, min_lvl, nlvl,
                       (unsigned long long*)tmp);
  // 0.01*percentile is formed in double and rounded to float, as the reference's weak typing does
  const float frac = (float)(0.01 * (double)min_percentile);
  float* gbuf = lds ? nullptr : (float*)((char*)tmp + (size_t)nlvl * 8);
  hipLaunchKernelGGL(k_reg_level, dim3(1), dim3(kRegThreads), lds ? (size_t)nlvl * 2 * sizeof(float) : 0, s,
                     (const unsigned long long*)tmp, nlvl, min_lvl, max_lvl, max_volume_fac, frac, gbuf, out_level);
  return check_launch("regularization level");
}

// CenterOfMass (tree.cuh:378-429): per node, the mass-weighted mean of the points of its range
// [isplit[i], isplit[i+1]) and their total mass, out = (centre[D], mass) per node.  One thread per
// node sums its points in order, with Kahan compensation when `kahan` (the reference's default):
// y = m * x - c for a coordinate (formed as one fused multiply-add, the form a contracting
// compiler gives the reference's m * x followed by "add - c"), y = m - c for the mass, then
// t = sum + y, c = (t - sum) - y, sum = t.  The centre is (1 / mass) * sum.  An empty range or a
// non-positive total mass gives NaN in every slot.
template <int D, typename T>
__global__ void k_center_of_mass(const int32_t* __restrict__ isplit, const T* __restrict__ pos,
                                 const T* __restrict__ mass, T* __restrict__ out, int64_t nnodes, int kahan) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  const int s0 = isplit[i], s1 = isplit[i + 1];
  const T nan = (T)__builtin_nan("");
  T sum[D + 1], c[D + 1];
#pragma unroll
  for (int d = 0; d <= D; d++) sum[d] = c[d] = T(0);
  for (int p = s0; p < s1; p++) {
    const T m = mass[p];
#pragma unroll
    for (int d = 0; d <= D; d++) {
      const T x = d < D ? pos[(int64_t)p * D + d] : T(1);
      if (kahan) {
        const T y = d < D ? fmaT(m, x, -c[d]) : m - c[d];
        const T t = sum[d] + y;
        c[d] = (t - sum[d]) - y;
        sum[d] = t;
      } else {
        sum[d] = d < D ? fmaT(m, x, sum[d]) : sum[d] + m;
      }
    }
  }
  if (s0 < s1 && sum[D] > T(0)) {
    const T inv = T(1) / sum[D];
#pragma unroll
    for (int d = 0; d < D; d++) out[i * (D + 1) + d] = inv * sum[d];
    out[i * (D + 1) + D] = sum[D];
  } else {
#pragma unroll
    for (int d = 0; d <= D; d++) out[i * (D + 1) + d] = nan;
  }
}

template <int D, typename T>
int com_impl(hipStream_t s, const int32_t* isplit, const void* pos, const void* mass, void* out, int64_t nnodes,
             int kahan) {
  if (nnodes <= 0) return 0;
  hipLaunchKernelGGL((k_center_of_mass<D, T>), dim3((unsigned)div_up(nnodes, 256)), dim3(256), 0, s, isplit,
                     (const T*)pos, (const T*)mass, (T*)out, nnodes, kahan);
  return check_launch("CenterOfMass");
}

// single-set tree kernels: dims 2-10 (the reference's instances, _generate_ffi.py:9-21)
#define JZT_DISPATCH_ND(fn, ...)                                                  \
  do {                                                                            \
    switch ((dtype == JZT_F64 ? 100 : 0) + dim) {                                 \
      case 2: return fn<2, float>(__VA_ARGS__);                                   \
      case 3: return fn<3, float>(__VA_ARGS__);                                   \
      case 4: return fn<4, float>(__VA_ARGS__);                                   \
      case 5: return fn<5, float>(__VA_ARGS__);                                   \
      case 6: return fn<6, float>(__VA_ARGS__);                                   \
      case 7: return fn<7, float>(__VA_ARGS__);                                   \
      case 8: return fn<8, float>(__VA_ARGS__);                                   \
      case 9: return fn<9, float>(__VA_ARGS__);                                   \
      case 10: return fn<10, float>(__VA_ARGS__);                                 \
      case 102: return fn<2, double>(__VA_ARGS__);                                \
      case 103: return fn<3, double>(__VA_ARGS__);                                \
      case 104: return fn<4, double>(__VA_ARGS__);                                \
      case 105: return fn<5, double>(__VA_ARGS__);                                \
      case 106: return fn<6, double>(__VA_ARGS__);                                \
      case 107: return fn<7, double>(__VA_ARGS__);                                \
      case 108: return fn<8, double>(__VA_ARGS__);                                \
      case 109: return fn<9, double>(__VA_ARGS__);                                \
      case 110: return fn<10, double>(__VA_ARGS__);                               \
      default: break;                                                             \
    }                                                                             \
    return fail(std::string(#fn) + ": unsupported dim/dtype");                    \
  } while (0)

#define JZT_DISPATCH(fn, ...)                                                     \
  do {                                                                            \
    if (dtype == JZT_F32 && dim == 3) return fn<3, float>(__VA_ARGS__);           \
    if (dtype == JZT_F32 && dim == 2) return fn<2, float>(__VA_ARGS__);           \
    if (dtype == JZT_F64 && dim == 3) return fn<3, double>(__VA_ARGS__);          \
    if (dtype == JZT_F64 && dim == 2) return fn<2, double>(__VA_ARGS__);          \
    return fail(std::string(#fn) + ": unsupported dim/dtype");                    \
  } while (0)

extern "C" int jzt_flag_leaf_boundaries(jzt_stream_t stream, const void* posz, const uint8_t* ptype,
                                        const int32_t* lvl_bound, const int32_t* npart, int8_t* split_flags,
                                        int32_t* lvl, int64_t size_part, int dim, int dtype, int max_size,
                                        int scan_size, int num_types) {
  if (num_types > 1 && ptype == nullptr) return fail("FlagLeafBoundaries: ptype required for num_types > 1");
  JZT_DISPATCH_ND(flag_leaf_impl, (hipStream_t)stream, posz, ptype, lvl_bound, npart, split_flags, lvl, size_part,
                  max_size, scan_size, num_types, (const int32_t*)nullptr);
}

extern "C" int jzt_flag_leaf_boundaries_sets(jzt_stream_t stream, const void* posz, const int32_t* set_id,
                                             const int32_t* lvl_bound, const int32_t* npart, int8_t* split_flags,
                                             int32_t* lvl, int64_t size_part, int dim, int dtype, int max_size,
                                             int scan_size) {
  if (set_id == nullptr) return fail("FlagLeafBoundaries (sets): set_id is required");
  JZT_DISPATCH(flag_leaf_impl, (hipStream_t)stream, posz, (const uint8_t*)nullptr, lvl_bound, npart, split_flags, lvl,
               size_part, max_size, scan_size, 1, set_id);
}

extern "C" int jzt_find_node_boundaries(jzt_stream_t stream, const void* pos_in, const void* pos_boundary,
                                        const int32_t* nleaves, int32_t* nodes_levels, int32_t* nodes_lbound,
                                        int32_t* nodes_rbound, int64_t size_nodes, int dim, int dtype, int lvl_max,
                                        int lvl_invalid) {
  (void)lvl_max;  // the reference passes it but its kernel never reads it outside dead stores (tree.cuh:144-196)
  if (size_nodes <= 0) return 0;
  JZT_DISPATCH_ND(find_node_impl, (hipStream_t)stream, pos_in, pos_boundary, nleaves, nodes_levels, nodes_lbound,
                  nodes_rbound, size_nodes, lvl_invalid, (const int32_t*)nullptr);
}

extern "C" int jzt_find_node_boundaries_sets(jzt_stream_t stream, const void* pos_in, const int32_t* set_id,
                                             const void* pos_boundary, const int32_t* nleaves, int32_t* nodes_levels,
                                             int32_t* nodes_lbound, int32_t* nodes_rbound, int64_t size_nodes, int dim,
                                             int dtype, int lvl_invalid) {
  if (set_id == nullptr) return fail("FindNodeBoundaries (sets): set_id is required");
  if (size_nodes <= 0) return 0;
  JZT_DISPATCH(find_node_impl, (hipStream_t)stream, pos_in, pos_boundary, nleaves, nodes_levels, nodes_lbound,
               nodes_rbound, size_nodes, lvl_invalid, set_id);
}

extern "C" int jzt_get_node_geometry(jzt_stream_t stream, const void* pos, const int32_t* lbound, const int32_t* rbound,
                                     const int32_t* nnodes, int32_t* level, void* center, void* extent,
                                     int64_t size_nodes, int64_t size_part, int dim, int dtype, int lvl_invalid,
                                     uint32_t mode_flags, int upper_extent) {
  if (size_nodes <= 0) return 0;
  JZT_DISPATCH_ND(geometry_impl, (hipStream_t)stream, pos, lbound, rbound, nnodes, level, center, extent, size_nodes,
                  size_part, lvl_invalid, mode_flags, upper_extent);
}

extern "C" size_t jzt_where_tmp_bytes(int64_t n) { return (size_t)(scan_tmp_words(n) + n) * 4; }

extern "C" int jzt_where_flags(jzt_stream_t stream, const int8_t* flags, int64_t n, int32_t* out, int64_t size,
                               int32_t fill, const int32_t* fill_dev, int32_t* count, void* tmp, size_t tmp_bytes) {
  if (tmp_bytes < jzt_where_tmp_bytes(n)) return fail("jzt_where_flags: tmp too small");
  return where_flags_i8((hipStream_t)stream, flags, n, out, size, fill, fill_dev, count, (int32_t*)tmp);
}

extern "C" int jzt_center_of_mass(jzt_stream_t stream, const int32_t* isplit, const void* pos, const void* mass,
                                  void* com_out, int64_t nnodes, int kahan, int dim, int dtype) {
  JZT_DISPATCH_ND(com_impl, (hipStream_t)stream, isplit, pos, mass, com_out, nnodes, kahan);
}
