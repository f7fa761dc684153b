// Exact k-nearest-neighbour search (replaces jztree's knn.cuh).
//
// KnnNode2Node (knn.cuh:435-694): for every child node Q of a parent, walk the parent's
//   r2-sorted interaction list and
//     FindRmax     accumulate (maxdist2, npart) bins until >= k particles are guaranteed; the
//                  bin radius (times 1 + 2^-16) bounds the k-th neighbour of every point in Q
//     Count/Insert every child T of the listed parents with mindist2(Q,T) <= rmax2(Q)
//     SegmentSort  each node's list ascending by (r2, src)
//   One wave64 workgroup per parent; its children sit in lanes; source children are staged in LDS.
//
// KnnLeaf2Leaf (knn.cuh:203-356): one wave64 per query leaf, one query per lane.  The leaf's
//   interaction list is walked in ascending r2 and stops once no lane could still accept
//   (r2T > kth*(1+2^-16) on all lanes).  Source particles are read at wave-uniform addresses
//   (scalar loads, broadcast to all lanes); each lane keeps its k best (r2, id) pairs sorted in
//   VGPRs as 64-bit keys (r2 bits << 32 | id), inserted with a static compare/select network.
//   k > 64 runs ceil(k/64) passes, each resuming after the previous pass's last key.
//
// Order: ascending by (canonical FP32 r2, reported id).  The reference inserts with strict '<' in
// visit order (knn.cuh:84-102), so its tie order depends on the traversal; here it does not.
#include "util.h"
#include "knn_common.h"
#include "knn_nd.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr float kRmaxSlack = 1.0f + kBoundSlack;
#ifndef JZT_LEAF_WAVES
#define JZT_LEAF_WAVES 4
#endif

template <int D, typename T> struct NodeT {
  Vec<D, T> cent;
  int level;
};

template <int D, typename T>
__device__ __forceinline__ NodeT<D, T> load_node(const T* __restrict__ nodes, int64_t i) {
  NodeT<D, T> n;
#pragma unroll
  for (int d = 0; d < D; d++) n.cent[d] = nodes[i * (D + 1) + d];
  T lv = nodes[i * (D + 1) + D];
  if constexpr (sizeof(T) == 4) n.level = __float_as_int(lv);
  else n.level = (int)__double_as_longlong(lv);
  return n;
}

template <int D, typename T>
__device__ __forceinline__ Vec<D, T> add(const Vec<D, T>& a, const Vec<D, T>& b) {
  Vec<D, T> r;
#pragma unroll
  for (int d = 0; d < D; d++) r[d] = a[d] + b[d];
  return r;
}

// ------------------------------------------------------------------------------------------------
// per-lane bins of (r2, count) for the guaranteed radius (knn.cuh:106-191, same semantics)
// ------------------------------------------------------------------------------------------------
template <int KB> struct RmaxBins {
  float r2s[KB];
  int cts[KB];
  int total, target;
  __device__ __forceinline__ RmaxBins(int tgt) {
#pragma unroll
    for (int i = 0; i < KB; i++) {
      r2s[i] = INFINITY;
      cts[i] = 0;
    }
    total = 0;
    target = tgt;
  }
  __device__ __forceinline__ float max_r2(int count) const {
    bool active = true;
    int sum = 0;
    float r2 = INFINITY;
#pragma unroll
    for (int i = 0; i < KB; i++) {
      sum += cts[i];
      r2 = (active && sum >= count) ? r2s[i] : r2;
      active = sum < count;
    }
    return r2;
  }
  __device__ __forceinline__ void insert_add(float r2, int num) {
    total += num;
    bool active = true;
#pragma unroll
    for (int i = 0; i < KB; i++) {
      bool a = active && (r2s[i] >= r2);
      cts[i] = a ? cts[i] + num : cts[i];
      active = active && !a;
    }
  }
  __device__ __forceinline__ void insert_shift(float r2, int num) {
    total = total + num - cts[KB - 1];
    bool active = true;
#pragma unroll
    for (int i = KB - 2; i >= 0; i--) {
      bool take = active && (r2 < r2s[i]);
      r2s[i + 1] = active ? (take ? r2s[i] : r2) : r2s[i + 1];
      cts[i + 1] = active ? (take ? cts[i] : num) : cts[i + 1];
      active = take;
    }
    r2s[0] = active ? r2 : r2s[0];
    cts[0] = active ? num : cts[0];
  }
  __device__ __forceinline__ void consider(float r2, int num) {
    if (r2 > r2s[KB - 1]) return;
    bool need_add = (cts[KB - 1] > 0) && (total + num - cts[KB - 1] < target);
    need_add |= r2 == r2s[KB - 1];
    if (need_add) insert_add(r2, num);
    else insert_shift(r2, num);
  }
};

#ifndef JZT_RMAX_GROUP
#define JZT_RMAX_GROUP 16
#endif
constexpr int kRmaxGroup = JZT_RMAX_GROUP;  // children per lane group of k_find_rmax_groups (8, 16 or 32)
static_assert(kRmaxGroup == 8 || kRmaxGroup == 16 || kRmaxGroup == 32, "lane group size");

template <int KB, int D, typename T>
__global__ __launch_bounds__(64) void k_find_rmax(const int32_t* __restrict__ pil_spl, const int32_t* __restrict__ pil_src,
                                                  const float* __restrict__ pil_r2, const int32_t* __restrict__ parent_spl,
                                                  const T* __restrict__ nodes, const int32_t* __restrict__ npart,
                                                  float* __restrict__ rmax_out, int k, T boxsize, int64_t pil_size, int64_t nitems,
                                                  int skip_le) {
  __shared__ Vec<D, T> s_x[64];
  __shared__ Vec<D, T> s_e[64];
  __shared__ int s_n[64];
  const int lane = threadIdx.x;
  const int64_t nact_ = active_items(parent_spl, nitems);
  for (int pq = blockIdx.x; pq < nact_; pq += gridDim.x) {
    const int q0 = parent_spl[pq], q1 = parent_spl[pq + 1];
    if (q1 - q0 <= skip_le) continue;  // done by k_find_rmax_groups
    const int e0 = pil_spl[pq], e1 = (int)min((int64_t)pil_spl[pq + 1], pil_size);
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int iq = min(qoff + lane, q1 - 1);
      NodeT<D, T> nq = load_node<D, T>(nodes, iq);
      Vec<D, T> eq = lvl_half_ext<D, T>(nq.level);
      RmaxBins<KB> bins(k);
      float rmax2 = INFINITY;
      for (int e = e0; e < e1; e++) {
        const int pt = pil_src[e];
        const float r2t = pil_r2[e];
        if (!wave_any(r2t <= rmax2)) break;
        const int t0 = parent_spl[pt], t1 = parent_spl[pt + 1];
        for (int toff = t0; toff < t1; toff += 64) {
          const int it = toff + lane;
          if (it < t1) {
            NodeT<D, T> nt = load_node<D, T>(nodes, it);
            s_x[lane] = nt.cent;
            s_e[lane] = lvl_half_ext<D, T>(nt.level);
            s_n[lane] = npart[it];
          }
          __syncthreads();
          const int nl = min(t1 - toff, 64);
          for (int j = 0; j < nl; j++) {
            float r2 = (float)maxdist2<D, T>(s_x[j], nq.cent, add<D, T>(eq, s_e[j]), boxsize);
            bins.consider(r2, s_n[j]);
          }
          __syncthreads();
        }
        rmax2 = bins.max_r2(k) * kRmaxSlack;
      }
      if (qoff + lane < q1) rmax_out[iq] = rmax2;
    }
  
  }
}

// FindRmax for parents with at most G children, 64 / G of them per wave: lane group g (G lanes)
// takes parent base + g, lane j of the group its child j.  Each group does exactly what
// k_find_rmax does for such a parent -- the same entries in list order, the same break test (no
// child of the parent reaches the next entry), the same targets in order -- so every child's bins
// see the same sequence and the radii are identical; only the lanes are shared.  The loop is
// wave-uniform: each step every unfinished group stages (up to G targets of) its current entry
// in its own LDS slice, and its lanes consider them.
template <int KB, int D, typename T, int G>
__global__ __launch_bounds__(64) void k_find_rmax_groups(const int32_t* __restrict__ pil_spl,
                                                         const int32_t* __restrict__ pil_src,
                                                         const float* __restrict__ pil_r2,
                                                         const int32_t* __restrict__ parent_spl,
                                                         const T* __restrict__ nodes, const int32_t* __restrict__ npart,
                                                         float* __restrict__ rmax_out, int k, T boxsize,
                                                         int64_t pil_size, int64_t nitems) {
  constexpr int NG = 64 / G;
  __shared__ Vec<D, T> s_x[64];
  __shared__ Vec<D, T> s_e[64];
  __shared__ int s_n[64];
  const int lane = threadIdx.x, g = lane / G, j = lane - g * G;
  const uint64_t gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (g * G);
  const int64_t nact_ = active_items(parent_spl, nitems);
  const int64_t nwaves = (nact_ + NG - 1) / NG;
  for (int64_t wv = blockIdx.x; wv < nwaves; wv += gridDim.x) {
    const int64_t pq = wv * NG + g;
    int q0 = 0, nc = 0, e = 0, e1 = 0;
    if (pq < nact_) {
      q0 = parent_spl[pq];
      nc = parent_spl[pq + 1] - q0;
      e = pil_spl[pq];
      e1 = (int)min((int64_t)pil_spl[pq + 1], pil_size);
    }
    // a group with more than G children (k_find_rmax's) or none has nothing to do
    bool gdone = nc <= 0 || nc > G;
    const bool valid = !gdone && j < nc;
    const int iq = q0 + min(j, max(nc - 1, 0));
    NodeT<D, T> nq = load_node<D, T>(nodes, valid ? iq : 0);
    const Vec<D, T> eq = lvl_half_ext<D, T>(nq.level);
    RmaxBins<KB> bins(k);
    float rmax2 = INFINITY;
    int toff = 0, t1 = 0;  // current entry's target range; toff >= t1: fetch the next entry
    while (__ballot(!gdone) != 0ull) {
      if (!gdone && toff >= t1) {
        bool go = e < e1;
        if (go) {
          const float r2t = pil_r2[e];
          go = (__ballot(valid && r2t <= rmax2) & gmask) != 0ull;
        }
        if (go) {
          const int pt = pil_src[e];
          toff = parent_spl[pt];
          t1 = parent_spl[pt + 1];
          e++;
        } else {
          gdone = true;
        }
      }
      const int it = toff + j;
      if (!gdone && it < t1 && j < G) {
        NodeT<D, T> nt = load_node<D, T>(nodes, it);
        s_x[lane] = nt.cent;
        s_e[lane] = lvl_half_ext<D, T>(nt.level);
        s_n[lane] = npart[it];
      }
      __syncthreads();
      if (!gdone) {
        const int nl = min(t1 - toff, G);
        if (valid) {
          for (int jj = 0; jj < nl; jj++) {
            const int sl = g * G + jj;
            float r2 = (float)maxdist2<D, T>(s_x[sl], nq.cent, add<D, T>(eq, s_e[sl]), boxsize);
            bins.consider(r2, s_n[sl]);
          }
        }
        toff += G;
        if (toff >= t1) rmax2 = bins.max_r2(k) * kRmaxSlack;  // the entry is done
      }
      __syncthreads();
    }
    if (valid) rmax_out[iq] = rmax2;
  }
}

// One wave per parent node; its children are staged in LDS, and for every entry of the parent's
// list the children of that entry are staged too.  Lanes take (child, target) pairs: with nt targets
// staged, lane l tests child c0 + l / nt against target l % nt, so a parent with 8 children and an
// entry with 8 targets keeps all 64 lanes busy (one lane per child would leave 56 idle).  A child's
// hits in one step are lanes of one contiguous block: counted with a ballot, and in the insert pass
// placed at the child's running offset plus the rank of the hit in its block.  The order inside a
// node's list is therefore not the reference's, but every list is sorted by (r2, src) afterwards
// (pairs are unique), so the lists are identical.
template <int PASS, int D, typename T>
__global__ __launch_bounds__(64) void k_count_insert(const int32_t* __restrict__ pil_spl, const int32_t* __restrict__ pil_src,
                                                     const float* __restrict__ pil_r2,
                                                     const int32_t* __restrict__ parent_spl, const T* __restrict__ nodes,
                                                     const float* __restrict__ rmax2, int32_t* __restrict__ count_out,
                                                     const int32_t* __restrict__ out_spl, int32_t* __restrict__ out_src,
                                                     float* __restrict__ out_r2, int64_t nmax, T boxsize,
                                                     int64_t pil_size, int64_t nitems) {
  __shared__ Vec<D, T> s_qc[64], s_qe[64], s_tc[64], s_te[64];
  __shared__ float s_qr[64];
  __shared__ int64_t s_run[64];  // per child: hits so far (count pass) or next output slot (insert pass)
  const int lane = threadIdx.x;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const int64_t nact_ = active_items(parent_spl, nitems);
  for (int pq = blockIdx.x; pq < nact_; pq += gridDim.x) {
    const int q0 = parent_spl[pq], q1 = parent_spl[pq + 1];
    const int e0 = pil_spl[pq], e1 = (int)min((int64_t)pil_spl[pq + 1], pil_size);
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int nc = min(q1 - qoff, 64);
      float rq = -INFINITY;
      if (lane < nc) {
        const NodeT<D, T> nq = load_node<D, T>(nodes, qoff + lane);
        s_qc[lane] = nq.cent;
        s_qe[lane] = lvl_half_ext<D, T>(nq.level);
        rq = rmax2[qoff + lane];
        s_qr[lane] = rq;
        s_run[lane] = (PASS == 1) ? (int64_t)out_spl[qoff + lane] : 0;
      }
      const float rq_max = wave_maxf(rq);
      __syncthreads();
      for (int e = e0; e < e1; e++) {
        const float r2t = pil_r2[e];
        if (!(r2t <= rq_max)) break;  // entries ascend by r2; no child reaches this one or any later
        const int pt = pil_src[e];
        const int t0 = parent_spl[pt], t1 = parent_spl[pt + 1];
        for (int toff = t0; toff < t1; toff += 64) {
          const int nt = min(t1 - toff, 64);
          if (lane < nt) {
            const NodeT<D, T> ntn = load_node<D, T>(nodes, toff + lane);
            s_tc[lane] = ntn.cent;
            s_te[lane] = lvl_half_ext<D, T>(ntn.level);
          }
          __syncthreads();
          const int cstep = nt >= 32 ? 1 : 64 / nt;
          const int my_c = lane / nt, my_t = lane - my_c * nt;
          const uint64_t block = (nt == 64 ? ~0ull : ((1ull << nt) - 1ull)) << (my_c * nt);
          for (int c0 = 0; c0 < nc; c0 += cstep) {
            const int c = c0 + my_c;
            const bool active = my_c < cstep && c < nc;
            bool hit = false;
            float r2 = 0.f;
            if (active) {
              const Vec<D, T> ew = add<D, T>(s_qe[c], s_te[my_t]);
              r2 = (float)mindist2<D, T>(s_tc[my_t], s_qc[c], ew, boxsize);
              hit = r2 <= s_qr[c];
              // Touching nodes are listed with a negative radius, -1 / maxdist2: they sort before every
              // node at a positive distance, the node itself first among them (smallest maxdist2), and
              // a consumer's "r2 <= k-th distance" test keeps them even when the k-th distance is
              // exactly 0 (k or more copies of a point): such a node may hold copies with smaller ids.
              // (A tiny positive radius, 1e-10 * maxdist2, failed that test and lost those ties.)
              if (PASS == 1 && hit && r2 == 0.f)
                r2 = -1.0f / fmaxf((float)maxdist2<D, T>(s_tc[my_t], s_qc[c], ew, boxsize), 1e-30f);
            }
            const uint64_t hits = __ballot(hit) & block;
            if (PASS == 1 && hit) {
              const int64_t o = s_run[c] + __popcll(hits & lt_mask);
              if (o < nmax) {
                out_r2[o] = r2;
                out_src[o] = toff + my_t;
              }
            }
            if (active && my_t == 0 && hits) s_run[c] += __popcll(hits);
          }
          __syncthreads();
        }
      }
      if (PASS == 0 && lane < nc) count_out[qoff + lane] = (int32_t)s_run[lane];
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------------------------------
// segmented sort by (key, val): bitonic network with only ascending comparators (the first stage of
// each merge compares mirrored pairs), so lengths need not be powers of two (virtual +inf tail)
// ------------------------------------------------------------------------------------------------
constexpr int kSegThreads = 256;
constexpr int kSegLds = 4096;  // entries held in LDS (32 KB)

__device__ __forceinline__ uint64_t kv_pack(float k, int32_t v) {
  uint32_t kb = __float_as_uint(k);
  kb = (kb & 0x80000000u) ? ~kb : (kb | 0x80000000u);  // total order of IEEE floats
  return ((uint64_t)kb << 32) | (uint32_t)v;
}
__device__ __forceinline__ float kv_key(uint64_t p) {
  uint32_t kb = (uint32_t)(p >> 32);
  kb = (kb & 0x80000000u) ? (kb & 0x7fffffffu) : ~kb;
  return __uint_as_float(kb);
}

template <typename Arr>
__device__ __forceinline__ void bitonic_sort_asc(Arr a, int len) {
  int np2 = 1;
  while (np2 < len) np2 <<= 1;
  for (int size = 2; size <= np2; size <<= 1) {
    // mirrored compare within each block of `size`
    for (int t = threadIdx.x; t < np2 / 2; t += blockDim.x) {
      int blk = t / (size / 2), off = t % (size / 2);
      int i = blk * size + off;
      int j = blk * size + size - 1 - off;
      if (j < len) {
        uint64_t x = a.load(i), y = a.load(j);
        if (y < x) {
          a.store(i, y);
          a.store(j, x);
        }
      }
    }
    __syncthreads();
    for (int stride = size / 4; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < np2 / 2; t += blockDim.x) {
        int blk = t / stride, off = t % stride;
        int i = blk * 2 * stride + off;
        int j = i + stride;
        if (j < len) {
          uint64_t x = a.load(i), y = a.load(j);
          if (y < x) {
            a.store(i, y);
            a.store(j, x);
          }
        }
      }
      __syncthreads();
    }
  }
}

struct LdsArr {
  uint64_t* p;
  __device__ __forceinline__ uint64_t load(int i) const { return p[i]; }
  __device__ __forceinline__ void store(int i, uint64_t x) const { p[i] = x; }
};
// (key, val) pairs sorted in place in global memory
struct GlobalKV {
  float* k;
  int32_t* v;
  __device__ __forceinline__ uint64_t load(int i) const { return kv_pack(k[i], v[i]); }
  __device__ __forceinline__ void store(int i, uint64_t x) const {
    k[i] = kv_key(x);
    v[i] = (int32_t)(uint32_t)x;
  }
};

// ---- wave-level segment sort ---------------------------------------------------------------
// One wave per segment of up to 64*E entries, the keys in registers: entry i*64 + lane lives in
// lane `lane`, register i.  Bitonic stages with stride >= 64 compare registers inside a lane,
// smaller strides exchange with a lane shuffle; no barriers, no LDS.
constexpr int kWaveSortMax = 1024;

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
  const int lo = __shfl_xor((int)(uint32_t)x, m, 64), hi = __shfl_xor((int)(uint32_t)(x >> 32), m, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <int E>
__device__ __forceinline__ void wave_bitonic_sort(uint64_t (&v)[E], int lane) {
  constexpr int N = 64 * E;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size / 2; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int rs = stride / 64;
#pragma unroll
        for (int i = 0; i < E; i++) {
          const int j = i ^ rs;
          if (j > i) {
            const bool asc = ((i * 64 + lane) & size) == 0;
            const uint64_t a = v[i], b = v[j];
            const bool sw = asc ? (b < a) : (a < b);
            v[i] = sw ? b : a;
            v[j] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < E; i++) {
          const uint64_t o = shfl_xor_u64(v[i], stride);
          const bool asc = ((i * 64 + lane) & size) == 0;
          const bool lower = (lane & stride) == 0;
          const bool take_min = lower == asc;
          const bool o_less = o < v[i];
          v[i] = (take_min == o_less) ? o : v[i];
        }
      }
    }
  }
}

template <int E>
__device__ __forceinline__ void wave_sort_segment(float* key, int32_t* val, const float* key_in, const int32_t* val_in,
                                                  int64_t b, int len, int lane) {
  uint64_t v[E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int idx = i * 64 + lane;
    v[i] = idx < len ? kv_pack(key_in[b + idx], val_in[b + idx]) : ~0ull;
  }
  wave_bitonic_sort<E>(v, lane);
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int idx = i * 64 + lane;
    if (idx < len) {
      key[b + idx] = kv_key(v[i]);
      val[b + idx] = (int32_t)(uint32_t)v[i];
    }
  }
}

// Segments of up to kWaveSortMax entries, one wave each (in place allowed: a wave reads its whole
// segment into registers before writing); longer ones are left to k_segment_sort.
__global__ __launch_bounds__(64) void k_segment_sort_wave(const int32_t* __restrict__ spl, float* key, int32_t* val,
                                                          const float* key_in, const int32_t* val_in, int64_t nmax, int64_t nitems) {
  const int64_t nact_ = active_items(spl, nitems);
  for (int seg = blockIdx.x; seg < nact_; seg += gridDim.x) {
    const int lane = threadIdx.x;
    int64_t b = spl[seg], e = spl[seg + 1];
    if (e > nmax) e = nmax;
    if (b >= e) continue;
    const int64_t len = e - b;
    if (len > kWaveSortMax) continue;
    if (len <= 64) wave_sort_segment<1>(key, val, key_in, val_in, b, (int)len, lane);
    else if (len <= 128) wave_sort_segment<2>(key, val, key_in, val_in, b, (int)len, lane);
    else if (len <= 256) wave_sort_segment<4>(key, val, key_in, val_in, b, (int)len, lane);
    else if (len <= 512) wave_sort_segment<8>(key, val, key_in, val_in, b, (int)len, lane);
    else wave_sort_segment<16>(key, val, key_in, val_in, b, (int)len, lane);
  
  }
}

// One workgroup per segment.  Segments up to kSegLds entries are sorted in LDS; longer ones in
// place in the output arrays (global memory, visible within the workgroup across barriers).
__global__ __launch_bounds__(kSegThreads) void k_segment_sort(const int32_t* __restrict__ spl, float* key,
                                                              int32_t* val, const float* key_in, const int32_t* val_in,
                                                              int64_t nmax, int skip_short, int64_t nitems) {
  __shared__ uint64_t s_a[kSegLds];
  const int64_t nact_ = active_items(spl, nitems);
  for (int seg = blockIdx.x; seg < nact_; seg += gridDim.x) {
    int64_t b = spl[seg], e = spl[seg + 1];
    if (e > nmax) e = nmax;
    if (b >= e) continue;
    const int len = (int)(e - b);
    if (len <= kWaveSortMax && skip_short) continue;  // done by k_segment_sort_wave
    if (len <= kSegLds) {
      for (int i = threadIdx.x; i < len; i += blockDim.x) s_a[i] = kv_pack(key_in[b + i], val_in[b + i]);
      __syncthreads();
      bitonic_sort_asc(LdsArr{s_a}, len);
      for (int i = threadIdx.x; i < len; i += blockDim.x) {
        key[b + i] = kv_key(s_a[i]);
        val[b + i] = (int32_t)(uint32_t)s_a[i];
      }
    } else {
      if (key != key_in) {
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
          key[b + i] = key_in[b + i];
          val[b + i] = val_in[b + i];
        }
        __syncthreads();
      }
      bitonic_sort_asc(GlobalKV{key + b, val + b}, len);
    }
    __syncthreads();  // the next segment reuses s_a
  
  }
}

// ------------------------------------------------------------------------------------------------
// leaf-vs-leaf top-k
// ------------------------------------------------------------------------------------------------
// One wave per query leaf.  Interaction-list metadata (source leaf, r2, particle range) is loaded
// 64 entries at a time, one per lane, and broadcast with readlane.  Source particles move in
// chunks of <= 64: the next chunk is loaded into registers (one particle per lane) while the
// current chunk is consumed from LDS, where every lane reads the same record (broadcast).
//
// Periodic boxes: a chunk is streamed without the wrap selects when, for every staged particle p
// and every dim, fl(qlo - p) >= -half and fl(qhi - p) < half, with [qlo, qhi] the query chunk's
// bounding box.  Rounding is monotone, so then every difference fl(q - p) lies in [-half, half)
// and the canonical wrap leaves it unchanged: the distances are bit-identical.
//
// Query leaves come from a per-XCD item queue (xcd_queue) when the caller's tmp holds its eight
// counters (jzt_knn_leaf2leaf_tmp_bytes includes them), else in the fixed xcd_items order: the
// leaves' costs vary widely, and with a fixed stride the last waves of the launch ran long after
// the others had finished (plain kNN at 2^26: 92.4 -> 87.8 ms with the queue).
//
// Optional start bound (init_r2, per query leaf, K <= 64 single pass): the slots start at
// (r2_0, id = 2^32-1), so only candidates with r2 <= r2_0 enter.  A leaf with r2_0 < 0 is skipped
// entirely (its rows and retry flags are not written), so a later pass can redo selected leaves.  A query that ends with fewer
// than K real neighbours is flagged in `retry` (its row holds the sentinel entries); every other
// query's row is exact, because all its K neighbours lie within r2_0 and the list holds every
// leaf within r2_0 (the caller's contract).
template <int K, bool PER, bool FLOOR, int D, int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) void k_knn_leaf(const int32_t* __restrict__ il_spl, const int32_t* __restrict__ il_src,
                                                 const float* __restrict__ il_r2, const int32_t* __restrict__ splT,
                                                 const float* __restrict__ xT, const int32_t* __restrict__ splQ,
                                                 const float* __restrict__ xQ, const int32_t* __restrict__ src_ids,
                                                 const int32_t* __restrict__ dst_rows,
                                                 const float* __restrict__ init_r2, int32_t* __restrict__ retry,
                                                 float* __restrict__ rad_out, int32_t* __restrict__ id_out,
                                                 uint64_t* __restrict__ resume, int koff, int ktot, float boxsize,
                                                 int64_t il_size, int64_t out_rows, int64_t nitems, int xcd,
                                                 uint32_t* __restrict__ queue) {
  __shared__ SrcRec s_rec[64];
  __shared__ uint64_t s_buf[kBufWords];
  const int lane = threadIdx.x;
  const int64_t nact = active_items(splQ, nitems);
  const ItemRange items = xcd_items(nact, xcd != 0);
  const ItemQueue iq_ = xcd_queue(nact, xcd != 0, queue);
  const int64_t lq_end = queue ? iq_.end : items.end;
  for (int64_t lq = queue ? queue_next(iq_) : items.first; lq < lq_end;
       lq = queue ? queue_next(iq_) : lq + items.step) {
    uint64_t* const row = s_buf + lane * kBufStride;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int q0 = splQ[lq], q1 = splQ[lq + 1];
    const int e0 = il_spl[lq], e1 = (int)min((int64_t)il_spl[lq + 1], il_size);
    const float half = 0.5f * boxsize;
    // a negative (or NaN) start bound marks a leaf that this launch leaves untouched
    const float r2_0 = init_r2 ? init_r2[lq] : 0.f;
    if (init_r2 && !(r2_0 >= 0.f)) continue;
    const uint64_t start_key = init_r2 ? (((uint64_t)__float_as_uint(r2_0) << 32) | 0xffffffffull) : kEmptyKey;
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int iq = min(qoff + lane, q1 - 1);
      const Vec<D, float> q = load_vec<D, float>(xQ, iq);
      const uint64_t floor_key = FLOOR ? resume[iq] : 0ull;
      Vec<D, float> qlo, qhi;  // bounding box of the query chunk
  #pragma unroll
      for (int d = 0; d < D; d++) {
        qlo[d] = wave_minf(q[d]);
        qhi[d] = wave_maxf(q[d]);
      }
      uint64_t best[K];
  #pragma unroll
      for (int i = 0; i < K; i++) best[i] = start_key;
      int cnt = 0;  // entries in this lane's buffer row
      bool done = false;
      // largest k-th distance of the wave (slack included); best[] changes only on flushes
      float kmax = wave_maxf(key_r2(best[K - 1])) * kRmaxSlack;
      for (int eb = e0; eb < e1 && !done; eb += 64) {
        const int ne = min(64, e1 - eb);
        int m_t0 = 0, m_t1 = 0;
        float m_r2 = INFINITY;
        if (lane < ne) {
          const int src = il_src[eb + lane];
          m_r2 = il_r2[eb + lane];
          m_t0 = splT[src];
          m_t1 = splT[src + 1];
        }
        // first chunk of this batch
        SrcRec cr;
        int cn;
        {
          const int t0 = rdlane(m_t0, 0), t1 = rdlane(m_t1, 0);
          cn = min(64, t1 - t0);
          if (lane < cn) {
            const Vec<D, float> p = load_vec<D, float>(xT, t0 + lane);
  #pragma unroll
            for (int d = 0; d < D; d++) cr.p[d] = p[d];
            cr.id = src_ids ? (uint32_t)src_ids[t0 + lane] : (uint32_t)(t0 + lane);
          }
        }
        int i = 0, off = 0;
        while (i < ne) {
          if (off == 0) {
            const float r2t = rdlane(m_r2, i);
            if (!wave_any(r2t <= key_r2(best[K - 1]) * kRmaxSlack)) {
              done = true;
              break;
            }
          }
          const int len = rdlane(m_t1, i) - rdlane(m_t0, i);
          int ni = i, noff = off + 64;
          if (noff >= len) {
            ni = i + 1;
            noff = 0;
          }
          const int n_cur = cn;
          // Source filter: p can enter some lane's top-k only if its gap to the query box is within
          // the largest current k-th distance of the wave (or a periodic wrap is possible); the
          // kept records are compacted into LDS.  Exact: by monotone rounding every canonical r2
          // from the box is >= the gap^2 computed here with the same FMA chain.
          bool need = false, reach = false;
          if (lane < n_cur) {
            float g2 = 0.f;
  #pragma unroll
            for (int d = 0; d < D; d++) {
              const float dlo = qlo[d] - cr.p[d], dhi = qhi[d] - cr.p[d];
              if (PER) need |= !(dlo >= -half) || !(dhi < half);
              const float gap = fmaxf(0.f, fmaxf(dlo, -dhi));
              g2 = fmaf(gap, gap, g2);
            }
            reach = need || g2 <= kmax;
          }
          const uint64_t keep = __ballot(reach);
          const bool wrap = PER && wave_any(need);
          const int n_keep = __popcll(keep);
          if (reach) s_rec[__popcll(keep & lt_mask)] = cr;
          __syncthreads();
          if (ni < ne) {
            const int t0 = rdlane(m_t0, ni) + noff, t1 = rdlane(m_t1, ni);
            cn = min(64, t1 - t0);
            if (lane < cn) {
              const Vec<D, float> p = load_vec<D, float>(xT, t0 + lane);
  #pragma unroll
              for (int d = 0; d < D; d++) cr.p[d] = p[d];
              cr.id = src_ids ? (uint32_t)src_ids[t0 + lane] : (uint32_t)(t0 + lane);
            }
          }
          if (consume_chunk<K, PER, FLOOR, D>(best, q, s_rec, n_keep, wrap, boxsize, half, floor_key, row, cnt))
            kmax = wave_maxf(key_r2(best[K - 1])) * kRmaxSlack;
          __syncthreads();
          i = ni;
          off = noff;
        }
      }
      if (wave_any(cnt > 0)) topk_flush<K>(best, row, cnt);
      if (qoff + lane < q1) {
        const int64_t orow = dst_rows ? (int64_t)dst_rows[iq] : (int64_t)iq;
        if (orow >= 0 && orow < out_rows) write_row<K>(best, rad_out, id_out, orow * ktot + koff, ktot - koff);
        // the last slot this pass reports (k need not be a power of two: slots past it may
        // keep the start key even when the query has all its neighbours)
        const int last = min(K, ktot - koff) - 1;
        uint64_t lk = best[0];
  #pragma unroll
        for (int i = 0; i < K; i++) lk = (i == last) ? best[i] : lk;
        if (retry) retry[iq] = (lk == start_key && init_r2) ? 1 : 0;
        if (resume) resume[iq] = lk;
      }
    }
  
  }
}

// Sort each segment by (key, val): segments up to kWaveSortMax entries one wave each in registers,
// longer ones by a workgroup (LDS up to kSegLds entries, else in place in global memory).
int segment_sort(hipStream_t s, const int32_t* spl, float* key, int32_t* val, const float* key_in,
                 const int32_t* val_in, int64_t nseg, int64_t nmax, const char* what) {
  if (nseg <= 0) return 0;
  hipLaunchKernelGGL(k_segment_sort_wave, dim3(grid_for(nseg)), dim3(64), 0, s, spl, key, val, key_in, val_in, nmax, nseg);
  if (check_launch(what)) return 1;
  // the two kernels touch disjoint segments
  hipLaunchKernelGGL(k_segment_sort, dim3(grid_for(nseg)), dim3(kSegThreads), 0, s, spl, key, val, key_in, val_in, nmax,
                     1, nseg);
  return check_launch(what);
}

__global__ void k_scale_rmax2(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * kRmaxSlack;
}

// Guaranteed k-neighbour radius^2 of the nodes of one plane, planes taken from the top down:
// cap[c] = squared cell diagonal of node c if it holds >= k particles, else its parent's cap (the
// parent: the last node P of the plane above with n2c_up[P] <= c, clamped into [0, size)), +inf
// at the top plane.  npart[c] = spl[n2l[c+1]] - spl[n2l[c]] (indices clamped to the spl array).
// The diagonal^2 sums 4^l_i over the per-dim levels of lvl_vec, in float64 (exact).
__global__ void k_plane_cap(const int32_t* __restrict__ n2l, const int32_t* __restrict__ spl, int64_t nspl,
                            const int32_t* __restrict__ lvl, const int32_t* __restrict__ n2c_up,
                            const double* __restrict__ cap_up, int64_t size, int k, int dim, double* __restrict__ cap) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= size) return;
  const int64_t a = min((int64_t)n2l[c], nspl - 1), b = min((int64_t)n2l[c + 1], nspl - 1);
  const int64_t np = (int64_t)spl[b] - (int64_t)spl[a];
  double r;
  if (np >= k) {
    const int o = (lvl[c] + 2000 * dim) / dim - 2000;
    const int rem = lvl[c] - o * dim;
    r = 0.0;
    for (int i = 0; i < dim; i++) r += ldexp(1.0, 2 * (o + (rem >= dim - i ? 1 : 0)));
  } else if (!n2c_up) {
    r = INFINITY;
  } else {
    int64_t lo = 0, hi = size + 1;  // count of n2c_up entries <= c
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (n2c_up[mid] <= c) lo = mid + 1;
      else hi = mid;
    }
    const int64_t parent = min(max(lo - 1, (int64_t)0), size - 1);
    r = cap_up[parent];
  }
  cap[c] = r;
}

// out[i] = max of v over the segment [spl[i], spl[i+1]) (clamped to [0, nv)), `empty` if it is empty
__global__ void k_segment_max(const int32_t* __restrict__ spl, const float* __restrict__ v, int64_t nv,
                              float* __restrict__ out, int64_t nseg, float empty) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  const int64_t a = min(max((int64_t)spl[i], (int64_t)0), nv), b = min(max((int64_t)spl[i + 1], (int64_t)0), nv);
  float m = empty;
  for (int64_t j = a; j < b; j++) m = fmaxf(m, v[j]);
  out[i] = m;
}

template <int D, typename T>
int node2node_impl(hipStream_t s, const int32_t* pil_spl, const int32_t* pil_src, const float* pil_r2,
                   const int32_t* parent_spl, const void* nodes, const int32_t* npart, const float* rmax2_in,
                   float* rmax2, int32_t* il_spl, int32_t* il_src, float* il_r2, int64_t size_parents,
                   int64_t size_nodes, int64_t pil_size, int64_t il_size, int k, float boxsize) {
  const T* nd = (const T*)nodes;
  if (fill_u32(s, il_spl, 0u, size_nodes + 1)) return 1;
  if (size_parents <= 0) return 0;
  dim3 g(grid_for(size_parents)), b(64);
  if (rmax2_in) {
    if (size_nodes > 0)
      hipLaunchKernelGGL(k_scale_rmax2, dim3((unsigned)div_up(size_nodes, 256)), dim3(256), 0, s, rmax2_in, rmax2,
                         size_nodes);
  } else {
    // parents with at most kRmaxGroup children: several per wave; the others: one wave each
    constexpr int G = kRmaxGroup;
    dim3 gg(grid_for(div_up(size_parents, 64 / G)));
#define JZT_FIND_RMAX(KB)                                                                                           \
  do {                                                                                                              \
    hipLaunchKernelGGL((k_find_rmax_groups<KB, D, T, G>), gg, b, 0, s, pil_spl, pil_src, pil_r2, parent_spl, nd,   \
                       npart, rmax2, k, (T)boxsize, pil_size, size_parents);                                        \
    hipLaunchKernelGGL((k_find_rmax<KB, D, T>), g, b, 0, s, pil_spl, pil_src, pil_r2, parent_spl, nd, npart, rmax2, \
                       k, (T)boxsize, pil_size, size_parents, G);                                                   \
  } while (0)
    if (k <= 16) JZT_FIND_RMAX(8);
    else if (k <= 32) JZT_FIND_RMAX(16);
    else JZT_FIND_RMAX(32);
#undef JZT_FIND_RMAX
  }
  if (check_launch("KnnNode2Node FindRmax")) return 1;
  hipLaunchKernelGGL((k_count_insert<0, D, T>), g, b, 0, s, pil_spl, pil_src, pil_r2, parent_spl, nd, rmax2, il_spl + 1,
                     (const int32_t*)nullptr, (int32_t*)nullptr, (float*)nullptr, il_size, (T)boxsize, pil_size, size_parents);
  // inclusive scan of the counts in place; the scan temp lives in the (not yet written) list
  if (scan_tmp_words(size_nodes) > il_size) return fail("KnnNode2Node: interaction list too small for scan temp");
  if (scan_i32(s, il_spl + 1, il_spl + 1, size_nodes, true, il_src)) return 1;
  hipLaunchKernelGGL((k_count_insert<1, D, T>), g, b, 0, s, pil_spl, pil_src, pil_r2, parent_spl, nd, rmax2,
                     (int32_t*)nullptr, il_spl, il_src, il_r2, il_size, (T)boxsize, pil_size, size_parents);
  if (check_launch("KnnNode2Node insert")) return 1;
  // sort each node's list by (r2, src) in place
  return segment_sort(s, il_spl, il_r2, il_src, il_r2, il_src, size_nodes, il_size, "KnnNode2Node sort");
}

template <int D>
int leaf2leaf_impl(hipStream_t s, const int32_t* il_spl, const int32_t* il_src, const float* il_r2, const int32_t* splT,
                   const void* xT, const int32_t* splQ, const void* xQ, const int32_t* src_ids,
                   const int32_t* dst_rows, const float* init_r2, int32_t* retry, void* rad, int32_t* ids, void* tmp,
                   size_t tmp_bytes, int64_t nleaves_q, int64_t npart_q, int64_t out_rows, int64_t il_size, int k,
                   float boxsize) {
  if (nleaves_q <= 0) return 0;
  if (init_r2 && k > 64) return fail("KnnLeaf2Leaf: a start bound needs k <= 64");
  uint64_t* resume = k > 64 ? (uint64_t*)tmp : nullptr;
  // item queue counters after the resume keys, when the caller's tmp holds them
  const size_t qoff = align256(knn_leaf_nd_tmp_bytes(npart_q, k, false));
  uint32_t* queue = tmp && tmp_bytes >= qoff + 256 ? (uint32_t*)((char*)tmp + qoff) : nullptr;
  dim3 g(grid_for8(nleaves_q)), b(64);
  const float* xt = (const float*)xT;
  const float* xq = (const float*)xQ;
  const int waves = tune_get(kTuneLeafWaves);
  const int xcd = tune_get(kTuneXcd);
  for (int koff = 0; koff < k; koff += 64) {
    const int kn = min(k - koff, 64);
    const bool floor = resume && koff > 0;
    if (queue && fill_u32(s, queue, 0u, 8)) return 1;
#define JZT_KNN_LAUNCH(KK, PER, FLOOR, W)                                                                              \
  hipLaunchKernelGGL((k_knn_leaf<KK, PER, FLOOR, D, W>), g, b, 0, s, il_spl, il_src, il_r2, splT, xt, splQ, xq, src_ids, \
                     dst_rows, init_r2, retry, (float*)rad, ids, resume, koff, k, boxsize, il_size, out_rows, nleaves_q, xcd, \
                     queue)
  // waves per EU (the register budget of the kernel): tunable for K <= 32; the multi-pass tail
  // (FLOOR) and K = 64 keep their fixed choice
#define JZT_KNN_LAUNCH_W(KK, PER)                                                  \
  do {                                                                             \
    constexpr int w2 = KK > 32 ? 1 : 2, w4 = KK > 32 ? 1 : JZT_LEAF_WAVES;         \
    if (KK > 32 || waves == 1) JZT_KNN_LAUNCH(KK, PER, false, 1);                  \
    else if (waves == 2) JZT_KNN_LAUNCH(KK, PER, false, w2);                       \
    else JZT_KNN_LAUNCH(KK, PER, false, w4);                                       \
  } while (0)
#define JZT_KNN_LAUNCH_K(KK)                                                            \
  do {                                                                                  \
    if (boxsize > 0.f) {                                                                \
      if (floor) JZT_KNN_LAUNCH(KK, true, true, (KK > 32 ? 1 : JZT_LEAF_WAVES));        \
      else JZT_KNN_LAUNCH_W(KK, true);                                                  \
    } else {                                                                            \
      if (floor) JZT_KNN_LAUNCH(KK, false, true, (KK > 32 ? 1 : JZT_LEAF_WAVES));       \
      else JZT_KNN_LAUNCH_W(KK, false);                                                 \
    }                                                                                   \
  } while (0)
    if (kn <= 8) JZT_KNN_LAUNCH_K(8);
    else if (kn <= 16) JZT_KNN_LAUNCH_K(16);
    else if (kn <= 32) JZT_KNN_LAUNCH_K(32);
    else JZT_KNN_LAUNCH_K(64);
#undef JZT_KNN_LAUNCH_K
#undef JZT_KNN_LAUNCH_W
#undef JZT_KNN_LAUNCH
    if (check_launch("KnnLeaf2Leaf")) return 1;
  }
  return 0;
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_knn_node2node(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                                 const float* parent_ilist_r2, const int32_t* parent_spl, const void* nodes,
                                 const int32_t* nodes_npart, float* node_rmax2, int32_t* node_ilist_spl,
                                 int32_t* node_ilist_isrc, float* node_ilist_r2, int64_t size_parents,
                                 int64_t size_nodes, int64_t parent_ilist_size, int64_t node_ilist_size, int k,
                                 float boxsize, int dim, int dtype) {
  return jzt_knn_node2node_ex(stream, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2, parent_spl, nodes,
                              nodes_npart, nullptr, node_rmax2, node_ilist_spl, node_ilist_isrc, node_ilist_r2,
                              size_parents, size_nodes, parent_ilist_size, node_ilist_size, k, boxsize, dim, dtype);
}

extern "C" int jzt_knn_node2node_ex(jzt_stream_t stream, const int32_t* parent_ilist_spl,
                                    const int32_t* parent_ilist_isrc, const float* parent_ilist_r2,
                                    const int32_t* parent_spl, const void* nodes, const int32_t* nodes_npart,
                                    const float* node_r2_in, float* node_rmax2, int32_t* node_ilist_spl,
                                    int32_t* node_ilist_isrc, float* node_ilist_r2, int64_t size_parents,
                                    int64_t size_nodes, int64_t parent_ilist_size, int64_t node_ilist_size, int k,
                                    float boxsize, int dim, int dtype) {
  hipStream_t s = (hipStream_t)stream;
  if (k < 1) return fail("KnnNode2Node: k must be >= 1");
  // dims 2-10, float or double node geometry (_generate_ffi.py:9-21)
#define JZT_N2N(D, T)                                                                                              \
  return node2node_impl<D, T>(s, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2, parent_spl, nodes,         \
                              nodes_npart, node_r2_in, node_rmax2, node_ilist_spl, node_ilist_isrc, node_ilist_r2, \
                              size_parents, size_nodes, parent_ilist_size, node_ilist_size, k, boxsize)
#define JZT_N2N_DIM(D)                  \
  case D:                               \
    if (dtype == JZT_F32) JZT_N2N(D, float); \
    JZT_N2N(D, double);
  if (dtype == JZT_F32 || dtype == JZT_F64) {
    switch (dim) {
      JZT_N2N_DIM(2) JZT_N2N_DIM(3) JZT_N2N_DIM(4) JZT_N2N_DIM(5) JZT_N2N_DIM(6) JZT_N2N_DIM(7) JZT_N2N_DIM(8)
      JZT_N2N_DIM(9) JZT_N2N_DIM(10)
      default: break;
    }
  }
#undef JZT_N2N_DIM
#undef JZT_N2N
  return fail("KnnNode2Node: unsupported dim/dtype");
}

extern "C" int jzt_segment_max_f32(jzt_stream_t stream, const int32_t* spl, const float* vals, int64_t nvals,
                                   float* out, int64_t nseg, float empty) {
  if (nseg <= 0) return 0;
  hipLaunchKernelGGL(k_segment_max, dim3((unsigned)div_up(nseg, 256)), dim3(256), 0, (hipStream_t)stream, spl, vals,
                     nvals, out, nseg, empty);
  return check_launch("SegmentMax");
}

extern "C" int jzt_plane_cap(jzt_stream_t stream, const int32_t* n2l, const int32_t* spl, int64_t nspl,
                             const int32_t* lvl, const int32_t* n2c_up, const double* cap_up, int64_t size, int k,
                             int dim, double* cap) {
  if (size <= 0) return 0;
  if (nspl < 1 || dim < 1) return fail("PlaneCap: empty splits or dim < 1");
  if ((n2c_up == nullptr) != (cap_up == nullptr)) return fail("PlaneCap: n2c_up and cap_up go together");
  hipLaunchKernelGGL(k_plane_cap, dim3((unsigned)div_up(size, 256)), dim3(256), 0, (hipStream_t)stream, n2l, spl, nspl,
                     lvl, n2c_up, cap_up, size, k, dim, cap);
  return check_launch("PlaneCap");
}

extern "C" size_t jzt_knn_leaf2leaf_tmp_bytes(int64_t size_part_query, int k, int dtype) {
  // float: the packed kernel's u64 resume key, or the (r2, id) pair of the any-dim kernel, both 8
  // bytes per query past 64 neighbours; double: a (double r2, u32 id) pair past 32
  // ... then 256 bytes of item-queue counters for the packed kernel
  return align256(knn_leaf_nd_tmp_bytes(size_part_query, k, dtype == JZT_F64)) + 256;
}

extern "C" int jzt_knn_leaf2leaf(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist_isrc,
                                 const float* ilist_r2, int64_t ilist_size, const int32_t* splT, const void* xT,
                                 const int32_t* splQ, const void* xQ, const int32_t* src_ids, void* knn_rad,
                                 int32_t* knn_id, void* tmp, size_t tmp_bytes, int64_t size_leaves_query,
                                 int64_t size_part_query, int k, float boxsize, int dim, int dtype) {
  return jzt_knn_leaf2leaf_ex(stream, ilist_spl, ilist_isrc, ilist_r2, ilist_size, splT, xT, splQ, xQ, src_ids,
                              nullptr, nullptr, nullptr, knn_rad, knn_id, tmp, tmp_bytes, size_leaves_query,
                              size_part_query, size_part_query, k, boxsize, dim, dtype);
}

extern "C" int jzt_knn_leaf2leaf_ex(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist_isrc,
                                    const float* ilist_r2, int64_t ilist_size, const int32_t* splT, const void* xT,
                                    const int32_t* splQ, const void* xQ, const int32_t* src_ids,
                                    const int32_t* dst_rows, const float* init_r2, int32_t* retry, void* knn_rad,
                                    int32_t* knn_id, void* tmp, size_t tmp_bytes, int64_t size_leaves_query,
                                    int64_t size_part_query, int64_t out_rows, int k, float boxsize, int dim,
                                    int dtype) {
  hipStream_t s = (hipStream_t)stream;
  if (k < 1) return fail("KnnLeaf2Leaf: k must be >= 1");
  // the multi-pass state is required; the work-queue counters after it are optional
  if (tmp_bytes < knn_leaf_nd_tmp_bytes(size_part_query, k, dtype == JZT_F64))
    return fail("KnnLeaf2Leaf: tmp too small");
  if (dtype != JZT_F32 && dtype != JZT_F64) return fail("KnnLeaf2Leaf: positions must be float32 or float64");
  if (dtype == JZT_F64 || dim > 3)
    return (dtype == JZT_F64 ? knn_leaf_nd_f64 : knn_leaf_nd_f32)(
        s, ilist_spl, ilist_isrc, ilist_r2, splT, xT, splQ, xQ, src_ids, dst_rows, init_r2, retry, knn_rad, knn_id, tmp,
        tmp_bytes, size_leaves_query, size_part_query, out_rows, ilist_size, k, boxsize, dim);
  if (dim == 3)
    return leaf2leaf_impl<3>(s, ilist_spl, ilist_isrc, ilist_r2, splT, xT, splQ, xQ, src_ids, dst_rows, init_r2, retry,
                             knn_rad, knn_id, tmp, tmp_bytes, size_leaves_query, size_part_query, out_rows, ilist_size, k,
                             boxsize);
  if (dim == 2)
    return leaf2leaf_impl<2>(s, ilist_spl, ilist_isrc, ilist_r2, splT, xT, splQ, xQ, src_ids, dst_rows, init_r2, retry,
                             knn_rad, knn_id, tmp, tmp_bytes, size_leaves_query, size_part_query, out_rows, ilist_size, k,
                             boxsize);
  return fail("KnnLeaf2Leaf: unsupported dim");
}

extern "C" int jzt_segment_sort(jzt_stream_t stream, const int32_t* spl, const float* key, const int32_t* val,
                                float* key_out, int32_t* val_out, int64_t size_segs, int64_t size_keys) {
  hipStream_t s = (hipStream_t)stream;
  if (size_segs <= 0) return 0;
  return segment_sort(s, spl, key_out, val_out, key, val, size_segs, size_keys, "SegmentSort");
}
