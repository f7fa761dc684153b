// Fused leaf-plane kNN: the last KnnNode2Node (leaf lists) and KnnLeaf2Leaf in one kernel.
//
// The reference builds, for every leaf, the r2-sorted list of leaves within its guaranteed radius
// (knn.cuh:602-694) and then walks that list with an early exit (knn.cuh:246-251).  Most of each
// list is never read.  Here one wave per query leaf Q walks the list of Q's *parent* (plane 1,
// already r2-sorted), and for each listed parent T_p evaluates its child leaves T on the fly:
//   - stop when no lane can still accept:  mindist2(P, T_p) > kth * (1 + 2^-16) on every lane;
//   - skip child T unless mindist2(Q, T) <= max over lanes of kth * (1 + 2^-16);
//   - otherwise stream T's particles through LDS (next chunk prefetched into registers) into the
//     per-lane sorted (r2, id) top-k.
// The top-k starts from a bound r2_0 (the parent's guaranteed radius, or a proposed radius): slots
// hold (r2_0, id = 2^32-1), so only candidates with r2 <= r2_0 enter.  If fewer than k real
// neighbours end up in the list the query is flagged, and a second launch with an infinite bound
// redoes exactly the flagged queries, so results never depend on the bound.
// Periodic wrapping: per (Q, T) pair, if the cell bounds show that no difference can reach half a
// box in any dim, the inner loop runs without the wrap selects; results are identical.
// Exactness: every particle within the final k-th r2 of a query lies in a child of a listed parent
// whose list position is not beyond the stop point, and in a child that is not skipped (the bounds
// are conservative by the slack), so the top-k equals the brute-force (r2, id) top-k.
#include "util.h"
#include "knn_common.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr float kSlack = 1.0f + kBoundSlack;

template <int D, typename T> struct Cell {
  Vec<D, T> c, h;
};
template <int D, typename T>
__device__ __forceinline__ Cell<D, T> load_cell(const T* __restrict__ nodes, int64_t i) {
  Cell<D, T> r;
#pragma unroll
  for (int d = 0; d < D; d++) r.c[d] = nodes[i * (D + 1) + d];
  T lv = nodes[i * (D + 1) + D];
  int level;
  if constexpr (sizeof(T) == 4) level = __float_as_int(lv);
  else level = (int)__double_as_longlong(lv);
  r.h = lvl_half_ext<D, T>(level);
  return r;
}

constexpr int kListCap = 512;  // LDS list entries per wave (4 KB)

// ascending bitonic sort of n <= kListCap u64 keys in LDS by one wave (only ascending comparators)
__device__ __forceinline__ void wave_sort(uint64_t* a, int n) {
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (int size = 2; size <= np2; size <<= 1) {
    for (int t = threadIdx.x; t < np2 / 2; t += 64) {
      const int blk = t / (size / 2), off = t % (size / 2);
      const int i = blk * size + off, j = blk * size + size - 1 - off;
      if (j < n) {
        const uint64_t x = a[i], y = a[j];
        if (y < x) {
          a[i] = y;
          a[j] = x;
        }
      }
    }
    __syncthreads();
    for (int stride = size / 4; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < np2 / 2; t += 64) {
        const int blk = t / stride, off = t % stride;
        const int i = blk * 2 * stride + off, j = i + stride;
        if (j < n) {
          const uint64_t x = a[i], y = a[j];
          if (y < x) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Consume the sorted list s_list[0..n) (key = mindist2 bits << 32 | wrap << 31 | leaf) nearest
// first; stop at the first leaf no lane can reach.  Particles are staged through LDS as 16-byte
// records, the next leaf's (first) chunk being loaded into registers while the current one is
// consumed.
template <int K, bool PER, int D>
__device__ __forceinline__ void consume_list(uint64_t (&best)[K], const Vec<D, float>& q, const uint64_t* s_list,
                                             int n, SrcRec* s_rec, const int32_t* __restrict__ l2p,
                                             const float* __restrict__ x, const int32_t* __restrict__ src_ids,
                                             float boxsize, float half, uint64_t* row, int& cnt) {
  const int lane = threadIdx.x;
  int e = 0, off = 0;
  SrcRec cr;
  int cn = 0;
  bool cwrap = false;
  auto load = [&](int ee, int oo) {
    const uint64_t v = s_list[ee];
    const int leaf = (int)(uint32_t)(v & 0x7fffffffull);
    const int a = l2p[leaf] + oo, b = l2p[leaf + 1];
    cn = min(64, b - a);
    cwrap = ((v >> 31) & 1ull) != 0;
    if (lane < cn) {
      const Vec<D, float> p = load_vec<D, float>(x, a + lane);
#pragma unroll
      for (int d = 0; d < D; d++) cr.p[d] = p[d];
      cr.id = src_ids ? (uint32_t)src_ids[a + lane] : (uint32_t)(a + lane);
    }
  };
  if (n <= 0) return;
  load(0, 0);
  while (e < n) {
    if (off == 0) {
      const float md = __uint_as_float((uint32_t)(s_list[e] >> 32));
      if (!wave_any(md <= key_r2(best[K - 1]) * kSlack)) return;
    }
    const int n_cur = cn;
    const bool w_cur = cwrap;
    int ne = e, noff = off + 64;
    {
      const uint64_t v = s_list[e];
      const int leaf = (int)(uint32_t)(v & 0x7fffffffull);
      if (noff >= l2p[leaf + 1] - l2p[leaf]) {
        ne = e + 1;
        noff = 0;
      }
    }
    if (lane < n_cur) s_rec[lane] = cr;
    __syncthreads();
    if (ne < n) load(ne, noff);
    consume_chunk<K, PER, false, D>(best, q, s_rec, n_cur, w_cur, boxsize, half, 0ull, row, cnt);
    __syncthreads();
    e = ne;
    off = noff;
  }
}

template <int K, bool PER, int D, typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 32 ? 4 : 1))) void k_knn_fused(
    const int32_t* __restrict__ pil_spl, const int32_t* __restrict__ pil_src, const float* __restrict__ pil_r2,
    int64_t pil_size, const int32_t* __restrict__ n2c, const int32_t* __restrict__ leaf_parent,
    const T* __restrict__ leaf_nodes, const int32_t* __restrict__ l2p, const T* __restrict__ x,
    const int32_t* __restrict__ src_ids, const int32_t* __restrict__ dst_rows, const float* __restrict__ init_r2,
    int32_t* __restrict__ retry, int only_retry, const float* __restrict__ retry_r2, T* __restrict__ rad_out,
    int32_t* __restrict__ id_out, int ktot, T boxsize, int64_t nitems, int64_t npart) {
  __shared__ SrcRec s_rec[64];
  __shared__ uint64_t s_list[kListCap];
  __shared__ uint64_t s_buf[kBufWords];
  uint64_t* const row = s_buf + threadIdx.x * kBufStride;
  const int lane = threadIdx.x;
  const int64_t nact_ = active_items(l2p, nitems);
  const T half = (T)(float)(T(0.5) * boxsize);
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  // the query chunk [qoff, min(qoff + 64, q1)) of query leaf lq
  auto run_chunk = [&](int lq, int qoff) {
    const int q1 = l2p[lq + 1];
    const int par = leaf_parent[lq];
    const int e0 = pil_spl[par], e1 = (int)min((int64_t)pil_spl[par + 1], pil_size);
    // the bounded pass starts at init_r2; the retry pass at retry_r2 (a bound every flagged query
    // is known to have k neighbours within), else unbounded
    const float r2_init = only_retry ? (retry_r2 ? retry_r2[lq] : INFINITY) : (init_r2 ? init_r2[lq] : INFINITY);
    {
      const int iq = min(qoff + lane, q1 - 1);
      const bool valid = qoff + lane < q1;
      if (only_retry && !wave_any(valid && retry[iq] != 0)) return;
      const Vec<D, T> q = load_vec<D, T>(x, iq);
      // A retry pass whose bound leaves some flagged query short redoes the chunk with no bound.
      // That second try is exact only if the parent lists hold every leaf it can need: true for
      // lists from the guaranteed-radius walk, and for lists built for retry_r2 itself as long as
      // retry_r2 is a true k-neighbour bound, so that the first try never comes up short.  The
      // caller's retry_r2 (the squared diagonal of the smallest enclosing cell with >= k
      // particles, widened by 2^-10) is such a bound: all k lie in that cell with the query.
      for (float r2_try = r2_init;; r2_try = INFINITY) {
      // In the retry pass only the flagged queries search: the others start full at key 0 (they
      // accept nothing and do not widen the wave's search radius) and write nothing.
      const bool searching = !only_retry || (valid && retry[iq] != 0);
      const uint64_t empty = searching ? (((uint64_t)__float_as_uint(r2_try) << 32) | 0xffffffffull) : 0ull;
      uint64_t best[K];
  #pragma unroll
      for (int i = 0; i < K; i++) best[i] = empty;
      int bcnt = 0;  // entries in this lane's top-k buffer row
      // Collect child leaves into s_list until the parent list is exhausted or the list is full,
      // then sort and consume them; a full list resumes collection at the child where it stopped.
      // (One consume site keeps one copy of the flush network in the kernel.)
      int eb = e0, ii = 0, cbo = 0;  // resume point: list batch, entry within it, child offset
      bool more = e0 < e1;
      while (true) {
        int cnt = 0;  // entries collected in s_list (wave-uniform)
        bool full = false;
        while (more && !full) {
          const Cell<D, T> cq = load_cell<D, T>(leaf_nodes, lq);  // (re)loaded here: not live across consume
          const int ne = min(64, e1 - eb);
          int m_c0 = 0, m_c1 = 0;
          float m_r2 = INFINITY;
          if (lane < ne) {
            const int src = pil_src[eb + lane];
            m_r2 = pil_r2[eb + lane];
            m_c0 = n2c[src];
            m_c1 = n2c[src + 1];
          }
          for (; ii < ne; ii++) {
            // the collection bound is the current k-th radius of the wave (the starting bound at first)
            const float kmax = wave_maxf(key_r2(best[K - 1])) * kSlack;
            if (!(rdlane(m_r2, ii) <= kmax)) {
              more = false;
              break;
            }
            const int c0 = rdlane(m_c0, ii), c1 = rdlane(m_c1, ii);
            for (int cb = c0 + cbo; cb < c1; cb += 64) {
              if (cnt + 64 > kListCap) {
                full = true;
                cbo = cb - c0;
                break;
              }
              const int nc = min(64, c1 - cb);
              bool keep = false;
              uint64_t item = 0;
              if (lane < nc) {
                const int lt = cb + lane;
                const Cell<D, T> ct = load_cell<D, T>(leaf_nodes, lt);
                Vec<D, T> w;
  #pragma unroll
                for (int d = 0; d < D; d++) w[d] = cq.h[d] + ct.h[d];
                const float md = (float)mindist2<D, T>(ct.c, cq.c, w, boxsize);
                keep = (md <= kmax) && (l2p[lt + 1] > l2p[lt]);
                int wrap = 0;
                if (PER) {
  #pragma unroll
                  for (int d = 0; d < D; d++) {
                    const T lo = (cq.c[d] - cq.h[d]) - (ct.c[d] + ct.h[d]);
                    const T hi = (cq.c[d] + cq.h[d]) - (ct.c[d] - ct.h[d]);
                    wrap |= !(lo >= -half && hi < half) ? 1 : 0;
                  }
                }
                item = ((uint64_t)__float_as_uint(md) << 32) | ((uint64_t)wrap << 31) | (uint32_t)lt;
              }
              const uint64_t mask = __ballot(keep);
              if (keep) s_list[cnt + __popcll(mask & lt_mask)] = item;
              cnt += __popcll(mask);
            }
            if (full) break;
            cbo = 0;
          }
          if (!full && more) {
            eb += 64;
            ii = 0;
            more = eb < e1;
          }
        }
        __syncthreads();
        wave_sort(s_list, cnt);
        consume_list<K, PER, D>(best, q, s_list, cnt, s_rec, l2p, x, src_ids, boxsize, half, row, bcnt);
        __syncthreads();
        if (!full) break;
      }

      if (wave_any(bcnt > 0)) topk_flush<K>(best, row, bcnt);
      const bool mine = valid && (!only_retry || retry[iq] != 0);
      uint64_t kth_key = best[0];
  #pragma unroll
      for (int i = 0; i < K; i++) kth_key = (i == ktot - 1) ? best[i] : kth_key;
      const bool short_list = kth_key == empty && r2_try != INFINITY;
      if (only_retry && wave_any(mine && short_list)) continue;
      if (mine) {
        if (retry && !only_retry) retry[iq] = short_list ? 1 : 0;
        if (!short_list) {
          const int64_t row = (int64_t)(dst_rows ? dst_rows[iq] : iq) * ktot;
  #pragma unroll
          for (int i = 0; i < K; i++) {
            if (i < ktot) {
              rad_out[row + i] = (T)sqrtf(key_r2(best[i]));
              id_out[row + i] = (int32_t)(uint32_t)best[i];
            }
          }
        }
      }
      break;
      }
    }
  };
  if (!only_retry) {
    for (int lq = blockIdx.x; lq < nact_; lq += gridDim.x) {
      const int q0 = l2p[lq], q1 = l2p[lq + 1];
      for (int qoff = q0; qoff < q1; qoff += 64) run_chunk(lq, qoff);
    }
    return;
  }
  // Retry pass: the flags are scanned 64 queries at a time (coalesced), and only the leaf chunks
  // holding a flagged query run.  A chunk that straddles two scanned blocks may run twice; it then
  // writes the same rows twice.
  for (int64_t qb = (int64_t)blockIdx.x * 64; qb < npart; qb += (int64_t)gridDim.x * 64) {
    uint64_t m = __ballot(qb + lane < npart && retry[qb + lane] != 0);
    while (m) {
      const int64_t qf = qb + __builtin_ctzll(m);
      int64_t lo = 0, hi = nact_;  // the leaf of qf: the last lq < nact_ with l2p[lq] <= qf
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (l2p[mid] <= qf) lo = mid;
        else hi = mid;
      }
      const int lq = (int)lo, q0 = l2p[lq], q1 = l2p[lq + 1];
      const int qoff = q0 + (int)((qf - q0) / 64) * 64;
      run_chunk(lq, qoff);
      // the chunk's queries: [qoff, min(qoff + 64, q1)) -- the leaf may end inside the block
      const int64_t b0 = max((int64_t)qoff - qb, (int64_t)0);
      const int64_t b1 = min((int64_t)min(qoff + 64, q1) - qb, (int64_t)64);
      const uint64_t done = (b1 >= 64 ? ~0ull : ((1ull << b1) - 1ull)) & ~((1ull << b0) - 1ull);
      m &= ~done;
    }
  }
}

// passes: bit 0 = bounded pass (flags short queries when init_r2 is given), bit 1 = retry pass
// (recomputes the queries flagged in `retry`, from retry_r2 when given, else without a bound)
template <int D, typename T>
int fused_impl(hipStream_t s, const int32_t* pil_spl, const int32_t* pil_src, const float* pil_r2, int64_t pil_size,
               const int32_t* n2c, const int32_t* leaf_parent, const void* leaf_nodes, const int32_t* l2p,
               const void* x, const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2, int32_t* retry,
               const float* retry_r2, void* rad, int32_t* ids, int64_t nleaves, int64_t npart, int k, float boxsize,
               int passes) {
  if (nleaves <= 0) return 0;
  if (init_r2 && !retry) return fail("KnnFused: a bound needs the retry buffer");
  dim3 g(grid_for(nleaves)), b(64);
  for (int pass = 0; pass < 2; pass++) {
    if (!(passes & (1 << pass))) continue;
    const int only = pass;
    const dim3 gp = pass == 0 ? g : dim3(grid_for(div_up(npart, 64)));  // the retry pass scans the queries
#define JZT_FUSED(KK, PP)                                                                                        \
  hipLaunchKernelGGL((k_knn_fused<KK, PP, D, T>), gp, b, 0, s, pil_spl, pil_src, pil_r2, pil_size, n2c, leaf_parent, \
                     (const T*)leaf_nodes, l2p, (const T*)x, src_ids, dst_rows, init_r2, retry, only, retry_r2, (T*)rad,  \
                     ids, k, (T)boxsize, nleaves, npart)
#define JZT_FUSED_K(PP)              \
  do {                               \
    if (k <= 8) JZT_FUSED(8, PP);    \
    else if (k <= 16) JZT_FUSED(16, PP); \
    else if (k <= 32) JZT_FUSED(32, PP); \
    else JZT_FUSED(64, PP);          \
  } while (0)
    if (boxsize > 0.f) JZT_FUSED_K(true);
    else JZT_FUSED_K(false);
#undef JZT_FUSED_K
#undef JZT_FUSED
    if (check_launch("KnnFused")) return 1;
  }
  return 0;
}

template <typename... A>
int fused_dispatch(int dim, int dtype, A... a) {
  if (dtype == JZT_F32 && dim == 3) return fused_impl<3, float>(a...);
  if (dtype == JZT_F32 && dim == 2) return fused_impl<2, float>(a...);
  return fail("KnnFused: unsupported dim/dtype");
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_knn_fused(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                             const float* parent_ilist_r2, int64_t parent_ilist_size, const int32_t* parent_child_spl,
                             const int32_t* leaf_parent, const void* leaf_nodes, const int32_t* leaf_spl,
                             const void* xz, const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2,
                             int32_t* retry, void* knn_rad, int32_t* knn_id, int64_t size_leaves, int64_t size_part,
                             int k, float boxsize, int dim, int dtype) {
  if (k < 1 || k > 64) return fail("KnnFused: k must be in [1, 64]");
  return fused_dispatch(dim, dtype, (hipStream_t)stream, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2,
                        parent_ilist_size, parent_child_spl, leaf_parent, leaf_nodes, leaf_spl, xz, src_ids, dst_rows,
                        init_r2, retry, (const float*)nullptr, knn_rad, knn_id, size_leaves, size_part, k, boxsize,
                        init_r2 ? 3 : 1);
}

extern "C" int jzt_knn_fused_retry(jzt_stream_t stream, const int32_t* parent_ilist_spl,
                                   const int32_t* parent_ilist_isrc, const float* parent_ilist_r2,
                                   int64_t parent_ilist_size, const int32_t* parent_child_spl,
                                   const int32_t* leaf_parent, const void* leaf_nodes, const int32_t* leaf_spl,
                                   const void* xz, const int32_t* src_ids, const int32_t* dst_rows,
                                   const float* bound_r2, const int32_t* retry, void* knn_rad, int32_t* knn_id,
                                   int64_t size_leaves, int64_t size_part, int k, float boxsize, int dim, int dtype) {
  if (k < 1 || k > 64) return fail("KnnFused: k must be in [1, 64]");
  if (!retry) return fail("KnnFused: retry flags required");
  return fused_dispatch(dim, dtype, (hipStream_t)stream, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2,
                        parent_ilist_size, parent_child_spl, leaf_parent, leaf_nodes, leaf_spl, xz, src_ids, dst_rows,
                        (const float*)nullptr, const_cast<int32_t*>(retry), bound_r2, knn_rad, knn_id, size_leaves, size_part,
                        k, boxsize, 2);
}
