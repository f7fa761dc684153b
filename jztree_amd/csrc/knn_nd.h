// KnnLeaf2Leaf for any dimension 2-10 and float or double positions (knn.cuh:203-356, instantiated
// by the reference for float/double x dims 2-10, _generate_ffi.py:9-21).  float32 positions in 2
// and 3 dimensions take the packed-key kernel of knn.hip; this one serves the rest.
//
// One wave64 per query leaf, one query per lane.  Each lane keeps its k best (r2, id) pairs
// ascending in registers -- r2 in the position type, id as u32 -- ordered by (canonical r2, id)
// like the packed kernel (ties by id, so the result is unique).  The leaf's interaction list is
// walked in ascending list r2 and stops once no lane could still accept an entry (r2T > kth *
// (1 + 2^-16) on every lane).  Source particles move through LDS in chunks of 64; a staged
// particle is kept only if its gap to the query chunk's bounding box is within the largest current
// k-th distance of the wave (or a periodic wrap is possible), and the kept ones are broadcast to
// all lanes from LDS.  Exactness of the filter: by monotone rounding, every canonical r2 from the
// box is >= the gap^2 computed with the same FMA chain.
//
// Query leaves are taken from a per-XCD item queue (xcd_queue) when the caller's tmp holds its
// counters, else in the fixed xcd_items order.
//
// k above the per-pass slot count KMAX (64 for float, 32 for double) runs ceil(k / KMAX) passes;
// pass p > 0 admits only pairs above the last pair of pass p-1 (kept per query in tmp: r2 array,
// then u32 id array).  Optional start bound init_r2 per query leaf (single pass): slots start at
// (r2_0, 2^32-1), a leaf with r2_0 < 0 is skipped, and a query that ends with fewer than k pairs
// within r2_0 is flagged in retry (the same contract as the packed kernel).
#pragma once
#include "knn_common.h"

namespace jzt {

template <typename T> __device__ __forceinline__ T wave_min_t(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
template <typename T> __device__ __forceinline__ T wave_max_t(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// (r2a, ia) < (r2b, ib) in the (r2, id) order; false whenever r2a is NaN
template <typename T> __device__ __forceinline__ bool pair_less(T r2a, uint32_t ia, T r2b, uint32_t ib) {
  return r2a < r2b || (r2a == r2b && ia < ib);
}

// Insert (r2, id) into the ascending pairs (caller guarantees it is below the last one); top-down
// compare/select network with statically indexed registers, leaving every 8 slots once no lane
// still shifts.
template <int K, typename T>
__device__ __forceinline__ void pair_insert(T (&br)[K], uint32_t (&bi)[K], T r2, uint32_t id) {
  bool hi = true;  // (r2, id) < slot i
#pragma unroll
  for (int i = K - 1; i >= 1; i--) {
    const bool lo = pair_less(r2, id, br[i - 1], bi[i - 1]);
    const T tr = hi ? r2 : br[i];
    const uint32_t ti = hi ? id : bi[i];
    br[i] = lo ? br[i - 1] : tr;
    bi[i] = lo ? bi[i - 1] : ti;
    hi = lo;
    if ((i & 7) == 0 && !wave_any(lo)) return;
  }
  if (hi) {
    br[0] = r2;
    bi[0] = id;
  }
}

template <int K, bool PER, int D, typename T>
__global__ __launch_bounds__(64) void k_knn_leaf_nd(const int32_t* __restrict__ il_spl, const int32_t* __restrict__ il_src,
                                                    const float* __restrict__ il_r2, const int32_t* __restrict__ splT,
                                                    const T* __restrict__ xT, const int32_t* __restrict__ splQ,
                                                    const T* __restrict__ xQ, const int32_t* __restrict__ src_ids,
                                                    const int32_t* __restrict__ dst_rows,
                                                    const float* __restrict__ init_r2, int32_t* __restrict__ retry,
                                                    T* __restrict__ rad_out, int32_t* __restrict__ id_out,
                                                    T* __restrict__ res_r2, uint32_t* __restrict__ res_id, int koff,
                                                    int ktot, float boxsize, int64_t il_size, int64_t out_rows,
                                                    int64_t nitems, int xcd, uint32_t* __restrict__ queue) {
  __shared__ Vec<D, T> s_p[64];
  __shared__ uint32_t s_id[64];
  const int lane = threadIdx.x;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const T box = (T)boxsize, half = (T)(0.5f * boxsize);
  const T slack = (T)(1.0f + kBoundSlack);
  const bool floor = res_r2 != nullptr && koff > 0;
  const int64_t nact = active_items(splQ, nitems);
  const ItemRange items = xcd_items(nact, xcd != 0);
  const ItemQueue iq_ = xcd_queue(nact, xcd != 0, queue);
  const int64_t lq_end = queue ? iq_.end : items.end;
  for (int64_t lq = queue ? queue_next(iq_) : items.first; lq < lq_end;
       lq = queue ? queue_next(iq_) : lq + items.step) {
    const int q0 = splQ[lq], q1 = splQ[lq + 1];
    const int e0 = il_spl[lq], e1 = (int)min((int64_t)il_spl[lq + 1], il_size);
    // a negative (or NaN) start bound marks a leaf that this launch leaves untouched
    const float r2_0 = init_r2 ? init_r2[lq] : 0.f;
    if (init_r2 && !(r2_0 >= 0.f)) continue;
    const T start_r2 = init_r2 ? (T)r2_0 : (T)INFINITY;
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int iq = min(qoff + lane, q1 - 1);
      const Vec<D, T> q = load_vec<D, T>(xQ, iq);
      T fr2 = (T)0;
      uint32_t fid = 0;
      if (floor) {
        fr2 = res_r2[iq];
        fid = res_id[iq];
      }
      Vec<D, T> qlo, qhi;  // bounding box of the query chunk
#pragma unroll
      for (int d = 0; d < D; d++) {
        qlo[d] = wave_min_t(q[d]);
        qhi[d] = wave_max_t(q[d]);
      }
      T br[K];
      uint32_t bi[K];
#pragma unroll
      for (int i = 0; i < K; i++) {
        br[i] = start_r2;
        bi[i] = 0xffffffffu;
      }
      T kmax = wave_max_t(br[K - 1]) * slack;
      bool done = false;
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
        for (int i = 0; i < ne; i++) {
          const float r2t = rdlane(m_r2, i);
          if (!wave_any((T)r2t <= br[K - 1] * slack)) {
            done = true;
            break;
          }
          const int t0 = rdlane(m_t0, i), t1 = rdlane(m_t1, i);
          for (int toff = t0; toff < t1; toff += 64) {
            const int cn = min(64, t1 - toff);
            bool need = false, reach = false;
            Vec<D, T> p;
            uint32_t pid = 0;
            if (lane < cn) {
              p = load_vec<D, T>(xT, toff + lane);
              pid = src_ids ? (uint32_t)src_ids[toff + lane] : (uint32_t)(toff + lane);
              T g2 = (T)0;
#pragma unroll
              for (int d = 0; d < D; d++) {
                const T dlo = qlo[d] - p[d], dhi = qhi[d] - p[d];
                if (PER) need |= !(dlo >= -half) || !(dhi < half);
                const T lo = dlo > -dhi ? dlo : -dhi;
                const T gap = lo > (T)0 ? lo : (T)0;
                g2 = fmaT(gap, gap, g2);
              }
              reach = need || g2 <= kmax;
            }
            const uint64_t keep = __ballot(reach);
            const bool wrap = PER && wave_any(need);
            const int n_keep = __popcll(keep);
            if (reach) {
              const int slot = __popcll(keep & lt_mask);
              s_p[slot] = p;
              s_id[slot] = pid;
            }
            __syncthreads();
            bool inserted = false;
            for (int j = 0; j < n_keep; j++) {
              const Vec<D, T> s = s_p[j];
              const uint32_t sid = s_id[j];
              const T r2 = (PER && wrap) ? dist2P<true, D, T>(q, s, box, half) : dist2P<false, D, T>(q, s, box, half);
              const bool acc = pair_less(r2, sid, br[K - 1], bi[K - 1]) && (!floor || pair_less(fr2, fid, r2, sid));
              if (wave_any(acc)) {
                inserted = true;
                if (acc) pair_insert<K, T>(br, bi, r2, sid);
              }
            }
            if (inserted) kmax = wave_max_t(br[K - 1]) * slack;
            __syncthreads();
          }
        }
      }
      if (qoff + lane < q1) {
        const int nk = min(K, ktot - koff);
        const int64_t orow = dst_rows ? (int64_t)dst_rows[iq] : (int64_t)iq;
        if (orow >= 0 && orow < out_rows) {
          const int64_t o = orow * ktot + koff;
#pragma unroll
          for (int i = 0; i < K; i++) {
            if (i < nk) {
              rad_out[o + i] = sqrt(br[i]);
              id_out[o + i] = (int32_t)bi[i];
            }
          }
        }
        // the last pair this pass reports
        T lr = br[0];
        uint32_t li = bi[0];
#pragma unroll
        for (int i = 0; i < K; i++) {
          lr = (i == nk - 1) ? br[i] : lr;
          li = (i == nk - 1) ? bi[i] : li;
        }
        if (retry) retry[iq] = (init_r2 && li == 0xffffffffu && lr == start_r2) ? 1 : 0;
        if (res_r2) {
          res_r2[iq] = lr;
          res_id[iq] = li;
        }
      }
    }
  }
}

// per-query tmp of the multi-pass search: (r2, id) of the last pair of the previous pass
inline size_t knn_leaf_nd_tmp_bytes(int64_t npart_q, int k, bool f64) {
  const int kmax = f64 ? 32 : 64;
  return k > kmax ? (size_t)npart_q * (f64 ? 12 : 8) : 0;
}

// host side: passes of up to KMAX slots over the whole query set (see the file comment)
template <int D, typename T>
int knn_leaf_nd_impl(hipStream_t s, const int32_t* il_spl, const int32_t* il_src, const float* il_r2,
                     const int32_t* splT, const void* xT, const int32_t* splQ, const void* xQ,
                     const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2, int32_t* retry,
                     void* rad, int32_t* ids, void* tmp, size_t tmp_bytes, int64_t nleaves_q, int64_t npart_q,
                     int64_t out_rows, int64_t il_size, int k, float boxsize) {
  constexpr int KMAX = sizeof(T) == 4 ? 64 : 32;
  if (nleaves_q <= 0) return 0;
  if (init_r2 && k > KMAX) return fail("KnnLeaf2Leaf: a start bound needs k <= " + std::to_string(KMAX));
  T* res_r2 = k > KMAX ? (T*)tmp : nullptr;
  uint32_t* res_id = k > KMAX ? (uint32_t*)((char*)tmp + (size_t)npart_q * sizeof(T)) : nullptr;
  // item-queue counters after the resume pairs, when the caller's tmp holds them
  const size_t qoff = align256(knn_leaf_nd_tmp_bytes(npart_q, k, sizeof(T) == 8));
  uint32_t* queue = tmp && tmp_bytes >= qoff + 256 ? (uint32_t*)((char*)tmp + qoff) : nullptr;
  const dim3 g(grid_for8(nleaves_q)), b(64);
  const int xcd = tune_get(kTuneXcd);
  for (int koff = 0; koff < k; koff += KMAX) {
    const int kn = min(k - koff, KMAX);
    if (queue && fill_u32(s, queue, 0u, 8)) return 1;
#define JZT_ND_LAUNCH(KK, PER)                                                                                     \
  hipLaunchKernelGGL((k_knn_leaf_nd<KK, PER, D, T>), g, b, 0, s, il_spl, il_src, il_r2, splT, (const T*)xT, splQ, \
                     (const T*)xQ, src_ids, dst_rows, init_r2, retry, (T*)rad, ids, res_r2, res_id, koff, k,       \
                     boxsize, il_size, out_rows, nleaves_q, xcd, queue)
#define JZT_ND_LAUNCH_K(KK)                   \
  do {                                        \
    if (boxsize > 0.f) JZT_ND_LAUNCH(KK, true); \
    else JZT_ND_LAUNCH(KK, false);            \
  } while (0)
    if (kn <= 8) JZT_ND_LAUNCH_K(8);
    else if (kn <= 16) JZT_ND_LAUNCH_K(16);
    else if (kn <= 32) JZT_ND_LAUNCH_K(32);
    else JZT_ND_LAUNCH_K(KMAX);
#undef JZT_ND_LAUNCH_K
#undef JZT_ND_LAUNCH
    if (check_launch("KnnLeaf2Leaf")) return 1;
  }
  return 0;
}

// the instantiations (knn_nd_f32.hip: dims 4-10, knn_nd_f64.hip: dims 2-10)
#define JZT_KNN_ND_ARGS                                                                                          \
  hipStream_t s, const int32_t *il_spl, const int32_t *il_src, const float *il_r2, const int32_t *splT,          \
      const void *xT, const int32_t *splQ, const void *xQ, const int32_t *src_ids, const int32_t *dst_rows,      \
      const float *init_r2, int32_t *retry, void *rad, int32_t *ids, void *tmp, size_t tmp_bytes,               \
      int64_t nleaves_q, int64_t npart_q, int64_t out_rows, int64_t il_size, int k, float boxsize, int dim
int knn_leaf_nd_f32(JZT_KNN_ND_ARGS);
int knn_leaf_nd_f64(JZT_KNN_ND_ARGS);


}  // namespace jzt
