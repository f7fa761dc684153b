// Leaf-plane kNN over query groups (extension): the queries of a plane-1 node in chunks of 64.
//
// KnnLeaf2Leaf (knn.cuh:309-356) runs one query leaf per wave.  Leaves are maximal Morton cells of
// at most `leaf size` particles and hold about 60% of that on clustered points, so a wave of 64
// lanes leaves many lanes idle.  Here a wave takes 64 consecutive queries of one plane-1 node P
// (they span several of P's child leaves) and walks P's r2-sorted list of plane-1 nodes T_p:
//   - stop when no lane can still accept:  r2(P, T_p) > kth * (1 + 2^-16) on every lane;
//   - for each child leaf T of T_p: skip it unless mindist2(box of the chunk, T) <= the wave's
//     largest kth * (1 + 2^-16); the chunk box is widened by 2^-20 of its coordinates, so rounding
//     of its centre and half extents can only make the test pass more often;
//   - stream T's particles 64 at a time: each lane tests one source against the chunk box (exact
//     reach filter and wrap test of the leaf kernel), the reachable ones are compacted into LDS
//     and consumed by every lane into its buffered top-k.
// Children of different list entries are disjoint, so every source is seen at most once.  The
// top-k starts from a bound r2_0 per group (the node's guaranteed radius, or a proposed one);
// queries left with fewer than k neighbours inside it are flagged in `retry`.
// Exactness: every particle within a query's final k-th r2 lies in a child of a listed node
// before the stop point (lists hold every node within r2_0, the stop test is conservative by the
// slack), in a child that is not skipped (the widened box and the slack make the skip test
// conservative), and passes the reach filter (exact by monotone rounding), so the top-k equals
// the brute-force (r2, id) top-k.
#include "util.h"
#include "knn_common.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr float kGroupSlack = 1.0f + kBoundSlack;

#ifndef JZT_GROUP_WAVES
#define JZT_GROUP_WAVES 4
#endif

template <int K, bool PER, int D, int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) void k_knn_group(
    const int32_t* __restrict__ pil_spl, const int32_t* __restrict__ pil_src, const float* __restrict__ pil_r2,
    int64_t pil_size, const int32_t* __restrict__ n2c, const float* __restrict__ leaf_nodes,
    const int32_t* __restrict__ l2p, const float* __restrict__ x, const int32_t* __restrict__ src_ids,
    const int32_t* __restrict__ dst_rows, const float* __restrict__ init_r2, int32_t* __restrict__ retry,
    float* __restrict__ rad_out, int32_t* __restrict__ id_out, int ktot, float boxsize, int64_t out_rows,
    int64_t nitems) {
  __shared__ SrcRec s_rec[64];
  __shared__ uint64_t s_buf[kBufWords];
  const int lane = threadIdx.x;
  uint64_t* const row = s_buf + lane * kBufStride;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const float half = 0.5f * boxsize;
  const int64_t nact_ = active_items(n2c, nitems);
  for (int pq = blockIdx.x; pq < nact_; pq += gridDim.x) {
    const int lc0 = n2c[pq], lc1 = n2c[pq + 1];
    if (lc1 <= lc0) continue;
    const int p0 = l2p[lc0], p1 = l2p[lc1];
    const float r2_0 = init_r2 ? init_r2[pq] : INFINITY;
    if (init_r2 && !(r2_0 >= 0.f)) continue;  // negative bound: this launch leaves the group alone
    const uint64_t start_key = ((uint64_t)__float_as_uint(r2_0) << 32) | 0xffffffffull;
    const int e0 = pil_spl[pq], e1 = (int)min((int64_t)pil_spl[pq + 1], pil_size);
    for (int qoff = p0; qoff < p1; qoff += 64) {
      const int iq = min(qoff + lane, p1 - 1);
      const Vec<D, float> q = load_vec<D, float>(x, iq);
      Vec<D, float> qlo, qhi, qc, qh;  // chunk box; centre and widened half extents for the skip test
#pragma unroll
      for (int d = 0; d < D; d++) {
        qlo[d] = wave_minf(q[d]);
        qhi[d] = wave_maxf(q[d]);
        qc[d] = 0.5f * (qlo[d] + qhi[d]);
        qh[d] = 0.5f * (qhi[d] - qlo[d]) * (1.0f + 0x1p-20f) + fmaxf(fabsf(qlo[d]), fabsf(qhi[d])) * 0x1p-20f;
      }
      uint64_t best[K];
#pragma unroll
      for (int i = 0; i < K; i++) best[i] = start_key;
      int cnt = 0;
      float kmax = wave_maxf(key_r2(best[K - 1])) * kGroupSlack;
      bool done = false;
      for (int eb = e0; eb < e1 && !done; eb += 64) {
        const int ne = min(64, e1 - eb);
        int m_c0 = 0, m_c1 = 0;
        float m_r2 = INFINITY;
        if (lane < ne) {
          const int src = pil_src[eb + lane];
          m_r2 = pil_r2[eb + lane];
          m_c0 = n2c[src];
          m_c1 = n2c[src + 1];
        }
        for (int i = 0; i < ne; i++) {
          if (!(rdlane(m_r2, i) <= kmax)) {  // r2-sorted: no later node is reachable either
            done = true;
            break;
          }
          const int c0 = rdlane(m_c0, i), c1 = rdlane(m_c1, i);
          for (int cb = c0; cb < c1; cb += 64) {
            bool keep = false;
            if (cb + lane < c1) {
              const int lt = cb + lane;
              Vec<D, float> tc, w;
#pragma unroll
              for (int d = 0; d < D; d++) tc[d] = leaf_nodes[(int64_t)lt * (D + 1) + d];
              const Vec<D, float> th = lvl_half_ext<D, float>(__float_as_int(leaf_nodes[(int64_t)lt * (D + 1) + D]));
#pragma unroll
              for (int d = 0; d < D; d++) w[d] = qh[d] + th[d];
              const float md = mindist2<D, float>(tc, qc, w, boxsize);
              keep = md <= kmax && l2p[lt + 1] > l2p[lt];
            }
            uint64_t kids = __ballot(keep);
            while (kids) {
              const int j = __builtin_ctzll(kids);
              kids &= kids - 1ull;
              const int lt = cb + j;
              const int t0 = l2p[lt], t1 = l2p[lt + 1];
              for (int toff = t0; toff < t1; toff += 64) {
                const int n_cur = min(64, t1 - toff);
                SrcRec cr;
                bool need = false, reach = false;
                if (lane < n_cur) {
                  const Vec<D, float> p = load_vec<D, float>(x, toff + lane);
                  float g2 = 0.f;
#pragma unroll
                  for (int d = 0; d < D; d++) {
                    cr.p[d] = p[d];
                    const float dlo = qlo[d] - p[d], dhi = qhi[d] - p[d];
                    if (PER) need |= !(dlo >= -half) || !(dhi < half);
                    const float gap = fmaxf(0.f, fmaxf(dlo, -dhi));
                    g2 = fmaf(gap, gap, g2);
                  }
                  cr.id = src_ids ? (uint32_t)src_ids[toff + lane] : (uint32_t)(toff + lane);
                  reach = need || g2 <= kmax;
                }
                const uint64_t keep_src = __ballot(reach);
                if (keep_src == 0ull) continue;
                const bool wrap = PER && wave_any(need);
                if (reach) s_rec[__popcll(keep_src & lt_mask)] = cr;
                __syncthreads();
                if (consume_chunk<K, PER, false, D>(best, q, s_rec, __popcll(keep_src), wrap, boxsize, half, 0ull,
                                                     row, cnt))
                  kmax = wave_maxf(key_r2(best[K - 1])) * kGroupSlack;
                __syncthreads();
              }
            }
          }
        }
      }
      if (wave_any(cnt > 0)) topk_flush<K>(best, row, cnt);
      if (qoff + lane < p1) {
        const int64_t orow = dst_rows ? (int64_t)dst_rows[iq] : (int64_t)iq;
        if (orow >= 0 && orow < out_rows) write_row<K>(best, rad_out, id_out, orow * ktot, ktot);
        if (retry) {  // the k-th slot, not the template's last (k need not be a power of two)
          uint64_t lk = best[0];
  #pragma unroll
          for (int i = 0; i < K; i++) lk = (i == ktot - 1) ? best[i] : lk;
          retry[iq] = (lk == start_key && init_r2) ? 1 : 0;
        }
      }
    }
  }
}

template <int D>
int group_impl(hipStream_t s, const int32_t* pil_spl, const int32_t* pil_src, const float* pil_r2, int64_t pil_size,
               const int32_t* n2c, const float* leaf_nodes, const int32_t* l2p, const float* x,
               const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2, int32_t* retry, float* rad,
               int32_t* ids, int64_t ngroups, int64_t out_rows, int k, float boxsize) {
  if (ngroups <= 0) return 0;
  if (k < 1 || k > 64) return fail("KnnGroup: 1 <= k <= 64");
  dim3 g(grid_for(ngroups)), b(64);
  const int waves = tune_get(kTuneLeafWaves);
#define JZT_GROUP_LAUNCH(KK, PER, W)                                                                                 \
  hipLaunchKernelGGL((k_knn_group<KK, PER, D, W>), g, b, 0, s, pil_spl, pil_src, pil_r2, pil_size, n2c, leaf_nodes, \
                     l2p, x, src_ids, dst_rows, init_r2, retry, rad, ids, k, boxsize, out_rows, ngroups)
#define JZT_GROUP_LAUNCH_W(KK, PER)                                               \
  do {                                                                            \
    constexpr int w2 = KK > 32 ? 1 : 2, w4 = KK > 32 ? 1 : JZT_GROUP_WAVES;       \
    if (KK > 32 || waves == 1) JZT_GROUP_LAUNCH(KK, PER, 1);                      \
    else if (waves == 2) JZT_GROUP_LAUNCH(KK, PER, w2);                           \
    else JZT_GROUP_LAUNCH(KK, PER, w4);                                           \
  } while (0)
#define JZT_GROUP_LAUNCH_K(KK)                       \
  do {                                               \
    if (boxsize > 0.f) JZT_GROUP_LAUNCH_W(KK, true); \
    else JZT_GROUP_LAUNCH_W(KK, false);              \
  } while (0)
  if (k <= 8) JZT_GROUP_LAUNCH_K(8);
  else if (k <= 16) JZT_GROUP_LAUNCH_K(16);
  else if (k <= 32) JZT_GROUP_LAUNCH_K(32);
  else JZT_GROUP_LAUNCH_K(64);
#undef JZT_GROUP_LAUNCH_K
#undef JZT_GROUP_LAUNCH_W
#undef JZT_GROUP_LAUNCH
  return check_launch("KnnGroup");
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_knn_group(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                             const float* parent_ilist_r2, int64_t parent_ilist_size, const int32_t* parent_to_leaf,
                             const void* leaf_nodes, const int32_t* leaf_to_part, const void* pos,
                             const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2, int32_t* retry,
                             void* knn_rad, int32_t* knn_id, int64_t ngroups, int64_t out_rows, int k,
                             float boxsize, int dim, int dtype) {
  hipStream_t s = (hipStream_t)stream;
  if (dtype != JZT_F32) return fail("KnnGroup: float32 positions only");
  if (init_r2 && !retry) return fail("KnnGroup: a start bound needs a retry buffer");
  if (dim == 3)
    return group_impl<3>(s, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2, parent_ilist_size, parent_to_leaf,
                         (const float*)leaf_nodes, leaf_to_part, (const float*)pos, src_ids, dst_rows, init_r2, retry,
                         (float*)knn_rad, knn_id, ngroups, out_rows, k, boxsize);
  if (dim == 2)
    return group_impl<2>(s, parent_ilist_spl, parent_ilist_isrc, parent_ilist_r2, parent_ilist_size, parent_to_leaf,
                         (const float*)leaf_nodes, leaf_to_part, (const float*)pos, src_ids, dst_rows, init_r2, retry,
                         (float*)knn_rad, knn_id, ngroups, out_rows, k, boxsize);
  return fail("KnnGroup: dim must be 2 or 3");
}
