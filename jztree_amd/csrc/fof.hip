// Friends-of-friends (replaces jztree's fof.cuh).
//
// Union-find over int32 labels with parent[x] <= x.  Hooking is lock-free with atomicMin: to join
// roots hi > lo, atomicMin(&parent[hi], lo) returns the old parent; if hi was no longer a root the
// displaced link (old) is re-joined with lo, so no link is lost (the reference retries a CAS,
// fof.cuh:22-36; both leave the group minimum as the root).  Pointer walks use ordinary (cached)
// loads: a stale value is still an ancestor (parents only decrease), and every retry restarts from
// the parent the atomic returned, so the larger root of the pair strictly decreases and the loop
// ends.  Contraction = pointer jumping to the root.
//
//  NodeToChildLabel  fof.cuh:42-71     children of an unlinked, too-large node label themselves,
//                                      all others inherit the first child of the parent's root
//  FofNode2Node      fof.cuh:73-284    link guaranteed node pairs, contract, count + scan + insert
//                                      the uncertain pairs into the child-level interaction list
//  FofLeaf2Leaf      fof.cuh:290-379   link particle pairs with canonical r2 <= r2link, contract
//
// Node-level decisions use the same 2^-16 relative slack as the kNN bounds: a pair is only linked
// as "guaranteed" if every particle pair certainly passes the particle predicate, and only dropped
// if none can.  Groups are therefore exactly the connected components of the particle predicate.
// One wave64 workgroup per parent node / query leaf; source nodes and particles staged in LDS.
#include "util.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr float kSlackHi = 1.0f + kBoundSlack;  // inflate for "may link" tests
constexpr float kSlackLo = 1.0f - kBoundSlack;  // deflate for "must link" tests

__device__ __forceinline__ int uf_parent(const int32_t* p, int x) { return p[x]; }

__device__ __forceinline__ int uf_find(const int32_t* p, int x) {
  while (true) {
    int q = uf_parent(p, x);
    if (q == x) return x;
    x = q;
  }
}

// Roots of the sets of a and b, following both parent chains in lock step so that their loads
// overlap (each step waits for one memory round trip, not two).
__device__ __forceinline__ void uf_find2(const int32_t* p, int& a, int& b) {
  while (true) {
    const int pa = uf_parent(p, a), pb = uf_parent(p, b);
    if (pa == a && pb == b) return;
    a = pa;
    b = pb;
  }
}

// Joins the sets of a and b (any members of them); returns a member of the joined set (the
// smaller index at the successful hook).  Parents always have smaller indices than their children
// (min-hooking), so the larger of a and b is hooked under the smaller at once, without finding
// roots first: if it was a root, that is the whole link (one atomic); if not, its old parent's set
// still has to be joined with the smaller one's, which the loop does from their roots.  Every set
// that was connected stays connected, so the components are those of the finds-first version.
__device__ __forceinline__ int uf_link(int32_t* p, int a, int b) {
  while (a != b) {
    int hi = a > b ? a : b, lo = a > b ? b : a;
    int old = atomicMin(p + hi, lo);
    if (old == hi || old == lo) return lo;  // hooked (or already joined to lo)
    // hi had been hooked to `old` meanwhile; join old's tree with lo and go on
    a = old;
    b = lo;
    uf_find2(p, a, b);
  }
  return a;
}

__global__ void k_contract(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = uf_find(p, (int)i);
}

template <int D, typename T> struct NodeX {
  Vec<D, T> c;
  Vec<D, T> h;  // half extent
};

template <int D, typename T>
__device__ __forceinline__ NodeX<D, T> load_nodex(const T* __restrict__ nodes, int64_t i) {
  NodeX<D, T> n;
#pragma unroll
  for (int d = 0; d < D; d++) n.c[d] = nodes[i * (D + 1) + d];
  T lv = nodes[i * (D + 1) + D];
  int level;
  if constexpr (sizeof(T) == 4) level = __float_as_int(lv);
  else level = (int)__double_as_longlong(lv);
  n.h = lvl_half_ext<D, T>(level);
  return n;
}

template <int D, typename T> __device__ __forceinline__ T norm2(const Vec<D, T>& v) {
  T r = v[0] * v[0];
#pragma unroll
  for (int d = 1; d < D; d++) r = fmaT(v[d], v[d], r);
  return r;
}

// ------------------------------------------------------------------------------------------------
template <int D, typename T>
__global__ __launch_bounds__(64) void k_node_to_child(const int32_t* __restrict__ pig, const uint8_t* __restrict__ local,
                                                      const int32_t* __restrict__ plvl, const int32_t* __restrict__ pspl,
                                                      int32_t* __restrict__ cig, int64_t size_parent, float r2link,
                                                      const float* __restrict__ r2v, int64_t nitems) {
  const int64_t nact_ = active_items(pspl, nitems);
  for (int64_t node = blockIdx.x; node < nact_; node += gridDim.x) {
    if (node >= size_parent || (local && !local[node])) continue;
    const int root = pig[node];
    if (root >= size_parent || (local && !local[root])) continue;
    const T L2 = norm2<D, T>(lvl_ext<D, T>(plvl[node]));
    const float r2 = r2v ? r2v[node] : r2link;
    const bool self = (root == node) && (L2 * (T)kSlackHi > (T)r2);
    const int c0 = pspl[node], c1 = pspl[node + 1];
    const int lab = pspl[root];
    for (int i = c0 + threadIdx.x; i < c1; i += 64) cig[i] = self ? i : lab;
  
  }
}

// PASS 0: link guaranteed pairs; PASS 1: count uncertain pairs; PASS 2: insert them.
// One wave per parent node; its children are staged in LDS, and for every entry of the parent's list
// (each parent pair once) the entry's children too.  Lanes take (child, target) pairs: with nt
// targets staged, lane l tests child c0 + l / nt against target l % nt (one lane per child would
// leave most lanes idle: nodes have about coarse_fac children).  A child's hits in one step are
// lanes of one contiguous block: counted with a ballot and, in the insert pass, placed at the
// child's running offset plus the rank of the hit in its block.  Lists are sets (the leaf kernel and
// the next plane visit every entry), and union-find with min-hooking reaches the same roots in any
// link order, so the labels do not depend on the order inside a list or of the links.
template <int PASS, int D, typename T>
__global__ __launch_bounds__(64) void k_fof_node(const int32_t* __restrict__ pil_spl, const int32_t* __restrict__ pil,
                                                 const int32_t* __restrict__ spl, const T* __restrict__ nodes,
                                                 const int32_t* __restrict__ il_spl, int32_t* ig, int32_t* __restrict__ count,
                                                 int32_t* __restrict__ il, float r2link, const float* __restrict__ r2v,
                                                 T boxsize, int64_t il_size, int64_t pil_size, int64_t nitems) {
  __shared__ NodeX<D, T> s_q[64], s_n[64];
  __shared__ int s_qg[64], s_g[64];
  __shared__ int64_t s_run[64];  // per child: hits so far (count pass) or next list slot (insert pass)
  const int lane = threadIdx.x;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const int64_t nact_ = active_items(spl, nitems);
  for (int pq = blockIdx.x; pq < nact_; pq += gridDim.x) {
    const int q0 = spl[pq], q1 = spl[pq + 1];
    const int e0 = pil_spl[pq], e1 = (int)min((int64_t)pil_spl[pq + 1], pil_size);
    const float r2 = r2v ? r2v[pq] : r2link;
    const T rl_hi = (T)r2 * (T)kSlackHi, rl_lo = (T)r2 * (T)kSlackLo;
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int nc = min(q1 - qoff, 64);
      if (lane < nc) {
        s_q[lane] = load_nodex<D, T>(nodes, qoff + lane);
        s_qg[lane] = ig[qoff + lane];
        s_run[lane] = (PASS == 2) ? (int64_t)il_spl[qoff + lane] : 0;
      }
      __syncthreads();
      for (int e = e0; e < e1; e++) {
        const int pt = pil[e];
        if (pt < pq) continue;  // each parent pair once
        const int t0 = spl[pt], t1 = spl[pt + 1];
        for (int toff = t0; toff < t1; toff += 64) {
          const int nt = min(t1 - toff, 64);
          if (lane < nt) {
            s_n[lane] = load_nodex<D, T>(nodes, toff + lane);
            s_g[lane] = ig[toff + lane];
          }
          __syncthreads();
          const int cstep = nt >= 32 ? 1 : 64 / nt;
          const int my_c = lane / nt, my_t = lane - my_c * nt;
          const uint64_t block = (nt == 64 ? ~0ull : ((1ull << nt) - 1ull)) << (my_c * nt);
          for (int c0 = 0; c0 < nc; c0 += cstep) {
            const int c = c0 + my_c;
            const bool active = my_c < cstep && c < nc;
            const int iq = qoff + c, jt = toff + my_t;
            bool hit = false;
            if (active && !(s_qg[c] == s_g[my_t] && iq != jt)) {  // equal labels: already linked
              const NodeX<D, T> nq = s_q[c], nt_ = s_n[my_t];
              Vec<D, T> w;
#pragma unroll
              for (int d = 0; d < D; d++) w[d] = nq.h[d] + nt_.h[d];
              const T r2max = maxdist2<D, T>(nt_.c, nq.c, w, boxsize);
              const T r2min = mindist2<D, T>(nt_.c, nq.c, w, boxsize);
              const T L2 = norm2<D, T>(nt_.h) * T(4);
              if (max(r2max, L2) <= rl_lo) {
                if (PASS == 0) uf_link(ig, iq, jt);
              } else if (PASS >= 1 && r2min <= rl_hi) {
                hit = true;
              }
            }
            if (PASS >= 1) {
              const uint64_t hits = __ballot(hit) & block;
              if (PASS == 2 && hit) {
                const int64_t o = s_run[c] + __popcll(hits & lt_mask);
                if (o < il_size) il[o] = jt;
              }
              if (active && my_t == 0 && hits) s_run[c] += __popcll(hits);
            }
          }
          __syncthreads();
        }
      }
      if (PASS == 1 && lane < nc) count[qoff + lane] = (int32_t)s_run[lane];
      __syncthreads();
    }
  }
}

template <int D, typename T> struct alignas(sizeof(T) == 4 ? 16 : 8) FofRec {
  T p[D];
  int32_t g;
};

template <typename T> __device__ __forceinline__ T wave_min_t(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T> __device__ __forceinline__ T wave_max_t(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Particle links inside each 64-particle chunk of a leaf, resolved before the leaf pairs (the
// first pass of FofLeaf2Leaf).  Per chunk: the within-rlink graph of its <= 64 particles (the
// same canonical r2 and test as every other pair), its connected components by bitmask closure,
// and then every particle that labels itself is pointed at the first particle of its component
// (a plain store: nothing else points at it, and only this wave writes it), while a particle
// carrying a label from the node levels is joined with that first particle by one union-find link
// per distinct (label, component) pair.  Afterwards the members of a chunk component share one
// parent, so the leaf-pair pass sees equal labels for them and links each pair of sets once.
template <bool PER, int D, typename T>
__global__ __launch_bounds__(64) void k_fof_chunks(const int32_t* __restrict__ spl, const T* __restrict__ pos,
                                                   int32_t* ig, float r2link_all, const float* __restrict__ r2v,
                                                   T boxsize, int64_t nitems, int xcd) {
  __shared__ FofRec<D, T> s_r[64];
  const int lane = threadIdx.x;
  const ItemRange items = xcd_items(active_items(spl, nitems), xcd != 0);
  for (int64_t lq = items.first; lq < items.end; lq += items.step) {
    const int q0 = spl[lq], q1 = spl[lq + 1];
    const T half = (T)(float)(T(0.5) * boxsize);
    const float r2link = r2v ? r2v[lq] : r2link_all;
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int n = min(64, q1 - qoff);
      const int iq = min(qoff + lane, q1 - 1);
      const bool ok = lane < n;
      const Vec<D, T> x = load_vec<D, T>(pos, iq);
      const int g = ig[iq];
      // one label on the whole chunk (a leaf linked on the node levels, or a single particle):
      // its particles are in one set already
      const int g0 = __shfl(g, 0, 64);
      if (__ballot(ok && g != g0) == 0ull) continue;
      bool need_wrap = false;
      Vec<D, T> lo, hi;
#pragma unroll
      for (int d = 0; d < D; d++) {
        lo[d] = wave_min_t<T>(x[d]);
        hi[d] = wave_max_t<T>(x[d]);
      }
      if (ok) {
#pragma unroll
        for (int d = 0; d < D; d++)
          if (PER) need_wrap |= !(lo[d] - x[d] >= -half) || !(hi[d] - x[d] < half);
        FofRec<D, T> r;
#pragma unroll
        for (int d = 0; d < D; d++) r.p[d] = x[d];
        r.g = g;
        s_r[lane] = r;
      }
      const bool wrap = PER && wave_any(need_wrap);
      // A chunk whose bounding box has a diagonal within rlink is one component: every pair's
      // canonical r2 is at most the box diagonal's (each |fl(q - p)| <= fl(hi - lo) by monotone
      // rounding, and the FMA chain is monotone in each difference), so no pair test is needed.
      const bool one_component = !wrap && dist2P<false, D, T>(hi, lo, boxsize, half) <= (T)r2link;
      int lroot = lane;
      if (one_component) {
        lroot = 0;
      } else {
        __syncthreads();
        uint64_t nb = 0ull;
        // four staged records per step (their LDS reads in flight together); the periodic and
        // plain distance loops are separate
        auto pairs = [&](auto wrap_c) {
          constexpr bool W = decltype(wrap_c)::value;
          for (int j0 = 0; j0 < n; j0 += 4) {
            FofRec<D, T> rr[4];
#pragma unroll
            for (int u = 0; u < 4; u++) rr[u] = s_r[min(j0 + u, n - 1)];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              Vec<D, T> xp;
#pragma unroll
              for (int d = 0; d < D; d++) xp[d] = rr[u].p[d];
              const int j = j0 + u;
              const T r2 = dist2P<W, D, T>(x, xp, boxsize, half);
              const bool h = (j < n) & (j != lane) & (r2 <= (T)r2link);
              nb |= (uint64_t)h << (j & 63);
            }
          }
        };
        if (ok) {
          if (wrap) pairs(std::integral_constant<bool, true>());
          else pairs(std::integral_constant<bool, false>());
        }
        // Connected components by breadth-first search over the neighbour rows, one component at a
        // time from its lowest particle (wave-uniform: a frontier's rows are read with readlane
        // and or-ed in scalar registers), so every particle's row is read once.  Only particles
        // with a neighbour take part; the root of a component is its lowest particle.
        uint64_t todo = __ballot(nb != 0ull);
        while (todo) {
          const int root = __builtin_ctzll(todo);
          uint64_t comp = 1ull << root, frontier = comp;
          while (frontier) {
            uint64_t next = 0ull;
            for (uint64_t f = frontier; f; f &= f - 1ull) {
              const int j = __builtin_ctzll(f);
              next |= ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(nb >> 32), j) << 32) |
                      (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)nb, j);
            }
            frontier = next & ~comp;
            comp |= next;
          }
          if ((comp >> lane) & 1ull) lroot = root;
          todo &= ~comp;
        }
      }
      __syncthreads();
      const int groot = __shfl(g, lroot, 64);  // label of the component's first particle
      if (ok && lroot != lane && g == iq) ig[iq] = groot;  // self-labelled: point at the component
      bool want = ok && lroot != lane && g != iq && g != groot;
      uint64_t wm = __ballot(want);
      while (wm) {
        const int L = __builtin_ctzll(wm);
        const int gL = __builtin_amdgcn_readlane(g, L), rL = __builtin_amdgcn_readlane(groot, L);
        const bool mine = want && g == gL && groot == rL;
        if (lane == L) uf_link(ig, gL, rL);
        want = want && !mine;
        wm &= ~__ballot(mine);
      }
    }
  }
}

// One wave per query leaf; each leaf pair once (source index >= query index).  Source particles
// are staged 64 at a time through LDS as (position, label) records.  Before staging, each lane
// tests its source particle p against the bounding box [qlo, qhi] of the query chunk:
//   - wrap: some difference fl(q - p), q in the box, may leave [-half, half) in some dim (by
//     monotone rounding it suffices to test fl(qlo - p) and fl(qhi - p)); only then does the chunk
//     use the periodic wrap (the distances are bit-identical either way);
//   - reach: p is within r2link * (1 + 2^-16) of the box (or needs the wrap).  A p that fails
//     cannot link with any query of the chunk, because the box gap bounds every |q - p| per dim
//     from below (the slack covers the rounding of the gap and of the canonical distance).
// Only reachable sources are visited; a chunk with none is skipped.
// Links are not made where they are found: the (query label, source label) pair goes into a queue
// in LDS, and the queue is worked off by all 64 lanes at once (one link per lane, their memory
// round trips overlapping) when it is half full and at the end.  Each lane remembers the labels
// its set was queued with, so the dedup tests stay valid: a queued pair is joined before the
// kernel ends, and the components are those of any order of the same links.  (Should a step find
// the queue full, its leader lane links at once.)
template <bool PER, int D, typename T>
__global__ __launch_bounds__(64) void k_fof_leaf(const int32_t* __restrict__ il_spl, const int32_t* __restrict__ il,
                                                 const int32_t* __restrict__ spl, const T* __restrict__ pos, int32_t* ig,
                                                 float r2link_all, const float* __restrict__ r2v, T boxsize,
                                                 int64_t il_size, int64_t nitems, int xcd, const T* __restrict__ cells) {
  constexpr int kQueue = 256;
  __shared__ FofRec<D, T> s_r[64];
  __shared__ int2 s_q[kQueue];
  const int lane = threadIdx.x;
  int qn = 0;  // queued pairs (wave-uniform)
  // every lane active: the queued links, one per lane
  auto flush = [&]() {
    __syncthreads();
    for (int i = lane; i < qn; i += 64) {
      const int2 pr = s_q[i];
      uf_link(ig, pr.x, pr.y);
    }
    qn = 0;
    __syncthreads();
  };
  const ItemRange items = xcd_items(active_items(spl, nitems), xcd != 0);
  for (int64_t lq = items.first; lq < items.end; lq += items.step) {
    const int q0 = spl[lq], q1 = spl[lq + 1];
    const int e0 = il_spl[lq], e1 = (int)min((int64_t)il_spl[lq + 1], il_size);
    const T half = (T)(float)(T(0.5) * boxsize);
    const float r2link = r2v ? r2v[lq] : r2link_all;
    const T r2skip = (T)(r2link * (1.f + kBoundSlack));
    for (int qoff = q0; qoff < q1; qoff += 64) {
      const int iq = min(qoff + lane, q1 - 1);
      const bool valid = qoff + lane < q1;
      const Vec<D, T> xq = load_vec<D, T>(pos, iq);
      Vec<D, T> qlo, qhi;
  #pragma unroll
      for (int d = 0; d < D; d++) {
        qlo[d] = wave_min_t<T>(xq[d]);
        qhi[d] = wave_max_t<T>(xq[d]);
      }
      // labels known to be in q's set: its own label and the last two labels it was queued with
      int gq = ig[iq];
      int glast = -1, glast2 = -1;
      for (int eb = e0; eb < e1; eb += 64) {
        // list entries 64 at a time, one per lane: the source leaf and its particle range (each
        // leaf pair once; a leaf with itself only from its chunk after this one, a chunk's own
        // links being k_fof_chunks'), and with the leaves' cells a test of the cell against the
        // query chunk's box: the cell (widened by 2^-20 of its coordinates against the rounding
        // of centre +- half extent) bounds every |fl(q - p)| from below by monotone rounding, so
        // an entry whose gap is beyond reach is dropped (only where no periodic wrap is needed)
        const int ie = eb + lane;
        int m0 = 0, m1 = 0;
        bool keep = false;
        if (ie < e1) {
          const int lt = il[ie];
          if (lt >= lq) {
            m0 = lt == lq ? qoff + 64 : spl[lt];
            m1 = spl[lt + 1];
            keep = m0 < m1;
            if (keep && cells && lt != lq) {
              const NodeX<D, T> nx = load_nodex<D, T>(cells, lt);
              bool wrap_c = false;
              T g2 = T(0);
#pragma unroll
              for (int d = 0; d < D; d++) {
                const T eps = (abs(nx.c[d]) + nx.h[d]) * T(1.0 / 1048576.0);
                const T lo = nx.c[d] - nx.h[d] - eps, hi = nx.c[d] + nx.h[d] + eps;
                if (PER) wrap_c |= !(qlo[d] - hi >= -half) || !(qhi[d] - lo < half);
                const T gap = max(T(0), max(lo - qhi[d], qlo[d] - hi));
                g2 = fmaT(gap, gap, g2);
              }
              keep = wrap_c || g2 <= r2skip;
            }
          }
        }
        for (uint64_t em = __ballot(keep); em; em &= em - 1ull) {
        const int E = __builtin_ctzll(em);
        const int t0 = __builtin_amdgcn_readlane(m0, E), t1 = __builtin_amdgcn_readlane(m1, E);
        for (int toff = t0; toff < t1; toff += 64) {
          const int it = toff + lane;
          FofRec<D, T> r;
          bool need_wrap = false, reach = false;
          if (it < t1) {
            const Vec<D, T> p = load_vec<D, T>(pos, it);
            T g2 = T(0);
  #pragma unroll
            for (int d = 0; d < D; d++) {
              r.p[d] = p[d];
              const T dlo = qlo[d] - p[d], dhi = qhi[d] - p[d];
              if (PER) need_wrap |= !(dlo >= -half) || !(dhi < half);
              const T gap = max(T(0), max(dlo, -dhi));
              g2 = fmaT(gap, gap, g2);
            }
            r.g = ig[it];
            reach = need_wrap || g2 <= r2skip;
          }
          uint64_t rem = __ballot(reach);
          if (rem == 0ull) continue;
          const bool wrap = PER && wave_any(need_wrap);
          if (qn > kQueue / 2) flush();
          s_r[lane] = r;
          __syncthreads();
          // The reachable sources label by label.  Labels are snapshots of set members: a query
          // whose own label, or one of the last two labels it was queued with, equals the label
          // needs no test of its sources.  For a label held by several sources, the others test
          // them two at a time until each has found one within rlink (one link per (query set,
          // label) is all it takes); labels held by a single source are collected and their
          // sources tested afterwards, two at a time.
          auto by_label = [&](auto wrap_c) {
            constexpr bool W = decltype(wrap_c)::value;
            auto r2_of = [&](int j, int& gj) {
              const FofRec<D, T> rj = s_r[j];
              gj = rj.g;
              Vec<D, T> xp;
#pragma unroll
              for (int d = 0; d < D; d++) xp[d] = rj.p[d];
              return dist2P<W, D, T>(xq, xp, boxsize, half);
            };
            // the next four set bits of a source mask (repeating the first where fewer are left)
            auto take4 = [](uint64_t& ms, int (&j)[4]) {
              j[0] = __builtin_ctzll(ms);
              ms &= ms - 1ull;
#pragma unroll
              for (int u = 1; u < 4; u++) {
                j[u] = ms ? __builtin_ctzll(ms) : j[0];
                ms &= ms - 1ull;
              }
            };
            auto knows = [&](int g) { return !valid || g == gq || g == glast || g == glast2; };
            // Lanes whose q holds the same label are in one set: a single link of that set with
            // the label's joins them all, so one pair is queued for them.
            auto queue_link = [&](bool want, int G) {
              uint64_t wm = __ballot(want);
              while (wm) {
                const int L = __builtin_ctzll(wm);
                const bool mine = want && gq == __builtin_amdgcn_readlane(gq, L);
                const uint64_t mm = __ballot(mine);
                if (qn < kQueue) {
                  if (lane == L) s_q[qn] = make_int2(gq, G);
                  qn++;
                } else if (lane == L) {
                  uf_link(ig, gq, G);
                }
                if (mine) {
                  glast2 = glast;
                  glast = G;
                  want = false;
                }
                wm &= ~mm;
              }
            };
            uint64_t single = 0ull;
            while (rem) {
              const int jg = __builtin_ctzll(rem);
              const int G = __builtin_amdgcn_readlane(r.g, jg);
              uint64_t ms = __ballot(reach && r.g == G);
              rem &= ~ms;
              if (ms == (1ull << jg)) {
                single |= ms;
                continue;
              }
              const bool need = !knows(G);
              if (!wave_any(need)) continue;
              // the lane's nearest source of the label so far, four sources per step
              T best = (T)__builtin_inff();
              while (ms) {
                int jj[4], gj;
                take4(ms, jj);
                T r2[4];
#pragma unroll
                for (int u = 0; u < 4; u++) r2[u] = r2_of(jj[u], gj);
                best = min(best, min(min(r2[0], r2[1]), min(r2[2], r2[3])));
                if (!wave_any(need && !(best <= (T)r2link))) break;
              }
              queue_link(need && best <= (T)r2link, G);
            }
            // single-source labels, four per step: distinct labels, so a link for one leaves the
            // others' tests as they were
            while (single) {
              const int n4 = min(4, __builtin_popcountll(single));
              int jj[4], gj[4];
              take4(single, jj);
              bool w[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const T r2 = r2_of(jj[u], gj[u]);
                w[u] = r2 <= (T)r2link && !knows(gj[u]);
              }
#pragma unroll
              for (int u = 0; u < 4; u++)
                if (u < n4) queue_link(w[u], gj[u]);
            }
          };
          if (wrap) by_label(std::integral_constant<bool, true>());
          else by_label(std::integral_constant<bool, false>());
          __syncthreads();
        }
        }
      }
    }
  }
  flush();
}

template <int D, typename T>
int node2child_impl(hipStream_t s, const int32_t* pig, const uint8_t* local, const int32_t* plvl, const int32_t* pspl,
                    int32_t* cig, int64_t size_parent, int64_t size_child, float r2link, const float* r2v) {
  if (fill_u32(s, cig, 0u, size_child)) return 1;
  if (size_parent <= 0) return 0;
  hipLaunchKernelGGL((k_node_to_child<D, T>), dim3(grid_for(size_parent)), dim3(64), 0, s, pig, local, plvl, pspl, cig,
                     size_parent, r2link, r2v, size_parent);
  return check_launch("NodeToChildLabel");
}

template <int D, typename T>
int fof_node_impl(hipStream_t s, const int32_t* pil_spl, const int32_t* pil, const int32_t* pspl, const void* nodes,
                  const int32_t* ig_in, int32_t* ig, int32_t* il_spl, int32_t* il, int64_t size_parent, int64_t size_node,
                  int64_t pil_size, int64_t il_size, float r2link, const float* r2v, float boxsize) {
  const T* nd = (const T*)nodes;
  if (fill_u32(s, il_spl, 0u, size_node + 1)) return 1;
  if (fill_u32(s, il, 0u, il_size)) return 1;
  if (ig != ig_in && hipMemcpyAsync(ig, ig_in, (size_t)size_node * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail("FofNode2Node: copy failed");
  if (size_parent > 0) {
    dim3 g(grid_for(size_parent)), b(64);
    hipLaunchKernelGGL((k_fof_node<0, D, T>), g, b, 0, s, pil_spl, pil, pspl, nd, (const int32_t*)nullptr, ig,
                       (int32_t*)nullptr, (int32_t*)nullptr, r2link, r2v, (T)boxsize, il_size, pil_size, size_parent);
    hipLaunchKernelGGL(k_contract, dim3((unsigned)div_up(size_node, 256)), dim3(256), 0, s, ig, size_node);
    hipLaunchKernelGGL((k_fof_node<1, D, T>), g, b, 0, s, pil_spl, pil, pspl, nd, (const int32_t*)nullptr, ig,
                       il_spl + 1, (int32_t*)nullptr, r2link, r2v, (T)boxsize, il_size, pil_size, size_parent);
    if (check_launch("FofNode2Node link/count")) return 1;
  }
  if (scan_tmp_words(size_node) > il_size) return fail("FofNode2Node: interaction list too small for scan temp");
  if (scan_i32(s, il_spl + 1, il_spl + 1, size_node, true, il)) return 1;
  if (fill_u32(s, il, 0u, scan_tmp_words(size_node))) return 1;
  if (size_parent > 0) {
    hipLaunchKernelGGL((k_fof_node<2, D, T>), dim3(grid_for(size_parent)), dim3(64), 0, s, pil_spl, pil, pspl, nd,
                       il_spl, ig, (int32_t*)nullptr, il, r2link, r2v, (T)boxsize, il_size, pil_size, size_parent);
  }
  return check_launch("FofNode2Node insert");
}

template <int D, typename T>
int fof_leaf_impl(hipStream_t s, const int32_t* il_spl, const int32_t* il, int64_t il_size, const int32_t* spl,
                  const void* pos, const int32_t* ig_in, int32_t* ig, int64_t size_leaves, int64_t size_part,
                  float r2link, const float* r2v, float boxsize, const void* cells) {
  if (ig != ig_in && hipMemcpyAsync(ig, ig_in, (size_t)size_part * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail("FofLeaf2Leaf: copy failed");
  if (size_leaves > 0) {
    const int xcd = tune_get(kTuneXcd);
    if (boxsize > 0.f)
      hipLaunchKernelGGL((k_fof_chunks<true, D, T>), dim3(grid_for8(size_leaves)), dim3(64), 0, s, spl, (const T*)pos,
                         ig, r2link, r2v, (T)boxsize, size_leaves, xcd);
    else
      hipLaunchKernelGGL((k_fof_chunks<false, D, T>), dim3(grid_for8(size_leaves)), dim3(64), 0, s, spl, (const T*)pos,
                         ig, r2link, r2v, (T)boxsize, size_leaves, xcd);
    if (boxsize > 0.f)
      hipLaunchKernelGGL((k_fof_leaf<true, D, T>), dim3(grid_for8(size_leaves)), dim3(64), 0, s, il_spl, il, spl,
                         (const T*)pos, ig, r2link, r2v, (T)boxsize, il_size, size_leaves, xcd, (const T*)cells);
    else
      hipLaunchKernelGGL((k_fof_leaf<false, D, T>), dim3(grid_for8(size_leaves)), dim3(64), 0, s, il_spl, il, spl,
                         (const T*)pos, ig, r2link, r2v, (T)boxsize, il_size, size_leaves, xcd, (const T*)cells);
  }
  if (size_part > 0)
    hipLaunchKernelGGL(k_contract, dim3((unsigned)div_up(size_part, 256)), dim3(256), 0, s, ig, size_part);
  return check_launch("FofLeaf2Leaf");
}

}  // namespace
}  // namespace jzt

using namespace jzt;

#define JZT_DISPATCH2(fn, ...)                                            \
  do {                                                                    \
    if (dtype == JZT_F32 && dim == 3) return fn<3, float>(__VA_ARGS__);   \
    if (dtype == JZT_F32 && dim == 2) return fn<2, float>(__VA_ARGS__);   \
    if (dtype == JZT_F64 && dim == 3) return fn<3, double>(__VA_ARGS__);  \
    if (dtype == JZT_F64 && dim == 2) return fn<2, double>(__VA_ARGS__);  \
    return fail(std::string(#fn) + ": unsupported dim/dtype");            \
  } while (0)

extern "C" int jzt_node_to_child_label(jzt_stream_t stream, const int32_t* parent_igroup, const uint8_t* parent_is_local,
                                       const int32_t* parent_lvl, const int32_t* parent_spl, int32_t* node_igroup,
                                       int64_t size_parent, int64_t size_child, float r2link, int dim, int dtype) {
  JZT_DISPATCH2(node2child_impl, (hipStream_t)stream, parent_igroup, parent_is_local, parent_lvl, parent_spl, node_igroup,
                size_parent, size_child, r2link, (const float*)nullptr);
}

extern "C" int jzt_node_to_child_label_sets(jzt_stream_t stream, const int32_t* parent_igroup, const int32_t* parent_lvl,
                                            const int32_t* parent_spl, const float* parent_r2link, int32_t* node_igroup,
                                            int64_t size_parent, int64_t size_child, int dim, int dtype) {
  if (parent_r2link == nullptr) return fail("NodeToChildLabel (sets): parent_r2link is required");
  JZT_DISPATCH2(node2child_impl, (hipStream_t)stream, parent_igroup, (const uint8_t*)nullptr, parent_lvl, parent_spl,
                node_igroup, size_parent, size_child, 0.f, parent_r2link);
}

extern "C" int jzt_fof_node2node(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist,
                                 const int32_t* parent_spl, const void* nodes, const int32_t* node_igroup_in,
                                 int32_t* node_igroup, int32_t* node_ilist_spl, int32_t* node_ilist, int64_t size_parent,
                                 int64_t size_node, int64_t parent_ilist_size, int64_t size_node_ilist, float r2link,
                                 float boxsize, int dim, int dtype) {
  JZT_DISPATCH2(fof_node_impl, (hipStream_t)stream, parent_ilist_spl, parent_ilist, parent_spl, nodes, node_igroup_in,
                node_igroup, node_ilist_spl, node_ilist, size_parent, size_node, parent_ilist_size, size_node_ilist,
                r2link, (const float*)nullptr, boxsize);
}

extern "C" int jzt_fof_node2node_sets(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist,
                                      const int32_t* parent_spl, const float* parent_r2link, const void* nodes,
                                      const int32_t* node_igroup_in, int32_t* node_igroup, int32_t* node_ilist_spl,
                                      int32_t* node_ilist, int64_t size_parent, int64_t size_node,
                                      int64_t parent_ilist_size, int64_t size_node_ilist, float boxsize, int dim,
                                      int dtype) {
  if (parent_r2link == nullptr) return fail("FofNode2Node (sets): parent_r2link is required");
  JZT_DISPATCH2(fof_node_impl, (hipStream_t)stream, parent_ilist_spl, parent_ilist, parent_spl, nodes, node_igroup_in,
                node_igroup, node_ilist_spl, node_ilist, size_parent, size_node, parent_ilist_size, size_node_ilist,
                0.f, parent_r2link, boxsize);
}

extern "C" int jzt_fof_leaf2leaf(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist,
                                 int64_t ilist_size, const int32_t* spl, const void* pos, const int32_t* part_igroup_in,
                                 int32_t* part_igroup, int64_t size_leaves, int64_t size_part, float r2link,
                                 float boxsize, int dim, int dtype) {
  JZT_DISPATCH2(fof_leaf_impl, (hipStream_t)stream, ilist_spl, ilist, ilist_size, spl, pos, part_igroup_in, part_igroup,
                size_leaves, size_part, r2link, (const float*)nullptr, boxsize, (const void*)nullptr);
}

extern "C" int jzt_fof_leaf2leaf_cells(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist,
                                       int64_t ilist_size, const int32_t* spl, const void* pos, const void* leaf_cells,
                                       const int32_t* part_igroup_in, int32_t* part_igroup, int64_t size_leaves,
                                       int64_t size_part, float r2link, float boxsize, int dim, int dtype) {
  if (leaf_cells == nullptr) return fail("FofLeaf2Leaf (cells): leaf_cells is required");
  JZT_DISPATCH2(fof_leaf_impl, (hipStream_t)stream, ilist_spl, ilist, ilist_size, spl, pos, part_igroup_in, part_igroup,
                size_leaves, size_part, r2link, (const float*)nullptr, boxsize, leaf_cells);
}

extern "C" int jzt_fof_leaf2leaf_sets(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist,
                                      int64_t ilist_size, const int32_t* spl, const float* leaf_r2link, const void* pos,
                                      const int32_t* part_igroup_in, int32_t* part_igroup, int64_t size_leaves,
                                      int64_t size_part, float boxsize, int dim, int dtype) {
  if (leaf_r2link == nullptr) return fail("FofLeaf2Leaf (sets): leaf_r2link is required");
  JZT_DISPATCH2(fof_leaf_impl, (hipStream_t)stream, ilist_spl, ilist, ilist_size, spl, pos, part_igroup_in, part_igroup,
                size_leaves, size_part, 0.f, leaf_r2link, boxsize, (const void*)nullptr);
}
