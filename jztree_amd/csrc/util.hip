// Device-wide scan / fill / compaction and the library's error state.
#include <mutex>
#include <string>

#include "util.h"

#include <vector>
#include "../../include/jztree_amd.h"

namespace jzt {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}
int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

int fill_u32(hipStream_t s, void* p, uint32_t v, int64_t n) {
  if (n <= 0) return 0;
  int64_t blocks = div_up(n, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, v, n);
  return check_launch("fill_u32");
}

// pass 1: per-tile sums
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                              int32_t* __restrict__ tile_sums) {
  __shared__ int sbuf[kScanThreads / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  int v = 0;
#pragma unroll
  for (int r = 0; r < kScanItems; r++) {
    int64_t i = base + (int64_t)r * kScanThreads + threadIdx.x;
    v += (i < n) ? in[i] : 0;
  }
  int tot;
  block_excl_scan<kScanThreads>(v, sbuf, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// pass 2: exclusive scan of the tile sums by one workgroup, in chunks with a running carry
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int32_t* __restrict__ tile_sums, int64_t ntiles) {
  __shared__ int sbuf[kScanThreads / 64];
  int carry = 0;
  for (int64_t c = 0; c < ntiles; c += kScanThreads) {
    int64_t i = c + threadIdx.x;
    int v = (i < ntiles) ? tile_sums[i] : 0;
    int tot;
    int ex = block_excl_scan<kScanThreads>(v, sbuf, &tot);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

// pass 3: scan within each tile (blocked per thread) and add the tile offset
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int32_t* in, int32_t* out, int64_t n,
                                                             const int32_t* __restrict__ tile_off, int inclusive) {
  __shared__ int sbuf[kScanThreads / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int x[kScanItems];
  int v = 0;
#pragma unroll
  for (int r = 0; r < kScanItems; r++) {
    int64_t i = base + r;
    x[r] = (i < n) ? in[i] : 0;
    v += x[r];
  }
  int tot;
  int ex = block_excl_scan<kScanThreads>(v, sbuf, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int r = 0; r < kScanItems; r++) {
    int64_t i = base + r;
    int nv = ex + x[r];
    if (i < n) out[i] = inclusive ? nv : ex;
    ex = nv;
  }
}

int scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, bool inclusive, int32_t* tmp) {
  if (n <= 0) return 0;
  int64_t nt = div_up(n, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, tmp, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, out, n, tmp,
                     inclusive ? 1 : 0);
  return check_launch("scan_i32");
}

// Compaction of int8 flags: where(flags != 0) in index order, as two passes over the flags (the
// flags are read as 16-byte vectors, 16 per thread, kWhereTile per workgroup): per-tile counts,
// their exclusive scan, then each tile writes its indices at its offset plus the in-tile rank.
constexpr int kWherePer = 16;
constexpr int kWhereTile = kScanThreads * kWherePer;  // 4096

__device__ __forceinline__ uint32_t load_flags16(const int8_t* __restrict__ f, int64_t i0, int64_t n, bool vec) {
  // bit k set iff flag i0 + k is nonzero (flags past n read as zero); vec: f is 16-byte aligned
  uint32_t bits = 0;
  if (vec && i0 + kWherePer <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(f + i0);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int b = 0; b < 4; b++) bits |= (((w[q] >> (8 * b)) & 0xffu) != 0u ? 1u : 0u) << (4 * q + b);
  } else {
    for (int k = 0; k < kWherePer && i0 + k < n; k++) bits |= (f[i0 + k] != 0 ? 1u : 0u) << k;
  }
  return bits;
}

__global__ __launch_bounds__(kScanThreads) void k_where_count(const int8_t* __restrict__ f, int64_t n,
                                                              int32_t* __restrict__ tile_counts) {
  __shared__ int sbuf[kScanThreads / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kWhereTile + (int64_t)threadIdx.x * kWherePer;
  const int c = __builtin_popcount(load_flags16(f, i0, n, ((uintptr_t)f & 15u) == 0));
  int tot;
  block_excl_scan<kScanThreads>(c, sbuf, &tot);
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_where_write(const int8_t* __restrict__ f, int64_t n, int64_t ntiles,
                                                              const int32_t* __restrict__ tile_off,
                                                              int32_t* __restrict__ out, int64_t size, int32_t fill,
                                                              const int32_t* fill_dev, int32_t* count_out) {
  __shared__ int sbuf[kScanThreads / 64];
  const int32_t total = tile_off[ntiles];
  if (blockIdx.x < ntiles) {
    const int64_t i0 = (int64_t)blockIdx.x * kWhereTile + (int64_t)threadIdx.x * kWherePer;
    uint32_t bits = load_flags16(f, i0, n, ((uintptr_t)f & 15u) == 0);
    int tot;
    int j = block_excl_scan<kScanThreads>(__builtin_popcount(bits), sbuf, &tot) + tile_off[blockIdx.x];
    for (; bits; bits &= bits - 1u, j++)
      if (j < size) out[j] = (int32_t)(i0 + __builtin_ctz(bits));
  }
  // the slots past the count get the fill value
  if (fill_dev) fill = fill_dev[0];
  const int64_t t0 = (int64_t)blockIdx.x * kWhereTile;
  for (int64_t i = t0 + threadIdx.x; i < t0 + kWhereTile && i < size; i += kScanThreads)
    if (i >= total) out[i] = fill;
  if (blockIdx.x == 0 && threadIdx.x == 0 && count_out) *count_out = total;
}

int where_flags_i8(hipStream_t s, const int8_t* flags, int64_t n, int32_t* out, int64_t size, int32_t fill,
                   const int32_t* fill_dev, int32_t* count_out, int32_t* tmp) {
  if (n <= 0) return fail("where_flags_i8: empty input");
  // tmp: ntiles + 1 words (at most scan_tmp_words(n), what callers provide)
  const int64_t ntiles = div_up(n, kWhereTile);
  hipLaunchKernelGGL(k_where_count, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, flags, n, tmp);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, tmp, ntiles);
  const int64_t grid = ntiles > div_up(size, kWhereTile) ? ntiles : div_up(size, kWhereTile);
  hipLaunchKernelGGL(k_where_write, dim3((unsigned)(grid > 0 ? grid : 1)), dim3(kScanThreads), 0, s, flags, n, ntiles,
                     tmp, out, size, fill, fill_dev, count_out);
  return check_launch("where_flags_i8");
}

// tree.py:868-930 grouped_dense_interaction_list as one device op (the node count lives on the
// device): super-node splits spl_super[i] = min(i*ngroup, nn) and list splits
// ispl[i] = min(i*nsuper, nint) for i <= size_super, entries ilist[j] = j % nsuper (and r2[j] = 0)
// for j < min(nint, size_ilist); entries past the list are not touched (readers stop at ispl).
__global__ void k_dense_ilist(const int32_t* __restrict__ nn_dev, int ngroup, int32_t* __restrict__ spl_super,
                              int32_t* __restrict__ ispl, int64_t size_super, int32_t* __restrict__ ilist,
                              float* __restrict__ r2, int64_t size_ilist, int32_t* __restrict__ nsuper_out) {
  const int64_t nn = nn_dev[0] > 0 ? nn_dev[0] : 0;
  const int64_t nsuper = (nn + ngroup - 1) / ngroup;
  const int64_t nint = nsuper * nsuper;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = tid; i <= size_super; i += st) {
    spl_super[i] = (int32_t)min(i * ngroup, nn);
    ispl[i] = (int32_t)min(min(i * nsuper, nint), (int64_t)INT32_MAX);
  }
  const int64_t nw = min(nint, size_ilist);
  for (int64_t j = tid; j < nw; j += st) {
    ilist[j] = (int32_t)(j % nsuper);
    if (r2) r2[j] = 0.f;
  }
  if (tid == 0) nsuper_out[0] = (int32_t)nsuper;
}

// Dense top-plane lists of a batch: the top nodes of set s (node_set non-decreasing over the first
// nn nodes) form ceil(c_s/ngroup) super-nodes, and a super-node interacts with every super-node
// of its own set and with no other.  tmp: F, S, I (int64, nsets+1 each) = exclusive scans of the
// node counts, super-node counts and list lengths per set.
__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t n, int32_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int64_t last_le_i64(const int64_t* a, int64_t n, int64_t v) {  // last i < n with a[i] <= v
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_dense_sets_prep(const int32_t* __restrict__ nn_dev,
                                                         const int32_t* __restrict__ node_set, int64_t nsets, int ngroup,
                                                         int64_t* __restrict__ F, int64_t* __restrict__ S,
                                                         int64_t* __restrict__ I) {
  __shared__ int64_t s_c[256], s_u[256], s_q[256];
  __shared__ int64_t carry[3];
  const int64_t nn = nn_dev[0] > 0 ? nn_dev[0] : 0;
  if (threadIdx.x == 0) carry[0] = carry[1] = carry[2] = 0;
  __syncthreads();
  for (int64_t b = 0; b < nsets; b += 256) {
    const int64_t s = b + threadIdx.x;
    int64_t c = 0;
    if (s < nsets) c = lower_bound_i32(node_set, nn, (int32_t)(s + 1)) - lower_bound_i32(node_set, nn, (int32_t)s);
    const int64_t u = (c + ngroup - 1) / ngroup;
    s_c[threadIdx.x] = c;
    s_u[threadIdx.x] = u;
    s_q[threadIdx.x] = u * u;
    __syncthreads();
    if (threadIdx.x == 0) {  // serial scan of 256 entries per chunk: nsets is at most a few 10^4
      int64_t c0 = carry[0], c1 = carry[1], c2 = carry[2];
      for (int t = 0; t < 256 && b + t < nsets; t++) {
        F[b + t] = c0; S[b + t] = c1; I[b + t] = c2;
        c0 += s_c[t]; c1 += s_u[t]; c2 += s_q[t];
      }
      carry[0] = c0; carry[1] = c1; carry[2] = c2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    F[nsets] = carry[0];
    S[nsets] = carry[1];
    I[nsets] = carry[2];
  }
}

__global__ void k_dense_sets_fill(const int32_t* __restrict__ nn_dev, int64_t nsets, int ngroup,
                                  const int64_t* __restrict__ F, const int64_t* __restrict__ S,
                                  const int64_t* __restrict__ I, int32_t* __restrict__ spl_super,
                                  int32_t* __restrict__ ispl, int64_t size_super, int32_t* __restrict__ ilist,
                                  float* __restrict__ r2, int64_t size_ilist, int32_t* __restrict__ nsuper_out) {
  const int64_t nn = nn_dev[0] > 0 ? nn_dev[0] : 0;
  const int64_t nsup = S[nsets], nint = I[nsets];
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t a = tid; a <= size_super; a += st) {
    if (a < nsup) {
      const int64_t s = last_le_i64(S, nsets + 1, a);
      const int64_t u = a - S[s], ns = S[s + 1] - S[s];
      spl_super[a] = (int32_t)min(F[s] + u * ngroup, nn);
      ispl[a] = (int32_t)min(I[s] + u * ns, (int64_t)INT32_MAX);
    } else {
      spl_super[a] = (int32_t)nn;
      ispl[a] = (int32_t)min(nint, (int64_t)INT32_MAX);
    }
  }
  const int64_t nw = min(nint, size_ilist);
  for (int64_t j = tid; j < nw; j += st) {
    const int64_t s = last_le_i64(I, nsets + 1, j);
    const int64_t ns = S[s + 1] - S[s];
    ilist[j] = (int32_t)(S[s] + (j - I[s]) % ns);
    if (r2) r2[j] = 0.f;
  }
  if (tid == 0) nsuper_out[0] = (int32_t)nsup;
}

// One plane of a packed per-plane array: out row i = data row (off + i) while off + i < end, else
// every element of the row = *fill.  off, end = ispl[level], ispl[level + 1] (int32 or int64 on the
// device).  Rows are row_words 32-bit words; the fill element is elem_words words.
__global__ void k_packed_get(const uint32_t* __restrict__ data, int64_t nrows, const void* __restrict__ ispl,
                             int ispl64, int level, const uint32_t* __restrict__ fill, int elem_words, int row_words,
                             uint32_t* __restrict__ out, int64_t size) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= size * row_words) return;
  const int64_t i = t / row_words;
  const int w = (int)(t - i * row_words);
  const int64_t off = ispl64 ? ((const int64_t*)ispl)[level] : (int64_t)((const int32_t*)ispl)[level];
  const int64_t end = ispl64 ? ((const int64_t*)ispl)[level + 1] : (int64_t)((const int32_t*)ispl)[level + 1];
  const int64_t r = off + i;
  out[t] = (r < end && r >= 0 && r < nrows) ? data[r * row_words + w] : fill[w % elem_words];
}

// flags[i] |= (lvl[i] >= *lvl_max && i <= *npart), as 0/1 bytes (regularization's extra splits)
__global__ void k_flags_or_level(int8_t* __restrict__ flags, const int32_t* __restrict__ lvl, int64_t n,
                                 const int32_t* __restrict__ lvl_max, const int32_t* __restrict__ npart) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool extra = lvl[i] >= lvl_max[0] && i < (int64_t)npart[0] + 1;
  flags[i] = (flags[i] != 0 || extra) ? 1 : 0;
}

// leaf kNN waves per EU (K <= 32), onesweep keys per thread, XCD-aware order of the leaf kernels
static int g_tune[kTuneCount] = {4, 12, 1};
int tune_get(int knob) { return (knob >= 0 && knob < kTuneCount) ? __atomic_load_n(&g_tune[knob], __ATOMIC_RELAXED) : 0; }

}  // namespace jzt

namespace jzt {
namespace {
// number of rows whose first coordinate is not NaN: per-block counts, one atomic per block
template <typename T>
__global__ __launch_bounds__(256) void k_count_not_nan(const T* __restrict__ pos, int64_t n, int dim,
                                                       uint32_t* __restrict__ out) {
  __shared__ uint32_t s_c;
  if (threadIdx.x == 0) s_c = 0;
  __syncthreads();
  uint32_t c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += __builtin_isnan(pos[i * dim]) ? 0u : 1u;
  for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0 && s_c) atomicAdd(out, s_c);
}
}  // namespace
}  // namespace jzt

extern "C" int jzt_count_not_nan(jzt_stream_t stream, const void* pos, int64_t n, int dim, int dtype, int32_t* out) {
  using namespace jzt;
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || n >= (int64_t)INT32_MAX) return fail("jzt_count_not_nan: n must be in [0, 2^31)");
  if (dim < 1) return fail("jzt_count_not_nan: dim < 1");
  if (fill_u32(s, out, 0u, 1)) return 1;
  if (n == 0) return 0;
  const unsigned g = (unsigned)std::min(div_up(n, 256), (int64_t)1024);
  if (dtype == JZT_F32)
    hipLaunchKernelGGL(k_count_not_nan<float>, dim3(g), dim3(256), 0, s, (const float*)pos, n, dim, (uint32_t*)out);
  else if (dtype == JZT_F64)
    hipLaunchKernelGGL(k_count_not_nan<double>, dim3(g), dim3(256), 0, s, (const double*)pos, n, dim, (uint32_t*)out);
  else
    return fail("jzt_count_not_nan: unsupported dtype");
  return check_launch("jzt_count_not_nan");
}

extern "C" int jzt_set_tuning(int knob, int value) {
  using namespace jzt;
  bool ok = false;
  if (knob == kTuneLeafWaves) ok = value == 1 || value == 2 || value == 4;
  else if (knob == kTuneSortKpt) ok = value == 8 || value == 12 || value == 16;
  else if (knob == kTuneXcd) ok = value == 0 || value == 1;
  else return fail("jzt_set_tuning: unknown knob " + std::to_string(knob));
  if (!ok) return fail("jzt_set_tuning: value " + std::to_string(value) + " not allowed for knob " + std::to_string(knob));
  __atomic_store_n(&g_tune[knob], value, __ATOMIC_RELAXED);
  return 0;
}

extern "C" int jzt_get_tuning(int knob) { return jzt::tune_get(knob); }

extern "C" const char* jzt_last_error(void) { return jzt::g_err.c_str(); }
extern "C" int jzt_abi_version(void) { return JZT_ABI_VERSION; }

extern "C" int jzt_dense_ilist(jzt_stream_t stream, const int32_t* nnodes, int ngroup, int32_t* spl_super,
                               int32_t* ispl, int64_t size_super, int32_t* ilist, float* r2, int64_t size_ilist,
                               int32_t* nsuper) {
  using namespace jzt;
  if (ngroup < 1) return fail("DenseIlist: ngroup must be >= 1");
  const int64_t work = size_super + 1 > size_ilist ? size_super + 1 : size_ilist;
  hipLaunchKernelGGL(k_dense_ilist, dim3(grid_for(div_up(work, 256))), dim3(256), 0, (hipStream_t)stream, nnodes,
                     ngroup, spl_super, ispl, size_super, ilist, r2, size_ilist, nsuper);
  return check_launch("DenseIlist");
}

extern "C" size_t jzt_dense_ilist_sets_tmp_bytes(int64_t nsets) { return (size_t)3 * 8 * (size_t)(nsets + 1); }

extern "C" int jzt_dense_ilist_sets(jzt_stream_t stream, const int32_t* nnodes, const int32_t* node_set,
                                    int64_t nsets, int ngroup, int32_t* spl_super, int32_t* ispl, int64_t size_super,
                                    int32_t* ilist, float* r2, int64_t size_ilist, int32_t* nsuper, void* tmp,
                                    size_t tmp_bytes) {
  using namespace jzt;
  if (ngroup < 1) return fail("DenseIlistSets: ngroup must be >= 1");
  if (nsets < 1 || nsets > INT32_MAX - 1) return fail("DenseIlistSets: need 1 <= nsets < 2^31 - 1");
  if (tmp_bytes < jzt_dense_ilist_sets_tmp_bytes(nsets)) return fail("DenseIlistSets: tmp buffer too small");
  int64_t* F = (int64_t*)tmp;
  int64_t* S = F + (nsets + 1);
  int64_t* I = S + (nsets + 1);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dense_sets_prep, dim3(1), dim3(256), 0, s, nnodes, node_set, nsets, ngroup, F, S, I);
  const int64_t work = size_super + 1 > size_ilist ? size_super + 1 : size_ilist;
  hipLaunchKernelGGL(k_dense_sets_fill, dim3(grid_for(div_up(work, 256))), dim3(256), 0, s, nnodes, nsets, ngroup, F, S,
                     I, spl_super, ispl, size_super, ilist, r2, size_ilist, nsuper);
  return check_launch("DenseIlistSets");
}

extern "C" int jzt_graph_count_nodes(void* graph, int64_t* counts) {
  using namespace jzt;
  if (graph == nullptr || counts == nullptr) return fail("GraphCountNodes: null argument");
  for (int i = 0; i < 4; i++) counts[i] = 0;
  size_t n = 0;
  if (hipGraphGetNodes((hipGraph_t)graph, nullptr, &n) != hipSuccess) return fail("GraphCountNodes: hipGraphGetNodes failed");
  std::vector<hipGraphNode_t> nodes(n);
  if (n && hipGraphGetNodes((hipGraph_t)graph, nodes.data(), &n) != hipSuccess)
    return fail("GraphCountNodes: hipGraphGetNodes failed");
  for (size_t i = 0; i < n; i++) {
    hipGraphNodeType t;
    if (hipGraphNodeGetType(nodes[i], &t) != hipSuccess) return fail("GraphCountNodes: hipGraphNodeGetType failed");
    counts[t == hipGraphNodeTypeKernel ? 0 : t == hipGraphNodeTypeMemcpy ? 1 : t == hipGraphNodeTypeMemset ? 2 : 3]++;
  }
  return 0;
}

extern "C" int jzt_packed_get(jzt_stream_t stream, const void* data, int64_t nrows, const void* ispl, int ispl64,
                              int level, const void* fill, int elem_bytes, int row_elems, void* out, int64_t size) {
  using namespace jzt;
  if (size <= 0) return 0;
  if ((elem_bytes != 4 && elem_bytes != 8) || row_elems < 1) return fail("PackedGet: elements of 4 or 8 bytes");
  if (!data || !ispl || !fill || !out || level < 0) return fail("PackedGet: null argument");
  const int ew = elem_bytes / 4, rw = ew * row_elems;
  const int64_t nt = size * rw;
  hipLaunchKernelGGL(k_packed_get, dim3((unsigned)div_up(nt, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const uint32_t*)data, nrows, ispl, ispl64, level, (const uint32_t*)fill, ew, rw, (uint32_t*)out,
                     size);
  return check_launch("PackedGet");
}

extern "C" int jzt_flags_or_level(jzt_stream_t stream, int8_t* flags, const int32_t* lvl, int64_t n,
                                  const int32_t* lvl_max, const int32_t* npart) {
  using namespace jzt;
  if (n <= 0) return 0;
  if (!flags || !lvl || !lvl_max || !npart) return fail("FlagsOrLevel: null argument");
  hipLaunchKernelGGL(k_flags_or_level, dim3((unsigned)div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, flags, lvl,
                     n, lvl_max, npart);
  return check_launch("FlagsOrLevel");
}
