// Friends-of-friends catalogue: per-group sums over particles in group order (replaces the
// reductions of jztree's _fof_catalogue_from_groups, src/jztree/fof.py:650-766).
//
// Input: particles in group order (every group a contiguous run that starts at its first z-order
// member, see jzt_fof_order) and the group size at the first member of each group (0 elsewhere).
// Groups with at least npart_min members are catalogued, in particle order.
//
//   1. keep flags + stable compaction: offset[g] (start of the g-th kept group; size_part past the
//      last), ngroups, count[g]
//   2. work items: every kept group split into runs of at most kCatChunk particles (exclusive scan
//      of the run counts); one 256-thread workgroup per run reduces it into float64 partial sums
//   3. per group, in item order (deterministic): total mass, centre of mass relative to the first
//      member (periodic differences), mean velocity, mean radial velocity, mean scale factor
//   4. a second pass of runs: mass-weighted squared distance to the centre of mass -> inertia radius
//
// The reference sums in the particle type with scatter-adds in an unspecified order; here each
// product is formed in the particle type exactly as the reference forms it and the sums run in
// float64, so results agree to the particle type's rounding.  Catalogue slots past ngroups hold
// count 0, offset size_part, mass 0 and NaN elsewhere (0/0, as in the reference).
#include "util.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr int kCatChunk = 8192;  // particles per work item
constexpr int kCatThreads = 256;
constexpr int kQ = 9;  // partial sums per item: m, m*dx[3], m*dv[3], m*v_rad, m/zplus1

// jnp.remainder (x - y*floor(x/y) evaluated as rem + sign fix-up, as XLA lowers it)
template <typename T> __device__ __forceinline__ T py_mod(T x, T y) {
  T r = fmod(x, y);
  if (r != T(0) && ((r < T(0)) != (y < T(0)))) r += y;
  return r;
}
// fof.py:715-719: periodic difference and position
template <typename T> __device__ __forceinline__ T wrap_dx(T dx, T box) {
  return box > T(0) ? py_mod<T>(dx + T(0.5) * box, box) - T(0.5) * box : dx;
}
template <typename T> __device__ __forceinline__ T wrap_pos(T x, T box) { return box > T(0) ? py_mod<T>(x, box) : x; }

__global__ void k_cat_keep(const int32_t* __restrict__ counts, int64_t n, int npart_min, int8_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = counts[i] >= npart_min ? 1 : 0;
}

// count[g], and the number of work items of each group (0 past ngroups)
__global__ void k_cat_groups(const int32_t* __restrict__ counts, const int32_t* __restrict__ offset,
                             const int32_t* __restrict__ ngroups, int64_t size_cata, int32_t* __restrict__ count,
                             int32_t* __restrict__ nitems) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= size_cata) return;
  const bool valid = g < (int64_t)ngroups[0];
  const int32_t c = valid ? counts[offset[g]] : 0;
  count[g] = c;
  nitems[g] = (c + kCatChunk - 1) / kCatChunk;
}

struct Item {
  int64_t g, begin, end;
};

__device__ __forceinline__ int64_t total_items(const int32_t* __restrict__ item_off, const int32_t* __restrict__ nit,
                                               int64_t size_cata) {
  return (int64_t)item_off[size_cata - 1] + nit[size_cata - 1];
}

// work item b < total_items: its group (the last g with item_off[g] <= b; kept groups have at
// least one item, the empty slots past ngroups all start at the total) and particle range
__device__ __forceinline__ Item item_of(int64_t b, const int32_t* __restrict__ item_off,
                                        const int32_t* __restrict__ offset, const int32_t* __restrict__ count,
                                        int64_t size_cata) {
  Item it;
  int64_t lo = 0, hi = size_cata - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if ((int64_t)item_off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int64_t s = offset[lo], c = count[lo];
  it.g = lo;
  it.begin = s + (b - item_off[lo]) * (int64_t)kCatChunk;
  it.end = min(s + c, it.begin + (int64_t)kCatChunk);
  return it;
}

__device__ __forceinline__ double block_sum(double v, double* s_red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kCatThreads / 64; i++) t += s_red[i];
  return t;
}

// pass 1: partial sums of one work item
template <int D, typename T>
__global__ __launch_bounds__(kCatThreads) void k_cat_pass1(
    const T* __restrict__ pos, const T* __restrict__ mass, T mass_scalar, const T* __restrict__ vel,
    const T* __restrict__ vrad, const T* __restrict__ zp1, const int32_t* __restrict__ item_off,
    const int32_t* __restrict__ nit, const int32_t* __restrict__ offset, const int32_t* __restrict__ count,
    int64_t size_cata, T box, double* __restrict__ part) {
  __shared__ double s_red[kCatThreads / 64];
  const int64_t nitems = total_items(item_off, nit, size_cata);
  for (int64_t b = blockIdx.x; b < nitems; b += gridDim.x) {
  const Item it = item_of(b, item_off, offset, count, size_cata);
  const int64_t s = offset[it.g];
  T p0[D], v0[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    p0[d] = pos[s * D + d];
    v0[d] = vel ? vel[s * D + d] : T(0);
  }
  double acc[kQ];
#pragma unroll
  for (int q = 0; q < kQ; q++) acc[q] = 0.0;
  for (int64_t i = it.begin + threadIdx.x; i < it.end; i += kCatThreads) {
    const T m = mass ? mass[i] : mass_scalar;
    acc[0] += (double)m;
#pragma unroll
    for (int d = 0; d < D; d++) acc[1 + d] += (double)(wrap_dx<T>(pos[i * D + d] - p0[d], box) * m);
    if (vel) {
#pragma unroll
      for (int d = 0; d < D; d++) acc[4 + d] += (double)((vel[i * D + d] - v0[d]) * m);
    }
    if (vrad) acc[7] += (double)(vrad[i] * m);
    if (zp1) acc[8] += (double)((T(1) / zp1[i]) * m);
  }
#pragma unroll
  for (int q = 0; q < kQ; q++) {
    const double t = block_sum(acc[q], s_red);
    if (threadIdx.x == 0) part[b * kQ + q] = t;
  }
  }
}

// per group: sums over its items in order, then the pass-1 properties
template <int D, typename T>
__global__ void k_cat_finish1(const T* __restrict__ pos, bool mass_array, T mass_scalar, const T* __restrict__ vel,
                              const int32_t* __restrict__ item_off, const int32_t* __restrict__ nit,
                              const int32_t* __restrict__ offset, const int32_t* __restrict__ count,
                              const int32_t* __restrict__ ngroups, int64_t size_cata, T box,
                              const double* __restrict__ part, T* __restrict__ gmass, T* __restrict__ mass_out,
                              T* __restrict__ com_pos, T* __restrict__ com_vel, T* __restrict__ vrad_out,
                              T* __restrict__ a_out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= size_cata) return;
  const T nan = (T)NAN;
  if (g >= (int64_t)ngroups[0]) {
    gmass[g] = T(0);
    if (mass_out) mass_out[g] = T(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (com_pos) com_pos[g * D + d] = nan;
      if (com_vel) com_vel[g * D + d] = nan;
    }
    if (vrad_out) vrad_out[g] = nan;
    if (a_out) a_out[g] = nan;
    return;
  }
  double S[kQ];
#pragma unroll
  for (int q = 0; q < kQ; q++) S[q] = 0.0;
  const int64_t b0 = item_off[g], b1 = b0 + nit[g];
  for (int64_t b = b0; b < b1; b++) {
#pragma unroll
    for (int q = 0; q < kQ; q++) S[q] += part[b * kQ + q];
  }
  // fof.py:722-728: a mass array is summed, a constant mass multiplies the count
  const T gm = mass_array ? (T)S[0] : (T)count[g] * mass_scalar;
  gmass[g] = gm;
  if (mass_out) mass_out[g] = gm;
  const int64_t s = offset[g];
#pragma unroll
  for (int d = 0; d < D; d++) {
    if (com_pos) com_pos[g * D + d] = wrap_pos<T>((T)S[1 + d] / gm + pos[s * D + d], box);
    if (com_vel) com_vel[g * D + d] = (T)S[4 + d] / gm + vel[s * D + d];
  }
  if (vrad_out) vrad_out[g] = (T)S[7] / gm;
  if (a_out) a_out[g] = (T)S[8] / gm;
}

// pass 2: mass-weighted squared distance to the centre of mass.  The reference adds dx[0]^2 +
// dx[1]^2 + dx[2]^2 (fof.py:736); on 2D data JAX clamps the index 2 to 1, so the second
// component is counted twice there, which this restates.
template <int D, typename T>
__global__ __launch_bounds__(kCatThreads) void k_cat_pass2(
    const T* __restrict__ pos, const T* __restrict__ mass, T mass_scalar, const int32_t* __restrict__ item_off,
    const int32_t* __restrict__ nit, const int32_t* __restrict__ offset, const int32_t* __restrict__ count,
    int64_t size_cata, T box, const T* __restrict__ com_pos, double* __restrict__ part) {
  __shared__ double s_red[kCatThreads / 64];
  const int64_t nitems = total_items(item_off, nit, size_cata);
  for (int64_t b = blockIdx.x; b < nitems; b += gridDim.x) {
  const Item it = item_of(b, item_off, offset, count, size_cata);
  T c[D];
#pragma unroll
  for (int d = 0; d < D; d++) c[d] = com_pos[it.g * D + d];
  double acc = 0.0;
  for (int64_t i = it.begin + threadIdx.x; i < it.end; i += kCatThreads) {
    const T m = mass ? mass[i] : mass_scalar;
    T dx[D];
#pragma unroll
    for (int d = 0; d < D; d++) dx[d] = wrap_dx<T>(pos[i * D + d] - c[d], box);
    const T r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[D > 2 ? 2 : D - 1] * dx[D > 2 ? 2 : D - 1];
    acc += (double)(r2 * m);
  }
  const double t = block_sum(acc, s_red);
  if (threadIdx.x == 0) part[b * kQ] = t;
  }
}

template <typename T>
__global__ void k_cat_finish2(const int32_t* __restrict__ item_off, const int32_t* __restrict__ nit,
                              const int32_t* __restrict__ ngroups, int64_t size_cata, const double* __restrict__ part,
                              const T* __restrict__ gmass, T* __restrict__ radius) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= size_cata) return;
  if (g >= (int64_t)ngroups[0]) {
    radius[g] = (T)NAN;
    return;
  }
  double S = 0.0;
  const int64_t b0 = item_off[g], b1 = b0 + nit[g];
  for (int64_t b = b0; b < b1; b++) S += part[b * kQ];
  radius[g] = sqrt((T)S / gmass[g]);
}

struct CatLayout {
  size_t keep, ngr, nit, ioff, gmass, com, part, scan, total;
  int64_t max_items;
};

CatLayout cat_layout(int64_t n, int64_t size_cata, int dim) {
  CatLayout L;
  size_t o = 0;
  L.max_items = size_cata + div_up(n, kCatChunk);
  L.keep = o; o = align256(o + (size_t)n);
  L.ngr = o; o = align256(o + 4);
  L.nit = o; o = align256(o + 4 * (size_t)size_cata);
  L.ioff = o; o = align256(o + 4 * (size_t)size_cata);
  L.gmass = o; o = align256(o + 8 * (size_t)size_cata);
  L.com = o; o = align256(o + 8 * (size_t)size_cata * (size_t)dim);
  L.part = o; o = align256(o + 8 * (size_t)kQ * (size_t)L.max_items);
  const int64_t sw = scan_tmp_words(n > size_cata ? n : size_cata) + (n > size_cata ? n : size_cata);
  L.scan = o; o = align256(o + 4 * (size_t)sw);
  L.total = o;
  return L;
}

template <int D, typename T>
int catalogue_impl(hipStream_t s, const T* pos, const T* mass, T mass_scalar, const T* vel, const T* vrad,
                   const T* zp1, const int32_t* counts, int64_t n, int64_t size_cata, int npart_min, T box,
                   int32_t* ngroups, int32_t* count, int32_t* offset, T* mass_out, T* com_pos, T* radius, T* com_vel,
                   T* vrad_out, T* a_out, char* tmp) {
  const CatLayout L = cat_layout(n, size_cata, D);
  int8_t* keep = (int8_t*)(tmp + L.keep);
  int32_t* nit = (int32_t*)(tmp + L.nit);
  int32_t* ioff = (int32_t*)(tmp + L.ioff);
  T* gmass = (T*)(tmp + L.gmass);
  T* com = com_pos ? com_pos : (T*)(tmp + L.com);  // pass 2 needs the centres even when not returned
  double* part = (double*)(tmp + L.part);
  int32_t* stmp = (int32_t*)(tmp + L.scan);
  hipLaunchKernelGGL(k_cat_keep, dim3((unsigned)div_up(n, 256)), dim3(256), 0, s, counts, n, npart_min, keep);
  if (where_flags_i8(s, keep, n, offset, size_cata, (int32_t)n, nullptr, ngroups, stmp)) return 1;
  const unsigned gg = (unsigned)div_up(size_cata, 256);
  hipLaunchKernelGGL(k_cat_groups, dim3(gg), dim3(256), 0, s, counts, offset, ngroups, size_cata, count, nit);
  if (scan_i32(s, nit, ioff, size_cata, false, stmp)) return 1;
  const unsigned gi = (unsigned)(L.max_items < 4096 ? L.max_items : 4096);
  hipLaunchKernelGGL((k_cat_pass1<D, T>), dim3(gi), dim3(kCatThreads), 0, s, pos, mass, mass_scalar, vel, vrad, zp1,
                     ioff, nit, offset, count, size_cata, box, part);
  hipLaunchKernelGGL((k_cat_finish1<D, T>), dim3(gg), dim3(256), 0, s, pos, mass != nullptr, mass_scalar, vel, ioff,
                     nit, offset, count, ngroups, size_cata, box, part, gmass, mass_out, com, com_vel, vrad_out, a_out);
  if (radius) {
    hipLaunchKernelGGL((k_cat_pass2<D, T>), dim3(gi), dim3(kCatThreads), 0, s, pos, mass, mass_scalar, ioff, nit,
                       offset, count, size_cata, box, com, part);
    hipLaunchKernelGGL((k_cat_finish2<T>), dim3(gg), dim3(256), 0, s, ioff, nit, ngroups, size_cata, part, gmass,
                       radius);
  }
  return check_launch("FofCatalogue");
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" size_t jzt_fof_catalogue_tmp_bytes(int64_t size_part, int64_t size_cata, int dim, int dtype) {
  (void)dtype;
  return cat_layout(size_part, size_cata, dim).total;
}

extern "C" int jzt_fof_catalogue(jzt_stream_t stream, const void* pos, const void* mass, double mass_scalar,
                                 const void* vel, const void* v_rad, const void* zplus1, const int32_t* counts,
                                 int64_t size_part, int64_t size_cata, int npart_min, double boxsize, int dim,
                                 int dtype, int32_t* ngroups, int32_t* count, int32_t* offset, void* mass_out,
                                 void* com_pos, void* com_inertia_radius, void* com_vel, void* v_rad_out,
                                 void* scale_factor, void* tmp, size_t tmp_bytes) {
  hipStream_t s = (hipStream_t)stream;
  if (size_part <= 0) return fail("FofCatalogue: no particles");
  if (size_part >= (int64_t)INT32_MAX) return fail("FofCatalogue: size_part must be < 2^31");
  if (npart_min < 1) return fail("FofCatalogue: npart_min must be >= 1");
  if (size_cata < 1) return fail("FofCatalogue: size_cata must be >= 1");
  if (tmp_bytes < cat_layout(size_part, size_cata, dim).total) return fail("FofCatalogue: tmp too small");
  if (cat_layout(size_part, size_cata, dim).max_items >= (int64_t)INT32_MAX)
    return fail("FofCatalogue: too many work items");
  char* t = (char*)tmp;
#define JZT_CAT(D, T)                                                                                               \
  return catalogue_impl<D, T>(s, (const T*)pos, (const T*)mass, (T)mass_scalar, (const T*)vel, (const T*)v_rad,    \
                              (const T*)zplus1, counts, size_part, size_cata, npart_min, (T)boxsize, ngroups, count, \
                              offset, (T*)mass_out, (T*)com_pos, (T*)com_inertia_radius, (T*)com_vel,              \
                              (T*)v_rad_out, (T*)scale_factor, t)
  if (dtype == JZT_F32) {
    if (dim == 3) JZT_CAT(3, float);
    if (dim == 2) JZT_CAT(2, float);
  } else if (dtype == JZT_F64) {
    if (dim == 3) JZT_CAT(3, double);
    if (dim == 2) JZT_CAT(2, double);
  }
#undef JZT_CAT
  return fail("FofCatalogue: unsupported dim/dtype (dim " + std::to_string(dim) + ")");
}
