// Learned radius proposer for the bounded exact kNN search (no reference counterpart).
//
// Per leaf, a small MLP predicts log2 of the largest k-th neighbour distance among the leaf's
// particles from 12 cheap tree features.  The prediction (plus a margin) only sets the starting
// bound of the exact search; queries that find fewer than k neighbours inside it are recomputed
// without a bound, so results never depend on the network.
//
// Network: 12 -> 64 -> 64 -> 64 -> 1, GELU (erf form) after each hidden layer, weights stored as
// bf16 (the trainable fp32 master copy is rounded to nearest-even after every Adam step), all
// arithmetic and accumulation in fp32.
//
// Work layout of the training kernel: one wave per tile of 64 samples.  Activations of a tile live
// in LDS transposed (row = neuron, column = sample, row stride 65 floats so that a wave writing one
// neuron row per lane hits 64 distinct banks); a layer is computed with lane = output neuron,
// reading each input row by broadcast and keeping the 64 per-sample accumulators in registers.
// Inference has no backward pass to feed and runs one sample per lane instead (k_mlp_infer).
#include "util.h"
#include "../../include/jztree_amd.h"

namespace jzt {
namespace {

constexpr int kNF = JZT_PROPOSER_NFEAT;  // 12 input features
constexpr int kNH = 64;                  // hidden width
constexpr int kOffW1 = 0;                // W1[kNF][kNH]  (input-major: W[j * kNH + k], j in, k out)
constexpr int kOffB1 = kOffW1 + kNF * kNH;
constexpr int kOffW2 = kOffB1 + kNH;     // W2[kNH][kNH]
constexpr int kOffB2 = kOffW2 + kNH * kNH;
constexpr int kOffW3 = kOffB2 + kNH;     // W3[kNH][kNH]
constexpr int kOffB3 = kOffW3 + kNH * kNH;
constexpr int kOffW4 = kOffB3 + kNH;     // W4[kNH]
constexpr int kOffB4 = kOffW4 + kNH;     // b4
constexpr int kNParam = kOffB4 + 1;      // 9217
static_assert(kNParam == JZT_PROPOSER_NPARAM, "parameter count");
constexpr int kLd = 65;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}

constexpr float kInvSqrt2 = 0.70710678118654752f;
constexpr float kInvSqrt2Pi = 0.39894228040143268f;
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * kInvSqrt2)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.f + erff(z * kInvSqrt2)) + z * kInvSqrt2Pi * __expf(-0.5f * z * z);
}

__device__ __forceinline__ float lg2n(int n) { return log2f((float)max(n, 1)); }

// ------------------------------------------------------------------------------------------------
// features
// ------------------------------------------------------------------------------------------------

// 12 features of leaf i (leaf slots >= *nleaves get zeros):
//   0 log2 n(i)          1 level(i)/3               2 log2 n(parent)      3 level(parent)/3
//   4 log2 n(grandpa)    5 level(grandpa)/3         6 log2 n(i-1)         7 log2 n(i+1)
//   8 level(i-1)/3 - f1  9 level(i+1)/3 - f1        10 log2 #leaves in parent
//   11 log2 of the particles in leaves i-2 .. i+2
// n = particle count, level = cell level of the z-order tree (a level step of 3 halves the cell
// side), parent/grandpa = enclosing nodes of planes 1 and 2 (the parent's values repeat when the
// tree has no plane 2); i-1, i+1 are the z-order neighbours (the leaf itself at the ends).
__global__ void k_leaf_features(const int32_t* __restrict__ npart0, const int32_t* __restrict__ lvl0,
                                const int32_t* __restrict__ par1, const int32_t* __restrict__ npart1,
                                const int32_t* __restrict__ lvl1, const int32_t* __restrict__ n2c1,
                                const int32_t* __restrict__ par2, const int32_t* __restrict__ npart2,
                                const int32_t* __restrict__ lvl2, const int32_t* __restrict__ nleaves_dev,
                                int64_t size, float* __restrict__ feat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  const int64_t nl = min((int64_t)nleaves_dev[0], size);
  float f[kNF];
  if (i >= nl) {
#pragma unroll
    for (int c = 0; c < kNF; c++) f[c] = 0.f;
  } else {
    const int p = min(max(par1[i], 0), (int)size - 1);
    const int g = par2 ? min(max(par2[p], 0), (int)size - 1) : -1;
    const int n = npart0[i];
    const float l = (float)lvl0[i] * (1.f / 3.f);
    const int64_t im = i > 0 ? i - 1 : i, ip = i + 1 < nl ? i + 1 : i;
    int nsum = 0;
    for (int64_t d = max(i - 2, (int64_t)0); d <= min(i + 2, nl - 1); d++) nsum += npart0[d];
    f[0] = lg2n(n);
    f[1] = l;
    f[2] = lg2n(npart1[p]);
    f[3] = (float)lvl1[p] * (1.f / 3.f);
    f[4] = g >= 0 ? lg2n(npart2[g]) : f[2];
    f[5] = g >= 0 ? (float)lvl2[g] * (1.f / 3.f) : f[3];
    f[6] = lg2n(npart0[im]);
    f[7] = lg2n(npart0[ip]);
    f[8] = (float)lvl0[im] * (1.f / 3.f) - l;
    f[9] = (float)lvl0[ip] * (1.f / 3.f) - l;
    f[10] = lg2n(n2c1[p + 1] - n2c1[p]);
    f[11] = lg2n(nsum);
  }
#pragma unroll
  for (int c = 0; c < kNF; c++) feat[i * kNF + c] = f[c];
}

// ------------------------------------------------------------------------------------------------
// MLP building blocks (one wave, one tile of 64 samples)
// ------------------------------------------------------------------------------------------------

// acc[s] = b[k] + sum_j in[j][s] * W[j][k] for lane k (in: LDS rows of stride kLd)
template <int NIN>
__device__ __forceinline__ void layer_acc(float (&acc)[64], const float* __restrict__ in, const uint16_t* __restrict__ w,
                                          const uint16_t* __restrict__ b, int k) {
  const float bk = bf16_to_f32(b[k]);
#pragma unroll
  for (int s = 0; s < 64; s++) acc[s] = bk;
  for (int j = 0; j < NIN; j++) {
    const float wj = bf16_to_f32(w[j * kNH + k]);
    const float* row = in + j * kLd;
#pragma unroll
    for (int s = 0; s < 64; s++) acc[s] = fmaf(row[s], wj, acc[s]);
  }
}

// ------------------------------------------------------------------------------------------------
// inference
// ------------------------------------------------------------------------------------------------

// norm = {feature mean[12], feature std[12], target mean, target std}
//
// Inference runs one sample per lane, so all lanes do useful work.  The weights, converted to fp32
// once per block, sit in LDS and every lane reads the same address (a broadcast); the lane's input
// activations sit in its own LDS column (written and read by that lane only) and the 64
// accumulators stay in registers.  Each output neuron accumulates b[k] + sum_j in[j] * W[j][k] in
// ascending j with fmaf, the order of layer_acc, so the results are the training forward pass's,
// bit for bit.
constexpr int kInferThreads = 128;
template <int NIN>
__device__ __forceinline__ void layer_lane(float (&out)[kNH], const float* __restrict__ col, const float* __restrict__ w,
                                           const float* __restrict__ b) {
#pragma unroll
  for (int k = 0; k < kNH; k++) out[k] = b[k];
#pragma unroll 1
  for (int j = 0; j < NIN; j++) {
    const float x = col[j * kInferThreads];
    const float4* wr = reinterpret_cast<const float4*>(w + j * kNH);
#pragma unroll
    for (int k4 = 0; k4 < kNH / 4; k4++) {
      const float4 w4 = wr[k4];
      out[4 * k4 + 0] = fmaf(x, w4.x, out[4 * k4 + 0]);
      out[4 * k4 + 1] = fmaf(x, w4.y, out[4 * k4 + 1]);
      out[4 * k4 + 2] = fmaf(x, w4.z, out[4 * k4 + 2]);
      out[4 * k4 + 3] = fmaf(x, w4.w, out[4 * k4 + 3]);
    }
  }
}

__global__ __launch_bounds__(kInferThreads) void k_mlp_infer(const uint16_t* __restrict__ wq,
                                                             const float* __restrict__ norm,
                                                             const float* __restrict__ feat,
                                                             const int32_t* __restrict__ npart0, int64_t n,
                                                             float margin, float* __restrict__ out_r2,
                                                             float* __restrict__ out_log2r) {
  __shared__ __align__(16) float s_w[(kNParam + 3) / 4 * 4];
  __shared__ float s_act[kNH * kInferThreads];  // column t: thread t's activations
  const int tid = threadIdx.x;
  for (int i = tid; i < kNParam; i += kInferThreads) s_w[i] = bf16_to_f32(wq[i]);
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * kInferThreads + tid;
  const bool live = s < n && !(npart0 && out_r2 && !out_log2r && npart0[s] <= 0);
  if (!__ballot(live)) {  // a wave of empty leaf slots only: their bound is 0
    if (s < n && out_r2) out_r2[s] = 0.f;
    return;
  }
  float* col = s_act + tid;
  const float ym = norm[2 * kNF], ys = norm[2 * kNF + 1];
#pragma unroll
  for (int c = 0; c < kNF; c++) col[c * kInferThreads] = s < n ? (feat[s * kNF + c] - norm[c]) / norm[kNF + c] : 0.f;
  float h[kNH];
  layer_lane<kNF>(h, col, s_w + kOffW1, s_w + kOffB1);
#pragma unroll
  for (int k = 0; k < kNH; k++) col[k * kInferThreads] = gelu(h[k]);
  layer_lane<kNH>(h, col, s_w + kOffW2, s_w + kOffB2);
#pragma unroll
  for (int k = 0; k < kNH; k++) col[k * kInferThreads] = gelu(h[k]);
  layer_lane<kNH>(h, col, s_w + kOffW3, s_w + kOffB3);
  float y = s_w[kOffB4];
#pragma unroll
  for (int j = 0; j < kNH; j++) y = fmaf(gelu(h[j]), s_w[kOffW4 + j], y);
  if (s < n) {
    const float lr = fmaf(y, ys, ym);
    if (out_log2r) out_log2r[s] = lr;
    if (out_r2) {
      const float r = exp2f(lr + margin);
      out_r2[s] = npart0 && npart0[s] <= 0 ? 0.f : r * r;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// training: forward + backward of the mean squared error over a batch, gradients accumulated
// ------------------------------------------------------------------------------------------------

// x: normalised features (B, 12), t: normalised targets (B).  grad (kNParam) and loss (1) must be
// zero on entry; they receive d/dW of (1/B) sum_s (y_s - t_s)^2 and that loss.  The gradient is
// taken with respect to the bf16 weights as the forward pass uses them.
__global__ __launch_bounds__(64) void k_mlp_grad(const uint16_t* __restrict__ wq, const float* __restrict__ x,
                                                 const float* __restrict__ t, int64_t batch, float inv_b,
                                                 float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float s_x[kNF * kLd];
  __shared__ float s_a1[kNH * kLd], s_a2[kNH * kLd], s_a3[kNH * kLd];  // activations gelu(z)
  __shared__ float s_g1[kNH * kLd], s_g2[kNH * kLd], s_g3[kNH * kLd];  // gelu'(z)
  __shared__ float s_d[kNH * kLd];                                     // dL/dz of the current layer
  __shared__ float s_dy[64];
  __shared__ uint16_t s_w[kNParam];
  const int lane = threadIdx.x;
  for (int i = lane; i < kNParam; i += 64) s_w[i] = wq[i];
  __syncthreads();
  float acc[64];
  for (int64_t tile = blockIdx.x; tile * 64 < batch; tile += gridDim.x) {
    const int64_t s = tile * 64 + lane;
    const bool valid = s < batch;
#pragma unroll
    for (int c = 0; c < kNF; c++) s_x[c * kLd + lane] = valid ? x[s * kNF + c] : 0.f;
    __syncthreads();
    // forward (lane = neuron)
    layer_acc<kNF>(acc, s_x, s_w + kOffW1, s_w + kOffB1, lane);
#pragma unroll
    for (int q = 0; q < 64; q++) {
      s_a1[lane * kLd + q] = gelu(acc[q]);
      s_g1[lane * kLd + q] = gelu_grad(acc[q]);
    }
    __syncthreads();
    layer_acc<kNH>(acc, s_a1, s_w + kOffW2, s_w + kOffB2, lane);
#pragma unroll
    for (int q = 0; q < 64; q++) {
      s_a2[lane * kLd + q] = gelu(acc[q]);
      s_g2[lane * kLd + q] = gelu_grad(acc[q]);
    }
    __syncthreads();
    layer_acc<kNH>(acc, s_a2, s_w + kOffW3, s_w + kOffB3, lane);
#pragma unroll
    for (int q = 0; q < 64; q++) {
      s_a3[lane * kLd + q] = gelu(acc[q]);
      s_g3[lane * kLd + q] = gelu_grad(acc[q]);
    }
    __syncthreads();
    // output and loss (lane = sample)
    float y = bf16_to_f32(s_w[kOffB4]);
    for (int j = 0; j < kNH; j++) y = fmaf(s_a3[j * kLd + lane], bf16_to_f32(s_w[kOffW4 + j]), y);
    const float e = valid ? y - t[s] : 0.f;
    s_dy[lane] = 2.f * e * inv_b;
    float l2 = e * e * inv_b;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o, 64);
    if (lane == 0) atomicAdd(loss, l2);
    __syncthreads();
    // output layer gradients (lane = hidden unit j) and dL/dz3 (lane = neuron k)
    {
      float gw = 0.f, gb = 0.f;
      const float w4 = bf16_to_f32(s_w[kOffW4 + lane]);
      for (int q = 0; q < 64; q++) {
        const float dy = s_dy[q];
        gw = fmaf(s_a3[lane * kLd + q], dy, gw);
        gb += dy;
        s_d[lane * kLd + q] = dy * w4 * s_g3[lane * kLd + q];
      }
      atomicAdd(&grad[kOffW4 + lane], gw);
      if (lane == 0) atomicAdd(&grad[kOffB4], gb);
    }
    __syncthreads();
    // hidden layers, top down: with s_d = dL/dz_l (row = neuron k), accumulate
    //   dW_l[j][k] = sum_s a_{l-1}[j][s] d[k][s],  db_l[k] = sum_s d[k][s]
    // then dL/dz_{l-1}[j][s] = g_{l-1}[j][s] * sum_k W_l[j][k] d[k][s]
    const float* a_in[3] = {s_x, s_a1, s_a2};
    const float* g_in[3] = {nullptr, s_g1, s_g2};
    const int offw[3] = {kOffW1, kOffW2, kOffW3};
    const int offb[3] = {kOffB1, kOffB2, kOffB3};
#pragma unroll
    for (int l = 2; l >= 0; l--) {
      const int nin = l == 0 ? kNF : kNH;
      // weight gradient column k = lane
      float dk[64];
#pragma unroll
      for (int q = 0; q < 64; q++) dk[q] = s_d[lane * kLd + q];
      float gb = 0.f;
#pragma unroll
      for (int q = 0; q < 64; q++) gb += dk[q];
      atomicAdd(&grad[offb[l] + lane], gb);
      for (int j = 0; j < nin; j++) {
        const float* row = a_in[l] + j * kLd;
        float gw = 0.f;
#pragma unroll
        for (int q = 0; q < 64; q++) gw = fmaf(row[q], dk[q], gw);
        atomicAdd(&grad[offw[l] + j * kNH + lane], gw);
      }
      if (l == 0) break;
      // back-propagate to z_{l-1}: lane = input unit j
#pragma unroll
      for (int q = 0; q < 64; q++) acc[q] = 0.f;
      for (int k = 0; k < kNH; k++) {
        const float wjk = bf16_to_f32(s_w[offw[l] + lane * kNH + k]);
        const float* row = s_d + k * kLd;
#pragma unroll
        for (int q = 0; q < 64; q++) acc[q] = fmaf(row[q], wjk, acc[q]);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 64; q++) s_d[lane * kLd + q] = acc[q] * g_in[l][lane * kLd + q];
      __syncthreads();
    }
    __syncthreads();
  }
}

__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wq, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  wq[i] = f32_to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wq, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wq[i] = f32_to_bf16(w[i]);
}

}  // namespace
}  // namespace jzt

using namespace jzt;

extern "C" int jzt_proposer_num_params(void) { return kNParam; }

extern "C" int jzt_leaf_features(jzt_stream_t stream, const int32_t* npart0, const int32_t* lvl0,
                                 const int32_t* leaf_parent, const int32_t* npart1, const int32_t* lvl1,
                                 const int32_t* parent_child_spl, const int32_t* parent_parent,
                                 const int32_t* npart2, const int32_t* lvl2, const int32_t* nleaves,
                                 int64_t size, float* feat) {
  if (size <= 0) return 0;
  if (parent_parent && (!npart2 || !lvl2)) return fail("LeafFeatures: plane-2 arrays missing");
  hipLaunchKernelGGL(k_leaf_features, dim3((unsigned)div_up(size, 256)), dim3(256), 0, (hipStream_t)stream, npart0,
                     lvl0, leaf_parent, npart1, lvl1, parent_child_spl, parent_parent, npart2, lvl2, nleaves, size,
                     feat);
  return check_launch("LeafFeatures");
}

extern "C" int jzt_mlp_infer(jzt_stream_t stream, const uint16_t* weights, const float* norm, const float* feat,
                             const int32_t* npart, int64_t n, float margin, float* out_r2, float* out_log2r) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)div_up(n, kInferThreads)), dim3(kInferThreads), 0,
                     (hipStream_t)stream, weights, norm, feat, npart, n, margin, out_r2, out_log2r);
  return check_launch("MlpInfer");
}

extern "C" int jzt_mlp_grad(jzt_stream_t stream, const uint16_t* weights, const float* x, const float* t,
                            int64_t batch, float* grad, float* loss) {
  hipStream_t s = (hipStream_t)stream;
  if (fill_u32(s, (uint32_t*)grad, 0u, kNParam) || fill_u32(s, (uint32_t*)loss, 0u, 1)) return 1;
  if (batch <= 0) return 0;
  const int64_t tiles = div_up(batch, 64);
  const unsigned grid = (unsigned)(tiles < 1024 ? tiles : 1024);
  hipLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(64), 0, s, weights, x, t, batch, 1.f / (float)batch, grad, loss);
  return check_launch("MlpGrad");
}

extern "C" int jzt_adam_step(jzt_stream_t stream, float* w, uint16_t* wq, float* m, float* v, const float* grad,
                             int64_t n, float lr, float beta1, float beta2, float eps, int step) {
  if (n <= 0) return 0;
  if (step < 1) return fail("AdamStep: step counts from 1");
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, w, wq, m, v, grad, n,
                     lr, beta1, beta2, eps, bc1, bc2);
  return check_launch("AdamStep");
}

extern "C" int jzt_to_bf16(jzt_stream_t stream, const float* w, uint16_t* wq, int64_t n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, w, wq, n);
  return check_launch("ToBf16");
}
