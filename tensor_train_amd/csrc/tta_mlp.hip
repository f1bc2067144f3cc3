// Rank-prediction MLP: n_in -> H -> H -> H -> 1 (ReLU hidden, linear out),
// FP32 parameters, trained by mean-squared error with Adam.  Forward,
// backward and the optimiser are the kernels below (tiled FP32 GEMMs with
// fused bias / ReLU, column sums, Adam).  No reference counterpart: it
// predicts, per bond, the truncation rank the plain rounding would choose,
// to set the target ranks of the randomized rounding.
//
// Parameter layout (flat FP32): for each layer l (in_l x out_l):
//   W_l (in_l x out_l, row-major) then b_l (out_l).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "../../include/tt_amd.h"
#include "tta_internal.hpp"
#include "tta_kernels.hpp"

struct tta_mlp
{
  int n_in = 0, hidden = 0, layers = 0;  // layers = number of hidden layers
  std::vector<int> dims;                  // n_in, H, ..., H, 1
  std::vector<size_t> woff, boff;         // offsets into params
  size_t nparams = 0;
  float* p = nullptr;   // params (device)
  float* g = nullptr;   // grads
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  long long step = 0;
  bool bf16_infer = false;  // inference with BF16 weights and activations
};

namespace tta
{
namespace
{
constexpr int TS = 16;

// C (M x N) = op(A) op(B) [+ bias] [relu], FP32.  opA: A(i,k) = trA ? A[k*lda+i] : A[i*lda+k].
// Split over K: block z sums k in [z*ks, min(K, (z+1)*ks)) into C + z*cstr
// (one z and ks = K for a plain product).
__global__ __launch_bounds__(256) void sgemm_kernel(int M, int N, int K, const float* A, int lda, int trA,
                                                    const float* B, int ldb, int trB, float* C, int ldc,
                                                    const float* bias, int relu, float beta, int ks, size_t cstr)
{
  __shared__ float As[TS][TS + 1];
  __shared__ float Bs[TS][TS + 1];
  const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
  const int i = blockIdx.y * TS + ty, j = blockIdx.x * TS + tx;
  const int kbeg = blockIdx.z * ks, kend = min(K, kbeg + ks);
  C += blockIdx.z * cstr;
  float acc = 0.f;
  for (int k0 = kbeg; k0 < kend; k0 += TS)
  {
    const int ka = k0 + tx, kb = k0 + ty;
    As[ty][tx] = (i < M && ka < kend) ? (trA ? A[(size_t)ka * lda + i] : A[(size_t)i * lda + ka]) : 0.f;
    Bs[ty][tx] = (kb < kend && j < N) ? (trB ? B[(size_t)j * ldb + kb] : B[(size_t)kb * ldb + j]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TS; ++kk) acc = fmaf(As[ty][kk], Bs[kk][tx], acc);
    __syncthreads();
  }
  if (i < M && j < N)
  {
    if (bias) acc += bias[j];
    if (relu) acc = fmaxf(acc, 0.f);
    float* c = C + (size_t)i * ldc + j;
    *c = beta == 0.f ? acc : acc + beta * (*c);
  }
}

// Round to the nearest BF16 value (ties to even), returned as FP32.
__device__ __forceinline__ float to_bf16(float x)
{
  unsigned int b = __float_as_uint(x);
  if ((b & 0x7f800000u) != 0x7f800000u) b += 0x7fffu + ((b >> 16) & 1u);
  return __uint_as_float(b & 0xffff0000u);
}

// Inference layer with BF16 operands: C (M x N) = bf16(A) bf16(W) + bias
// [relu], products accumulated in FP32; A row-major (lda), W row-major (N
// wide).  The weights are rounded once per launch as they are staged.
__global__ __launch_bounds__(256) void bf16_layer_kernel(int M, int N, int K, const float* A, int lda, const float* W,
                                                         const float* bias, int relu, float* C, int ldc)
{
  __shared__ float As[TS][TS + 1];
  __shared__ float Bs[TS][TS + 1];
  const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
  const int i = blockIdx.y * TS + ty, j = blockIdx.x * TS + tx;
  float acc = 0.f;
  for (int k0 = 0; k0 < K; k0 += TS)
  {
    const int ka = k0 + tx, kb = k0 + ty;
    As[ty][tx] = (i < M && ka < K) ? to_bf16(A[(size_t)i * lda + ka]) : 0.f;
    Bs[ty][tx] = (kb < K && j < N) ? to_bf16(W[(size_t)kb * N + j]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TS; ++kk) acc = fmaf(As[ty][kk], Bs[kk][tx], acc);
    __syncthreads();
  }
  if (i < M && j < N)
  {
    acc += bias ? to_bf16(bias[j]) : 0.f;
    if (relu) acc = fmaxf(acc, 0.f);
    C[(size_t)i * ldc + j] = acc;
  }
}

// dH *= (H > 0)
__global__ void relu_mask_kernel(float* dH, const float* H, long n)
{
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    if (H[t] <= 0.f) dH[t] = 0.f;
}

// dout = 2 (out - y) / n; loss partials.
__global__ __launch_bounds__(256) void mse_kernel(const float* out, const float* y, float* dout, int n, float* loss)
{
  __shared__ float red[256];
  float acc = 0.f;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256)
  {
    const float e = out[t] - y[t];
    dout[t] = 2.f * e / (float)n;
    acc += e * e;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1)
  {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}

// Partial column sums of dY (n x m): block (x, z) sums rows
// [z*ks, min(n, (z+1)*ks)) of columns [64x, 64x + 64) into part + z*pstr.
__global__ __launch_bounds__(256) void colsum_part_kernel(const float* dY, int n, int m, int ks, float* part,
                                                          size_t pstr)
{
  __shared__ float red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + tx;
  const int r0 = blockIdx.y * ks, r1 = min(n, r0 + ks);
  float s = 0.f;
  if (j < m)
    for (int i = r0 + ty; i < r1; i += 4) s += dY[(size_t)i * m + j];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && j < m) part[blockIdx.y * pstr + j] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

// out[e] = sum_{z < np} part[z*pstr + e] in order z = 0, 1, ... (deterministic).
__global__ void sum_parts_kernel(const float* part, int np, size_t pstr, long len, float* out)
{
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < len; e += (long)gridDim.x * blockDim.x)
  {
    float s = 0.f;
    for (int z = 0; z < np; ++z) s += part[z * pstr + e];
    out[e] = s;
  }
}

// Rows of the batch per slice of the batch reduction behind dW and db.
int batch_slice(int n) { return std::max(512, (n + 127) / 128 + TS - 1) / TS * TS; }

__global__ void adam_kernel(float* p, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2,
                            float c1, float c2, float epsa)
{
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
  {
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gt;
    const float b = b2 * m2[t] + (1.f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = b;
    p[t] -= lr * (a / c1) / (sqrtf(b / c2) + epsa);
  }
}

// Training keeps the parameters of the lowest loss seen.  The loss of step s
// is that of the parameters before its update, so the copy is taken between
// the backward pass and the update.  All blocks compare against the same best
// (it is updated by the separate set_best_kernel afterwards).
__global__ void keep_best_kernel(const float* loss, const float* best, const float* p, float* bp, long n)
{
  if (!(*loss < *best)) return;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) bp[t] = p[t];
}

__global__ void set_best_kernel(const float* loss, float* best)
{
  if (threadIdx.x == 0 && *loss < *best) *best = *loss;
}

__global__ void fill_kernel(float* x, float v) { *x = v; }

// After the last step: back to the best parameters when the final ones are
// worse.
__global__ void restore_best_kernel(float* loss, const float* best, float* p, const float* bp, long n)
{
  if (!(*best < *loss)) return;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = bp[t];
}

void sgemm(hipStream_t st, int M, int N, int K, const float* A, int lda, int trA, const float* B, int ldb, int trB,
           float* C, int ldc, const float* bias = nullptr, int relu = 0, float beta = 0.f)
{
  dim3 grid((N + TS - 1) / TS, (M + TS - 1) / TS);
  hipLaunchKernelGGL(sgemm_kernel, grid, dim3(TS * TS), 0, st, M, N, K, A, lda, trA, B, ldb, trB, C, ldc, bias, relu,
                     beta, K, (size_t)0);
}

struct mlp_ws
{
  std::vector<float*> act;  // act[0] = X, act[l+1] = output of layer l
  std::vector<float*> dact;
  float* y = nullptr;
  float* loss = nullptr;
  float* part = nullptr;  // per-slice partial dW and db of one layer
};

// Forward pass over n rows; activations kept for backward.
void forward(tta_ctx* c, tta_mlp* m, int n, mlp_ws& w)
{
  const int L = (int)m->dims.size() - 1;
  for (int l = 0; l < L; ++l)
    sgemm(c->stream, n, m->dims[l + 1], m->dims[l], w.act[l], m->dims[l], 0, m->p + m->woff[l], m->dims[l + 1], 0,
          w.act[l + 1], m->dims[l + 1], m->p + m->boff[l], l < L - 1 ? 1 : 0);
}

// Gradients of the MSE loss into m->g; loss (device float) accumulated.
void backward(tta_ctx* c, tta_mlp* m, int n, mlp_ws& w)
{
  const int L = (int)m->dims.size() - 1;
  hip_check(hipMemsetAsync(w.loss, 0, sizeof(float), c->stream), "memset");
  hipLaunchKernelGGL(mse_kernel, dim3(std::min(64, (n + 255) / 256)), dim3(256), 0, c->stream, w.act[L], w.y,
                     w.dact[L], n, w.loss);
  for (int l = L - 1; l >= 0; --l)
  {
    const int in = m->dims[l], out = m->dims[l + 1];
    // dW = act_l^T dact_{l+1}; db = colsum(dact_{l+1}).  Both reduce over the
    // n rows of the batch: each slice of ks rows writes its partial [dW | db]
    // (db follows dW in the parameter layout), then one pass sums the slices.
    const int ks = batch_slice(n), np = (n + ks - 1) / ks;
    const size_t len = (size_t)in * out + out;
    dim3 grid((out + TS - 1) / TS, (in + TS - 1) / TS, np);
    hipLaunchKernelGGL(sgemm_kernel, grid, dim3(TS * TS), 0, c->stream, in, out, n, w.act[l], in, 1, w.dact[l + 1],
                       out, 0, w.part, out, nullptr, 0, 0.f, ks, len);
    hipLaunchKernelGGL(colsum_part_kernel, dim3((out + 63) / 64, np), dim3(256), 0, c->stream, w.dact[l + 1], n, out,
                       ks, w.part + (size_t)in * out, len);
    hipLaunchKernelGGL(sum_parts_kernel, dim3((int)std::min<size_t>((len + 255) / 256, 1024)), dim3(256), 0,
                       c->stream, w.part, np, len, (long)len, m->g + m->woff[l]);
    if (l > 0)
    {
      // dact_l = dact_{l+1} W^T, masked by relu'(act_l).
      sgemm(c->stream, n, in, out, w.dact[l + 1], out, 0, m->p + m->woff[l], out, 1, w.dact[l], in);
      const long tot = (long)n * in;
      hipLaunchKernelGGL(relu_mask_kernel, dim3((int)std::min<long>((tot + 255) / 256, 4096)), dim3(256), 0,
                         c->stream, w.dact[l], w.act[l], tot);
    }
  }
}

// Inference forward pass with BF16 operands (FP32 accumulation).
void forward_bf16(tta_ctx* c, tta_mlp* m, int n, mlp_ws& w)
{
  const int L = (int)m->dims.size() - 1;
  for (int l = 0; l < L; ++l)
  {
    const int in = m->dims[l], out = m->dims[l + 1];
    dim3 grid((out + TS - 1) / TS, (n + TS - 1) / TS);
    hipLaunchKernelGGL(bf16_layer_kernel, grid, dim3(TS * TS), 0, c->stream, n, out, in, w.act[l], in,
                       m->p + m->woff[l], m->p + m->boff[l], l < L - 1 ? 1 : 0, w.act[l + 1], out);
  }
}

struct ws_owner
{
  tta_ctx* c;
  std::vector<void*> b;
  explicit ws_owner(tta_ctx* c_) : c(c_) {}
  ~ws_owner()
  {
    for (void* p : b) c->mem.put(p);
  }
  float* f(size_t n)
  {
    void* p = c->mem.get(std::max<size_t>(n, 1) * sizeof(float));
    b.push_back(p);
    return static_cast<float*>(p);
  }
};

void make_ws(tta_ctx* c, tta_mlp* m, int n, ws_owner& O, mlp_ws& w, bool grads)
{
  const int L = (int)m->dims.size() - 1;
  w.act.assign(L + 1, nullptr);
  w.dact.assign(L + 1, nullptr);
  for (int l = 0; l <= L; ++l)
  {
    w.act[l] = O.f((size_t)n * m->dims[l]);
    if (grads) w.dact[l] = O.f((size_t)n * m->dims[l]);
  }
  w.y = O.f((size_t)n);
  w.loss = O.f(1);
  if (grads)
  {
    size_t len = 0;
    for (int l = 0; l < L; ++l) len = std::max(len, (size_t)m->dims[l] * m->dims[l + 1] + m->dims[l + 1]);
    const int ks = batch_slice(n);
    w.part = O.f(len * (size_t)((n + ks - 1) / ks));
  }
}
}  // namespace
}  // namespace tta

using namespace tta;

extern "C"
{

int tta_mlp_create(tta_ctx* c, int n_in, int hidden, int layers, unsigned long long seed, tta_mlp** out)
{
  if (!c || !out || n_in < 1 || hidden < 1 || layers < 1) return TTA_ERR_ARG;
  try
  {
    auto* m = new tta_mlp();
    m->n_in = n_in;
    m->hidden = hidden;
    m->layers = layers;
    m->dims.push_back(n_in);
    for (int l = 0; l < layers; ++l) m->dims.push_back(hidden);
    m->dims.push_back(1);
    size_t off = 0;
    for (size_t l = 0; l + 1 < m->dims.size(); ++l)
    {
      m->woff.push_back(off);
      off += (size_t)m->dims[l] * m->dims[l + 1];
      m->boff.push_back(off);
      off += (size_t)m->dims[l + 1];
    }
    m->nparams = off;
    // He initialisation (normal, std sqrt(2 / fan_in)), zero biases.
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (size_t l = 0; l + 1 < m->dims.size(); ++l)
    {
      std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)m->dims[l]));
      for (size_t e = 0; e < (size_t)m->dims[l] * m->dims[l + 1]; ++e) h[m->woff[l] + e] = nd(rng);
    }
    m->p = static_cast<float*>(c->mem.get(off * 4));
    m->g = static_cast<float*>(c->mem.get(off * 4));
    m->m1 = static_cast<float*>(c->mem.get(off * 4));
    m->m2 = static_cast<float*>(c->mem.get(off * 4));
    hip_check(hipMemcpyAsync(m->p, h.data(), off * 4, hipMemcpyHostToDevice, c->stream), "params");
    hip_check(hipMemsetAsync(m->m1, 0, off * 4, c->stream), "memset");
    hip_check(hipMemsetAsync(m->m2, 0, off * 4, c->stream), "memset");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    *out = m;
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}

int tta_mlp_free(tta_ctx* c, tta_mlp* m)
{
  if (!c || !m) return TTA_ERR_ARG;
  c->mem.put(m->p);
  c->mem.put(m->g);
  c->mem.put(m->m1);
  c->mem.put(m->m2);
  delete m;
  return TTA_OK;
}

long long tta_mlp_num_params(const tta_mlp* m) { return m ? (long long)m->nparams : -1; }

int tta_mlp_get_params(tta_ctx* c, const tta_mlp* m, float* host)
{
  if (!c || !m || !host) return TTA_ERR_ARG;
  if (hipMemcpyAsync(host, m->p, m->nparams * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return TTA_ERR_HIP;
  return hipStreamSynchronize(c->stream) == hipSuccess ? TTA_OK : TTA_ERR_HIP;
}

int tta_mlp_set_params(tta_ctx* c, tta_mlp* m, const float* host)
{
  if (!c || !m || !host) return TTA_ERR_ARG;
  if (hipMemcpyAsync(m->p, host, m->nparams * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) return TTA_ERR_HIP;
  return hipStreamSynchronize(c->stream) == hipSuccess ? TTA_OK : TTA_ERR_HIP;
}

// Loss and gradient (flat, same layout as the parameters) on n rows.
int tta_mlp_grad(tta_ctx* c, tta_mlp* m, const float* X, const float* y, int n, float* grad_out, float* loss_out)
{
  if (!c || !m || !X || !y || n < 1) return TTA_ERR_ARG;
  try
  {
    ws_owner O(c);
    mlp_ws w;
    make_ws(c, m, n, O, w, true);
    hip_check(hipMemcpyAsync(w.act[0], X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, c->stream), "X");
    hip_check(hipMemcpyAsync(w.y, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream), "y");
    forward(c, m, n, w);
    backward(c, m, n, w);
    float l = 0.f;
    hip_check(hipMemcpyAsync(&l, w.loss, 4, hipMemcpyDeviceToHost, c->stream), "loss");
    if (grad_out) hip_check(hipMemcpyAsync(grad_out, m->g, m->nparams * 4, hipMemcpyDeviceToHost, c->stream), "g");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    if (loss_out) *loss_out = l;
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}

// Full-batch Adam for `steps` steps (beta1 0.9, beta2 0.999) with the step
// size decayed from lr to lr / 100 on a half cosine over the call, keeping the
// parameters of the lowest training loss seen (full-batch Adam at a constant
// step size can leave a good minimum late in training: on 20k bond samples a
// loss of 6.6e-5 at step 1200 jumped to 1.1e-3 by step 1500).  Returns the loss
// of the parameters kept.
int tta_mlp_train(tta_ctx* c, tta_mlp* m, const float* X, const float* y, int n, int steps, float lr,
                  float* loss_out)
{
  if (!c || !m || !X || !y || n < 1 || steps < 0) return TTA_ERR_ARG;
  try
  {
    ws_owner O(c);
    mlp_ws w;
    make_ws(c, m, n, O, w, true);
    hip_check(hipMemcpyAsync(w.act[0], X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, c->stream), "X");
    hip_check(hipMemcpyAsync(w.y, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream), "y");
    const float b1 = 0.9f, b2 = 0.999f;
    const long np_ = (long)m->nparams;
    const dim3 pg((int)std::min<size_t>((m->nparams + 255) / 256, 1024));
    float* bp = O.f(m->nparams);
    float* best = O.f(1);
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(1), 0, c->stream, best, INFINITY);
    hip_check(hipMemcpyAsync(bp, m->p, m->nparams * 4, hipMemcpyDeviceToDevice, c->stream), "params");
    const double pi = 3.14159265358979323846;
    for (int s = 0; s < steps; ++s)
    {
      forward(c, m, n, w);
      backward(c, m, n, w);
      hipLaunchKernelGGL(keep_best_kernel, pg, dim3(256), 0, c->stream, w.loss, best, m->p, bp, np_);
      hipLaunchKernelGGL(set_best_kernel, dim3(1), dim3(64), 0, c->stream, w.loss, best);
      ++m->step;
      const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
      const float lr_s = (float)(0.01 * lr + 0.99 * lr * 0.5 * (1.0 + std::cos(pi * s / std::max(steps, 1))));
      hipLaunchKernelGGL(adam_kernel, pg, dim3(256), 0, c->stream, m->p, m->g, m->m1, m->m2, np_, lr_s, b1, b2, c1,
                         c2, 1e-8f);
    }
    forward(c, m, n, w);
    backward(c, m, n, w);
    hipLaunchKernelGGL(restore_best_kernel, pg, dim3(256), 0, c->stream, w.loss, best, m->p, bp, np_);
    forward(c, m, n, w);  // the loss of the parameters kept
    backward(c, m, n, w);
    float l = 0.f;
    hip_check(hipMemcpyAsync(&l, w.loss, 4, hipMemcpyDeviceToHost, c->stream), "loss");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    if (loss_out) *loss_out = l;
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}

}  // extern "C"

namespace tta
{
// Forward pass on n feature rows already in device memory; predictions to
// the host (throws tta::error).
void mlp_predict_dev(tta_ctx* c, tta_mlp* m, const float* X_dev, int n, float* y_host)
{
  if (n <= 0) return;
  ws_owner O(c);
  mlp_ws w;
  make_ws(c, m, n, O, w, false);
  w.act[0] = const_cast<float*>(X_dev);
  if (m->bf16_infer)
    forward_bf16(c, m, n, w);
  else
    forward(c, m, n, w);
  hip_check(hipMemcpyAsync(y_host, w.act.back(), (size_t)n * 4, hipMemcpyDeviceToHost, c->stream), "y");
  ctx_sync(c);
}
}  // namespace tta

extern "C"
{

int tta_mlp_set_bf16_inference(tta_ctx* c, tta_mlp* m, int on)
{
  if (!c || !m) return TTA_ERR_ARG;
  m->bf16_infer = on != 0;
  return TTA_OK;
}

int tta_mlp_predict(tta_ctx* c, tta_mlp* m, const float* X, int n, float* y_out)
{
  if (!c || !m || !X || !y_out || n < 0) return TTA_ERR_ARG;
  if (n == 0) return TTA_OK;
  try
  {
    ws_owner O(c);
    mlp_ws w;
    make_ws(c, m, n, O, w, false);
    hip_check(hipMemcpyAsync(w.act[0], X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, c->stream), "X");
    if (m->bf16_infer)
      forward_bf16(c, m, n, w);
    else
      forward(c, m, n, w);
    hip_check(hipMemcpyAsync(y_out, w.act.back(), (size_t)n * 4, hipMemcpyDeviceToHost, c->stream), "y");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  }
  catch (const tta::error& e)
  {
    c->last_error = e.what();
    return e.code;
  }
  return TTA_OK;
}

}  // extern "C"
