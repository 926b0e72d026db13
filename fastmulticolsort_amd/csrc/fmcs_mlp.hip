// fmcs_mlp.hip -- the learned plan-cost model: a small fp32 MLP
// (nin -> h1 -> h2 -> 1, ReLU) trained and evaluated by hand-written kernels.
// It predicts log(ms) of a massaging plan from request and plan features;
// the planner picks the argmin.  No counterpart in the reference, whose
// plan costs are analytic (src/stitch_composer.cpp:2186-2378).
//
// Parameter layout (floats): W1[h1][nin], b1[h1], W2[h2][h1], b2[h2],
// W3[h2], b3.  Backward is deterministic: each block writes the gradient
// sums of its samples to its own slot and a second kernel adds the slots
// in block order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmcs.h"
#include "../../include/fmcs_plan.h"
#include <initializer_list>
#include <utility>

#include "fmcs_internal.h"

namespace fmcs {
namespace {

constexpr int MLP_MAX = 64;   // max nin, h1, h2
constexpr int FWD_NT = 64;    // one sample per thread
constexpr int BWD_SB = 16;    // samples per backward block
constexpr int BWD_NT = 256;

struct Shape {
  uint32_t nin, h1, h2;
  __host__ __device__ uint32_t oW1() const { return 0; }
  __host__ __device__ uint32_t ob1() const { return h1 * nin; }
  __host__ __device__ uint32_t oW2() const { return ob1() + h1; }
  __host__ __device__ uint32_t ob2() const { return oW2() + h2 * h1; }
  __host__ __device__ uint32_t oW3() const { return ob2() + h2; }
  __host__ __device__ uint32_t ob3() const { return oW3() + h2; }
  __host__ __device__ uint32_t count() const { return ob3() + 1; }
};

__global__ __launch_bounds__(FWD_NT) void k_mlp_forward(Shape sh, const float* __restrict__ prm,
                                                        const float* __restrict__ x, uint32_t batch,
                                                        float* __restrict__ y, float* __restrict__ acts) {
  __shared__ float sW1[MLP_MAX * MLP_MAX];
  __shared__ float sW2[MLP_MAX * MLP_MAX];
  __shared__ float sb1[MLP_MAX], sb2[MLP_MAX], sW3[MLP_MAX];
  for (uint32_t i = threadIdx.x; i < sh.h1 * sh.nin; i += FWD_NT) sW1[i] = prm[sh.oW1() + i];
  for (uint32_t i = threadIdx.x; i < sh.h2 * sh.h1; i += FWD_NT) sW2[i] = prm[sh.oW2() + i];
  for (uint32_t i = threadIdx.x; i < sh.h1; i += FWD_NT) sb1[i] = prm[sh.ob1() + i];
  for (uint32_t i = threadIdx.x; i < sh.h2; i += FWD_NT) {
    sb2[i] = prm[sh.ob2() + i];
    sW3[i] = prm[sh.oW3() + i];
  }
  __syncthreads();
  const uint32_t s = blockIdx.x * FWD_NT + threadIdx.x;
  if (s >= batch) return;
  float xin[MLP_MAX], h1[MLP_MAX];
  for (uint32_t k = 0; k < sh.nin; ++k) xin[k] = x[(uint64_t)s * sh.nin + k];
  for (uint32_t i = 0; i < sh.h1; ++i) {
    float a = sb1[i];
    for (uint32_t k = 0; k < sh.nin; ++k) a += sW1[i * sh.nin + k] * xin[k];
    h1[i] = a > 0.f ? a : 0.f;
    if (acts) acts[(uint64_t)s * (sh.h1 + sh.h2) + i] = h1[i];
  }
  float out = prm[sh.ob3()];
  for (uint32_t j = 0; j < sh.h2; ++j) {
    float a = sb2[j];
    for (uint32_t i = 0; i < sh.h1; ++i) a += sW2[j * sh.h1 + i] * h1[i];
    float h = a > 0.f ? a : 0.f;
    if (acts) acts[(uint64_t)s * (sh.h1 + sh.h2) + sh.h1 + j] = h;
    out += sW3[j] * h;
  }
  y[s] = out;
}

// Gradients of L = sum_s 0.5 (y_s - t_s)^2 / batch for the block's samples,
// written to partial[blockIdx.x][0..P) and the loss to partial[..][P].
__global__ __launch_bounds__(BWD_NT) void k_mlp_backward(Shape sh, const float* __restrict__ prm,
                                                         const float* __restrict__ x,
                                                         const float* __restrict__ acts,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ t, uint32_t batch,
                                                         float* __restrict__ partial) {
  __shared__ float sW2[MLP_MAX * MLP_MAX];
  __shared__ float sW3[MLP_MAX];
  __shared__ float sx[BWD_SB][MLP_MAX];
  __shared__ float sh1[BWD_SB][MLP_MAX];
  __shared__ float sh2[BWD_SB][MLP_MAX];
  __shared__ float sdz2[BWD_SB][MLP_MAX];
  __shared__ float sdz1[BWD_SB][MLP_MAX];
  __shared__ float sdy[BWD_SB];
  __shared__ float sloss[BWD_SB];
  const uint32_t tid = threadIdx.x;
  const uint32_t s0 = blockIdx.x * BWD_SB;
  const uint32_t ns = min((uint32_t)BWD_SB, batch - s0);
  const uint32_t H = sh.h1 + sh.h2;
  for (uint32_t i = tid; i < sh.h2 * sh.h1; i += BWD_NT) sW2[i] = prm[sh.oW2() + i];
  for (uint32_t i = tid; i < sh.h2; i += BWD_NT) sW3[i] = prm[sh.oW3() + i];
  for (uint32_t i = tid; i < BWD_SB * MLP_MAX; i += BWD_NT) {
    uint32_t s = i / MLP_MAX, k = i % MLP_MAX;
    bool ok = s < ns;
    sx[s][k] = (ok && k < sh.nin) ? x[(uint64_t)(s0 + s) * sh.nin + k] : 0.f;
    sh1[s][k] = (ok && k < sh.h1) ? acts[(uint64_t)(s0 + s) * H + k] : 0.f;
    sh2[s][k] = (ok && k < sh.h2) ? acts[(uint64_t)(s0 + s) * H + sh.h1 + k] : 0.f;
  }
  if (tid < BWD_SB) {
    float d = 0.f, l = 0.f;
    if (tid < ns) {
      float e = y[s0 + tid] - t[s0 + tid];
      d = e / (float)batch;
      l = 0.5f * e * e / (float)batch;
    }
    sdy[tid] = d;
    sloss[tid] = l;
  }
  __syncthreads();
  // dz2 = dy * W3 * [h2 > 0]
  for (uint32_t i = tid; i < BWD_SB * MLP_MAX; i += BWD_NT) {
    uint32_t s = i / MLP_MAX, j = i % MLP_MAX;
    sdz2[s][j] = (j < sh.h2 && sh2[s][j] > 0.f) ? sdy[s] * sW3[j] : 0.f;
  }
  __syncthreads();
  // dz1 = (W2^T dz2) * [h1 > 0]
  for (uint32_t i = tid; i < BWD_SB * MLP_MAX; i += BWD_NT) {
    uint32_t s = i / MLP_MAX, k = i % MLP_MAX;
    float a = 0.f;
    if (k < sh.h1 && sh1[s][k] > 0.f)
      for (uint32_t j = 0; j < sh.h2; ++j) a += sW2[j * sh.h1 + k] * sdz2[s][j];
    sdz1[s][k] = a;
  }
  __syncthreads();
  float* out = partial + (uint64_t)blockIdx.x * (sh.count() + 1);
  // dW1, db1
  for (uint32_t p = tid; p < sh.h1 * sh.nin; p += BWD_NT) {
    uint32_t i = p / sh.nin, k = p % sh.nin;
    float a = 0.f;
    for (uint32_t s = 0; s < BWD_SB; ++s) a += sdz1[s][i] * sx[s][k];
    out[sh.oW1() + p] = a;
  }
  for (uint32_t i = tid; i < sh.h1; i += BWD_NT) {
    float a = 0.f;
    for (uint32_t s = 0; s < BWD_SB; ++s) a += sdz1[s][i];
    out[sh.ob1() + i] = a;
  }
  // dW2, db2
  for (uint32_t p = tid; p < sh.h2 * sh.h1; p += BWD_NT) {
    uint32_t j = p / sh.h1, i = p % sh.h1;
    float a = 0.f;
    for (uint32_t s = 0; s < BWD_SB; ++s) a += sdz2[s][j] * sh1[s][i];
    out[sh.oW2() + p] = a;
  }
  for (uint32_t j = tid; j < sh.h2; j += BWD_NT) {
    float a = 0.f, b = 0.f;
    for (uint32_t s = 0; s < BWD_SB; ++s) {
      a += sdz2[s][j];
      b += sdy[s] * sh2[s][j];
    }
    out[sh.ob2() + j] = a;
    out[sh.oW3() + j] = b;
  }
  if (tid == 0) {
    float a = 0.f, l = 0.f;
    for (uint32_t s = 0; s < BWD_SB; ++s) {
      a += sdy[s];
      l += sloss[s];
    }
    out[sh.ob3()] = a;
    out[sh.count()] = l;
  }
}

// out[p] = sum over blocks b (in order) of partial[b * stride + p], p < count
__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t nblocks, uint32_t stride, uint32_t count,
                             float* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  float a = 0.f;
  for (uint32_t b = 0; b < nblocks; ++b) a += partial[(uint64_t)b * stride + p];
  out[p] = a;
}

__global__ void k_mlp_adam(float* __restrict__ prm, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float gp = g[p];
  float mp = b1 * m[p] + (1.f - b1) * gp;
  float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  prm[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

bool shape_ok(const fmcs_mlp_shape* s) {
  return s && s->nin >= 1 && s->nin <= MLP_MAX && s->h1 >= 1 && s->h1 <= MLP_MAX && s->h2 >= 1 && s->h2 <= MLP_MAX;
}

}  // namespace
}  // namespace fmcs

using namespace fmcs;

namespace {

// The network's buffers are device memory the caller trains in place: they
// are checked (no staging) before any launch.
int check_bufs(fmcs_ctx* ctx, std::initializer_list<std::pair<const void*, const char*>> bufs) {
  internal::Stager st(ctx, false);
  for (const auto& b : bufs) {
    void* p = const_cast<void*>(b.first);
    const int rc = st.arg(p, 1, internal::Stager::IN, b.second);
    if (rc != FMCS_OK) return rc;
  }
  return FMCS_OK;
}

}  // namespace

extern "C" {

int fmcs_mlp_param_count(const fmcs_mlp_shape* s, uint32_t* count) {
  if (!shape_ok(s) || !count) return FMCS_E_ARG;
  Shape sh{s->nin, s->h1, s->h2};
  *count = sh.count();
  return FMCS_OK;
}

int fmcs_mlp_forward(fmcs_ctx* ctx, const fmcs_mlp_shape* s, const float* params_d, const float* x_d,
                     uint32_t batch, float* y_d, float* acts_d) {
  if (!ctx || !shape_ok(s) || !params_d || !x_d || !y_d) return FMCS_E_ARG;
  if (batch == 0) return FMCS_OK;
  const int crc = check_bufs(ctx, {{params_d, "params_d"}, {x_d, "x_d"}, {y_d, "y_d"}, {acts_d, "acts_d"}});
  if (crc != FMCS_OK) return crc;
  Shape sh{s->nin, s->h1, s->h2};
  hipLaunchKernelGGL(k_mlp_forward, dim3((batch + FWD_NT - 1) / FWD_NT), dim3(FWD_NT), 0,
                     internal::stream_of(ctx), sh, params_d, x_d, batch, y_d, acts_d);
  if (hipGetLastError() != hipSuccess) return internal::set_error(ctx, FMCS_E_HIP, "k_mlp_forward launch");
  return FMCS_OK;
}

int fmcs_mlp_backward(fmcs_ctx* ctx, const fmcs_mlp_shape* s, const float* params_d, const float* x_d,
                      const float* acts_d, const float* y_d, const float* target_d, uint32_t batch, float* grad_d,
                      float* loss_d) {
  if (!ctx || !shape_ok(s) || !params_d || !x_d || !acts_d || !y_d || !target_d || !grad_d || batch == 0)
    return FMCS_E_ARG;
  const int crc = check_bufs(ctx, {{params_d, "params_d"}, {x_d, "x_d"}, {acts_d, "acts_d"}, {y_d, "y_d"},
                                   {target_d, "target_d"}, {grad_d, "grad_d"}, {loss_d, "loss_d"}});
  if (crc != FMCS_OK) return crc;
  Shape sh{s->nin, s->h1, s->h2};
  const uint32_t width = sh.count() + 1;
  const uint32_t nblk = (batch + BWD_SB - 1) / BWD_SB;
  void* part;
  int rc = internal::scratch(ctx, 23, (size_t)nblk * width * sizeof(float), &part);
  if (rc != FMCS_OK) return rc;
  hipStream_t st = internal::stream_of(ctx);
  hipLaunchKernelGGL(k_mlp_backward, dim3(nblk), dim3(BWD_NT), 0, st, sh, params_d, x_d, acts_d, y_d, target_d,
                     batch, (float*)part);
  if (hipGetLastError() != hipSuccess) return internal::set_error(ctx, FMCS_E_HIP, "k_mlp_backward launch");
  // gradient (P words) to grad_d; the loss word goes to loss_d when given
  hipLaunchKernelGGL(k_mlp_reduce, dim3((width - 1 + 255) / 256), dim3(256), 0, st, (const float*)part, nblk,
                     width, width - 1, grad_d);
  if (loss_d)
    hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(64), 0, st, (const float*)part + (width - 1), nblk, width, 1u,
                       loss_d);
  if (hipGetLastError() != hipSuccess) return internal::set_error(ctx, FMCS_E_HIP, "k_mlp_reduce launch");
  return FMCS_OK;
}

int fmcs_mlp_adam(fmcs_ctx* ctx, uint32_t nparams, float* params_d, const float* grad_d, float* m_d, float* v_d,
                  float lr, float beta1, float beta2, float eps, uint32_t step) {
  if (!ctx || !params_d || !grad_d || !m_d || !v_d || step == 0) return FMCS_E_ARG;
  if (nparams == 0) return FMCS_OK;
  const int crc = check_bufs(ctx, {{params_d, "params_d"}, {grad_d, "grad_d"}, {m_d, "m_d"}, {v_d, "v_d"}});
  if (crc != FMCS_OK) return crc;
  float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((nparams + 255) / 256), dim3(256), 0, internal::stream_of(ctx), params_d,
                     grad_d, m_d, v_d, nparams, lr, beta1, beta2, eps, c1, c2);
  if (hipGetLastError() != hipSuccess) return internal::set_error(ctx, FMCS_E_HIP, "k_mlp_adam launch");
  return FMCS_OK;
}

}  // extern "C"
