// Learned splitter network: a 65 -> 32 (ReLU) -> 1 (sigmoid) fp32 MLP.
//
// Inputs (built by the first-pass splitter from the sorted sample): a window
// of 64 consecutive sample keys around the target quantile, normalised to
// [0, 1] by the window's extremes, and the target's fractional position in the
// window.  Output: where between the two sample keys that bracket the target
// the true quantile lies.  Trained on exact k-th values produced by the plain
// method; the first pass builds its ranges from the predictions and the
// linear slope inside each range does the rest, so a poor prediction only
// costs work, never exactness.
//
// Parameter layout (row-major fp32): W1[32][65], b1[32], W2[32], b2.
//
// Kernels: batched forward; loss + gradient in two steps (per-sample
// back-propagation to the hidden layer, then one thread per parameter summing
// over a chunk of samples); an Adam step.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "ggms.h"

namespace ggms {

constexpr int MI = 65, MH = 32, MP = MH * MI + MH + MH + 1;

namespace {
std::mutex g_w_mu;
std::vector<float*> g_w;  // per device
}  // namespace

const float* learned_weights_device(int device) {
  std::lock_guard<std::mutex> lk(g_w_mu);
  if (device < 0 || device >= (int)g_w.size()) return nullptr;
  return g_w[device];
}

__device__ __forceinline__ float sigmoidf_(float o) { return 1.0f / (1.0f + __expf(-o)); }

__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ w, const float* __restrict__ x,
                                                     uint32_t nb, float* __restrict__ out) {
  __shared__ float sw[MP];
  for (int i = threadIdx.x; i < MP; i += blockDim.x) sw[i] = w[i];
  __syncthreads();
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const float* xb = x + (size_t)b * MI;
  float in[MI];
  for (int i = 0; i < MI; ++i) in[i] = xb[i];
  const float* W1 = sw;
  const float* b1 = sw + MH * MI;
  const float* W2 = b1 + MH;
  float o = W2[MH];
  for (int h = 0; h < MH; ++h) {
    float z = b1[h];
    for (int i = 0; i < MI; ++i) z += W1[h * MI + i] * in[i];
    o += W2[h] * fmaxf(z, 0.0f);
  }
  out[b] = sigmoidf_(o);
}

// Backward, step 1 (one thread per sample): forward pass, then the per-sample
// quantities the weight gradient needs: dz[b][h] = dL/dz_h (zero where the
// ReLU is off), a[b][h] = relu(z_h), d_o[b] = dL/do.  Loss is the mean of
// 0.5 (sigmoid(o) - y)^2 over the batch.
__global__ void __launch_bounds__(256) k_mlp_backprop(const float* __restrict__ w, const float* __restrict__ x,
                                                      const float* __restrict__ y, uint32_t nb,
                                                      float* __restrict__ dz, float* __restrict__ act,
                                                      float* __restrict__ d_out, float* __restrict__ loss) {
  __shared__ float sw[MP];
  __shared__ float sl;
  for (int i = threadIdx.x; i < MP; i += blockDim.x) sw[i] = w[i];
  if (threadIdx.x == 0) sl = 0.0f;
  __syncthreads();
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    const float* xb = x + (size_t)b * MI;
    float in[MI];
    for (int i = 0; i < MI; ++i) in[i] = xb[i];
    const float* W1 = sw;
    const float* b1 = sw + MH * MI;
    const float* W2 = b1 + MH;
    float z[MH];
    float o = W2[MH];
    for (int h = 0; h < MH; ++h) {
      float s = b1[h];
      for (int i = 0; i < MI; ++i) s += W1[h * MI + i] * in[i];
      z[h] = s;
      o += W2[h] * fmaxf(s, 0.0f);
    }
    float p = sigmoidf_(o);
    float e = p - y[b];
    atomicAdd(&sl, 0.5f * e * e / (float)nb);
    float d_o = e * p * (1.0f - p) / (float)nb;
    d_out[b] = d_o;
    for (int h = 0; h < MH; ++h) {
      act[(size_t)b * MH + h] = fmaxf(z[h], 0.0f);
      dz[(size_t)b * MH + h] = z[h] > 0.0f ? d_o * W2[h] : 0.0f;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, sl);
}

// Backward, step 2 (one thread per parameter, one workgroup row per chunk of
// samples): grad_W1 = sum_b dz_b x_b^T, grad_b1 = sum_b dz_b,
// grad_W2 = sum_b d_o_b a_b, grad_b2 = sum_b d_o_b.
constexpr uint32_t kGradChunk = 1024;
__global__ void __launch_bounds__(256) k_mlp_wgrad(const float* __restrict__ x, const float* __restrict__ dz,
                                                   const float* __restrict__ act, const float* __restrict__ d_out,
                                                   uint32_t nb, float* __restrict__ grad) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)MP) return;
  uint32_t b0 = blockIdx.y * kGradChunk;
  uint32_t b1 = b0 + kGradChunk < nb ? b0 + kGradChunk : nb;
  float acc = 0.0f;
  if (p < (uint32_t)(MH * MI)) {
    uint32_t h = p / MI, i = p % MI;
    for (uint32_t b = b0; b < b1; ++b) acc += dz[(size_t)b * MH + h] * x[(size_t)b * MI + i];
  } else if (p < (uint32_t)(MH * MI + MH)) {
    uint32_t h = p - MH * MI;
    for (uint32_t b = b0; b < b1; ++b) acc += dz[(size_t)b * MH + h];
  } else if (p < (uint32_t)(MH * MI + 2 * MH)) {
    uint32_t h = p - MH * MI - MH;
    for (uint32_t b = b0; b < b1; ++b) acc += d_out[b] * act[(size_t)b * MH + h];
  } else {
    for (uint32_t b = b0; b < b1; ++b) acc += d_out[b];
  }
  atomicAdd(&grad[p], acc);
}

__global__ void k_fill_f32(float* __restrict__ p, uint32_t n, float v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace ggms

using namespace ggms;

extern "C" {

uint32_t ggms_mlp_param_count(void) { return (uint32_t)MP; }

ggms_status ggms_set_learned_weights(int device, const float* w_host, uint32_t n) {
  if (!w_host || n != (uint32_t)MP) return GGMS_EINVAL;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return GGMS_EDEVICE;
  std::lock_guard<std::mutex> lk(g_w_mu);
  if ((int)g_w.size() <= device) g_w.resize(device + 1, nullptr);
  if (hipSetDevice(device) != hipSuccess) return GGMS_EDEVICE;
  ggms_status s = GGMS_OK;
  if (!g_w[device] && hipMalloc(&g_w[device], sizeof(float) * MP) != hipSuccess) {
    g_w[device] = nullptr;
    s = GGMS_ENOMEM;
  }
  if (s == GGMS_OK && hipMemcpy(g_w[device], w_host, sizeof(float) * MP, hipMemcpyHostToDevice) != hipSuccess)
    s = GGMS_EDEVICE;
  (void)hipSetDevice(cur);
  return s;
}

ggms_status ggms_clear_learned_weights(int device) {
  std::lock_guard<std::mutex> lk(g_w_mu);
  if (device < 0 || device >= (int)g_w.size() || !g_w[device]) return GGMS_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(device);
  (void)hipFree(g_w[device]);
  g_w[device] = nullptr;
  (void)hipSetDevice(cur);
  return GGMS_OK;
}

ggms_status ggms_mlp_forward(const float* d_w, const float* d_x, uint32_t nb, float* d_out, void* stream) {
  if (!d_w || (!d_x && nb) || (!d_out && nb)) return GGMS_EINVAL;
  if (nb == 0) return GGMS_OK;
  k_mlp_forward<<<(nb + 255) / 256, 256, 0, (hipStream_t)stream>>>(d_w, d_x, nb, d_out);
  return hipGetLastError() == hipSuccess ? GGMS_OK : GGMS_EDEVICE;
}

uint64_t ggms_mlp_workspace_floats(uint32_t nb) { return (uint64_t)nb * (2 * MH + 1); }

ggms_status ggms_mlp_loss_grad(const float* d_w, const float* d_x, const float* d_y, uint32_t nb, float* d_grad,
                               float* d_loss, float* d_work, void* stream) {
  if (!d_w || !d_grad || !d_loss || (nb && (!d_x || !d_y || !d_work))) return GGMS_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  k_fill_f32<<<(MP + 255) / 256, 256, 0, st>>>(d_grad, MP, 0.0f);
  k_fill_f32<<<1, 32, 0, st>>>(d_loss, 1, 0.0f);
  if (nb) {
    float* dz = d_work;
    float* act = d_work + (size_t)nb * MH;
    float* dout = act + (size_t)nb * MH;
    k_mlp_backprop<<<(nb + 255) / 256, 256, 0, st>>>(d_w, d_x, d_y, nb, dz, act, dout, d_loss);
    dim3 g((MP + 255) / 256, (nb + kGradChunk - 1) / kGradChunk);
    k_mlp_wgrad<<<g, 256, 0, st>>>(d_x, dz, act, dout, nb, d_grad);
  }
  return hipGetLastError() == hipSuccess ? GGMS_OK : GGMS_EDEVICE;
}

ggms_status ggms_mlp_adam_step(float* d_w, const float* d_grad, float* d_m, float* d_v, uint32_t step, float lr,
                               float beta1, float beta2, float eps, void* stream) {
  if (!d_w || !d_grad || !d_m || !d_v || step == 0) return GGMS_EINVAL;
  float c1 = 1.0f - powf(beta1, (float)step);
  float c2 = 1.0f - powf(beta2, (float)step);
  k_adam<<<(MP + 255) / 256, 256, 0, (hipStream_t)stream>>>(d_w, d_grad, d_m, d_v, MP, lr, beta1, beta2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? GGMS_OK : GGMS_EDEVICE;
}

}  // extern "C"
