// Learned-closure kernels: feature extraction, batched MLP inference, one training step (forward,
// backward and weight gradients on BF16 MFMA with FP32 accumulation) and an Adam update that
// re-packs the BF16 weights.  The fused inference inside the cumulant sweep uses the same
// mlp_forward_wave (closure.h).
//
// Training loss over n samples:  L = (1/n) sum (softplus(o) - y)^2,  y = nu_t / nu_scale.
// Master parameters (FP32, flat): w1t[32][16] | w2t[32][32] | w3[32] | b1[32] | b2[32] | b3.
#include <string>

#include "lbm_common.h"
#include "closure.h"

namespace lbm355 {

constexpr int OFF_W1T = 0, OFF_W2T = 512, OFF_W3 = 1536, OFF_B1 = 1568, OFF_B2 = 1600, OFF_B3 = 1632;
constexpr int MASTER_SIZE = 1633;

__global__ void __launch_bounds__(256) features_kernel(const float* __restrict__ pi6, const float* __restrict__ rho,
                                                       int64_t n, float* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float p[6] = {pi6[6 * i], pi6[6 * i + 1], pi6[6 * i + 2], pi6[6 * i + 3], pi6[6 * i + 4], pi6[6 * i + 5]};
    float f[CLOSURE_IN];
    closure_features(p, rho[i], f);
#pragma unroll
    for (int k = 0; k < CLOSURE_IN; ++k) x[CLOSURE_IN * i + k] = f[k];
  }
}

// persistent waves over 64-sample tiles; lanes past n compute on zeros and store nothing
__global__ void __launch_bounds__(256) forward_kernel(const ClosureParams* __restrict__ P, const float* __restrict__ X,
                                                      int64_t n, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t ntiles = (n + 63) / 64;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t i = tile * 64 + lane;
    const bool valid = i < n;
    float x[CLOSURE_IN];
#pragma unroll
    for (int k = 0; k < CLOSURE_IN; ++k) x[k] = valid ? X[CLOSURE_IN * i + k] : 0.f;
    const float o = mlp_forward_wave(P, x, nullptr, nullptr);
    if (valid) out[i] = o;
  }
}

__global__ void zero_kernel(float* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// one wavefront per block; gradients accumulate in registers over the block's tiles and are added
// to `grad` (FP32 atomics) at the end
__global__ void __launch_bounds__(64) train_kernel(const ClosureParams* __restrict__ P, const float* __restrict__ X,
                                                   const float* __restrict__ Y, int64_t n, float inv_n,
                                                   float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float s_a0[64][CLOSURE_IN_PAD];
  __shared__ float s_a1[64][CLOSURE_H];
  __shared__ float s_h2[64][CLOSURE_H];
  __shared__ float s_d2[64][CLOSURE_H];
  __shared__ float s_d1[64][CLOSURE_H];
  __shared__ float s_g[64];
  const int lane = threadIdx.x;
  const int half = lane >> 5, col = lane & 31;
  v16f gW1, gW2;
#pragma unroll
  for (int v = 0; v < 16; ++v) gW1[v] = gW2[v] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, lsum = 0.f;
  // W2 as an A operand [k1][k2] = w2t[k2][k1], per K block
  v4s w2ta[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int t = 0; t < 4; ++t) w2ta[kb][t] = (short)P->w2t[8 * kb + 4 * half + t][col];
  float w3r[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) w3r[v] = bf2f(P->w3[c_row(v, lane)]);

  const int64_t ntiles = (n + 63) / 64;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * 64 + lane;
    const bool valid = i < n;
    float x[CLOSURE_IN];
#pragma unroll
    for (int k = 0; k < CLOSURE_IN; ++k) x[k] = valid ? X[CLOSURE_IN * i + k] : 0.f;
    const float y = valid ? Y[i] : 0.f;
    float h1[2][16], h2[2][16];
    const float o = mlp_forward_wave(P, x, h1, h2);
    const float diff = softplus(o) - y;
    const float g = valid ? 2.f * inv_n * diff * sigmoid(o) : 0.f;
    lsum += valid ? diff * diff : 0.f;
#pragma unroll
    for (int f = 0; f < CLOSURE_IN_PAD; ++f)
      s_a0[lane][f] = f < CLOSURE_IN ? bf2f(f2bf((x[f] - P->in_shift[f]) * P->in_scale[f])) : 0.f;
    s_g[lane] = g;
    const float g_other = __shfl_xor(g, 32, 64);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float gc = (c == half) ? g : g_other;        // dL/do of cell 32c + col
      const int cell = 32 * c + col;
      v4s db[4];
      float d2[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        d2[v] = gc * w3r[v] * (1.f - h2[c][v] * h2[c][v]);
        const int row = c_row(v, lane);
        s_a1[cell][row] = bf2f(f2bf(h1[c][v]));
        s_h2[cell][row] = h2[c][v];
        s_d2[cell][row] = d2[v];
        db[v >> 2][v & 3] = (short)f2bf(d2[v]);
      }
      v16f acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w2ta[kb], db[kb], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) s_d1[cell][c_row(v, lane)] = acc[v] * (1.f - h1[c][v] * h1[c][v]);
    }
    __syncthreads();
    // weight gradients: K = the 64 cells of the tile, 8 blocks of 8
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
      v4s ad2, ba1, ad1, ba0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = 8 * kb + 4 * half + t;
        ad2[t] = (short)f2bf(s_d2[k][col]);
        ba1[t] = (short)f2bf(s_a1[k][col]);
        ad1[t] = (short)f2bf(s_d1[k][col]);
        ba0[t] = col < CLOSURE_IN_PAD ? (short)f2bf(s_a0[k][col]) : (short)0;
      }
      gW2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ad2, ba1, gW2, 0, 0, 0);
      gW1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ad1, ba0, gW1, 0, 0, 0);
    }
    if (lane < 32) {
      for (int k = 0; k < 64; ++k) {
        gb1 += s_d1[k][lane];
        gb2 += s_d2[k][lane];
        gw3 += s_g[k] * s_h2[k][lane];
      }
    } else if (lane == 32) {
      for (int k = 0; k < 64; ++k) gb3 += s_g[k];
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = c_row(v, lane);
    atomicAdd(&grad[OFF_W2T + row * CLOSURE_H + col], gW2[v]);
    if (col < CLOSURE_IN_PAD) atomicAdd(&grad[OFF_W1T + row * CLOSURE_IN_PAD + col], gW1[v]);
  }
  if (lane < 32) {
    atomicAdd(&grad[OFF_B1 + lane], gb1);
    atomicAdd(&grad[OFF_B2 + lane], gb2);
    atomicAdd(&grad[OFF_W3 + lane], gw3);
  } else if (lane == 32) {
    atomicAdd(&grad[OFF_B3], gb3);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if (lane == 0) atomicAdd(loss, lsum * inv_n);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, then BF16 re-pack into the device block.
// The padded layer-1 inputs 10..15 stay zero.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float beta1, float beta2, float eps, float bc1, float bc2,
                            ClosureParams* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MASTER_SIZE) return;
  float gi = g[i];
  if (i < OFF_W2T && (i % CLOSURE_IN_PAD) >= CLOSURE_IN) gi = 0.f;
  const float mi = beta1 * m[i] + (1.f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  if (i < OFF_W2T && (i % CLOSURE_IN_PAD) >= CLOSURE_IN) wi = 0.f;
  w[i] = wi;
  if (i < OFF_W2T) P->w1t[i / CLOSURE_IN_PAD][i % CLOSURE_IN_PAD] = f2bf(wi);
  else if (i < OFF_W3) P->w2t[(i - OFF_W2T) / CLOSURE_H][(i - OFF_W2T) % CLOSURE_H] = f2bf(wi);
  else if (i < OFF_B1) P->w3[i - OFF_W3] = f2bf(wi);
  else if (i < OFF_B2) P->b1[i - OFF_B1] = wi;
  else if (i < OFF_B3) P->b2[i - OFF_B2] = wi;
  else P->b3 = wi;
}

}  // namespace lbm355

using namespace lbm355;

namespace {
int cerr(int code, const std::string& msg) { return lbm355_set_error(code, msg.c_str()); }
int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cerr(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}
int grid_for(int64_t n, int per_block) {
  const int64_t b = (n + per_block - 1) / per_block;
  return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}
}  // namespace

extern "C" {

int64_t lbm355_closure_params_size(void) { return (int64_t)sizeof(ClosureParams); }
int32_t lbm355_closure_master_size(void) { return MASTER_SIZE; }

int lbm355_closure_features(const float* pi_neq, const float* density, int64_t n, float* features, void* stream) {
  if (!pi_neq || !density || !features || n < 0) return cerr(LBM355_ERR_ARG, "bad closure_features arguments");
  if (n == 0) return LBM355_OK;
  hipLaunchKernelGGL(features_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, pi_neq, density, n,
                     features);
  return launched();
}

int lbm355_closure_forward(const void* params, const float* features, int64_t n, float* out, void* stream) {
  if (!params || !features || !out || n < 0) return cerr(LBM355_ERR_ARG, "bad closure_forward arguments");
  if (n == 0) return LBM355_OK;
  hipLaunchKernelGGL(forward_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const ClosureParams*)params, features, n, out);
  return launched();
}

int lbm355_closure_train_step(const void* params, const float* features, const float* targets, int64_t n,
                              float* grad, float* loss, int32_t blocks, void* stream) {
  if (!params || !features || !targets || !grad || !loss || n <= 0 || blocks <= 0)
    return cerr(LBM355_ERR_ARG, "bad closure_train_step arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_kernel, dim3((MASTER_SIZE + 255) / 256), dim3(256), 0, s, grad, MASTER_SIZE);
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, loss, 1);
  const int64_t tiles = (n + 63) / 64;
  const int grid = (int)(blocks < tiles ? blocks : tiles);
  hipLaunchKernelGGL(train_kernel, dim3(grid), dim3(64), 0, s, (const ClosureParams*)params, features, targets, n,
                     1.f / (float)n, grad, loss);
  return launched();
}

int lbm355_closure_adam(float* master, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                        float eps, int32_t step, void* params, void* stream) {
  if (!master || !grad || !m || !v || !params || step < 1) return cerr(LBM355_ERR_ARG, "bad closure_adam arguments");
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((MASTER_SIZE + 255) / 256), dim3(256), 0, (hipStream_t)stream, master, grad, m,
                     v, lr, beta1, beta2, eps, bc1, bc2, (ClosureParams*)params);
  return launched();
}

}  // extern "C"
