// Training and inference kernels of the refinement-prediction network (mlp.hpp), and their
// C-ABI (include/agimp.h, agimp_mlp_*).  All pointers are device pointers; the caller owns
// the memory and orders the calls on `stream` (NULL: the default stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "agimp.h"
#include "common.hpp"
#include "mlp.hpp"

namespace agimp {

constexpr int MLP_T = 256;  // 4 wavefronts per workgroup

__global__ void __launch_bounds__(MLP_T) k_mlp_forward(const float* __restrict__ X, int64_t rows,
                                                       const float* __restrict__ params, float* __restrict__ Z) {
  __shared__ __align__(16) MlpLds W;
  mlp_load(W, params);
  __syncthreads();
  const int lane = __lane_id();
  const int64_t ntiles = (rows + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * (MLP_T / 64) + threadIdx.x / 64;
  const int64_t nwaves = (int64_t)gridDim.x * (MLP_T / 64);
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    MlpTile t;
    mlp_forward_tile(W, X, rows, tile * 16, t);
    const int64_t row = tile * 16 + lane;
    if (lane < 16 && row < rows) *reinterpret_cast<float4*>(Z + row * MLP_OUT) = make_float4(t.z[0], t.z[1], t.z[2], t.z[3]);
  }
}

// per-wavefront LDS for the row-major -> K-major transposes of the weight-gradient products;
// element (unit h, row r) of a [units][16 rows] BF16 matrix lives at h*16 + r
struct WaveScratch {
  uint16_t x[MLP_IN * 16];
  uint16_t h1[MLP_H * 16];
  uint16_t h2[MLP_H * 16];
  uint16_t d2[MLP_H * 16];
  uint16_t d1[MLP_H * 16];
  uint16_t dz[16 * 16];  // class rows 4..15 stay zero
};

// lane layout of an accumulator / B operand: (unit 16*mb + 4*(lane/16) + j, row lane%16)
__device__ __forceinline__ void stash(uint16_t* S, int mb, bf16x4 v) {
  const int lane = __lane_id();
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) S[(mb * 16 + 4 * g + j) * 16 + n] = (uint16_t)v[j];
}

// Mean (weighted) softmax cross-entropy and its gradient with respect to the parameters:
//   loss = scale * sum_rows cw[y] * (-log softmax(z)[y]),  grad likewise.
// grad [5508] and loss [1] are accumulated with atomics (zeroed by the caller).
__global__ void __launch_bounds__(MLP_T) k_mlp_grad(const float* __restrict__ X, const int* __restrict__ Y,
                                                    const int* __restrict__ idx, int64_t rows,
                                                    const float* __restrict__ params,
                                                    const float* __restrict__ cw, float scale,
                                                    float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ __align__(16) MlpLds W;
  __shared__ __align__(16) WaveScratch SS[MLP_T / 64];
  mlp_load(W, params);
  for (int i = threadIdx.x; i < (int)(sizeof(WaveScratch) * (MLP_T / 64) / 4); i += MLP_T)
    reinterpret_cast<uint32_t*>(SS)[i] = 0u;
  __syncthreads();
  WaveScratch& S = SS[threadIdx.x / 64];
  const int lane = __lane_id();
  const int n = lane & 15, g = lane >> 4;
  const int64_t ntiles = (rows + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * (MLP_T / 64) + threadIdx.x / 64;
  const int64_t nwaves = (int64_t)gridDim.x * (MLP_T / 64);
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 gW1[4], gW2[4][4], gW3[4];
  float gb1[4][4], gb2[4][4], gb3[4], lsum = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = zero;
    gW3[a] = zero;
    gb3[a] = 0.f;
#pragma unroll
    for (int b = 0; b < 4; ++b) { gW2[a][b] = zero; gb1[a][b] = 0.f; gb2[a][b] = 0.f; }
  }
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    MlpTile t;
    mlp_forward_tile(W, X, rows, tile * 16, t, idx);
    const int64_t row = tile * 16 + n;
    // dz = scale * cw[y] * (softmax(z) - onehot(y)); lanes 0..15 own one row each
    float dz[4] = {0.f, 0.f, 0.f, 0.f};
    if (g == 0 && row < rows) {
      const int y = Y[idx ? (int64_t)idx[row] : row];
      const float w = (cw ? cw[y] : 1.f) * scale;
      float mx = fmaxf(fmaxf(t.z[0], t.z[1]), fmaxf(t.z[2], t.z[3]));
      float e[4], se = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) { e[j] = __expf(t.z[j] - mx); se += e[j]; }
      const float inv = 1.f / se;
#pragma unroll
      for (int j = 0; j < 4; ++j) dz[j] = w * (e[j] * inv - (j == y ? 1.f : 0.f));
      lsum += w * (__logf(se) + mx - t.z[y]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) gb3[j] += dz[j];
    const bf16x4 dzb = pack4(dz[0], dz[1], dz[2], dz[3]);  // B operand of W3^T dz^T (k = class)
    // dh2^T = W3^T dz^T, masked by the ReLU of layer 2
    bf16x4 d2b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      bf16x4 a = {0, 0, 0, 0};
      if (g == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = (short)W.w3[j * MLP_H + mb * 16 + n];
      }
      f32x4 d = mfma16(a, dzb, zero);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        d[j] = (t.a2[mb][j] > 0.f) ? d[j] : 0.f;
        gb2[mb][j] += d[j];
      }
      d2b[mb] = pack4(d[0], d[1], d[2], d[3]);
    }
    // dh1^T = W2^T dh2^T, masked by the ReLU of layer 1
    bf16x4 d1b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 d = zero;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = (short)W.w2[(kb * 16 + 4 * g + j) * MLP_H + mb * 16 + n];
        d = mfma16(a, d2b[kb], d);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        d[j] = (t.a1[mb][j] > 0.f) ? d[j] : 0.f;
        gb1[mb][j] += d[j];
      }
      d1b[mb] = pack4(d[0], d[1], d[2], d[3]);
    }
    // transposes for the weight gradients (K = the 16 rows of the tile)
    stash(S.x, 0, t.xb);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      stash(S.h1, mb, t.h1[mb]);
      stash(S.h2, mb, t.h2[mb]);
      stash(S.d2, mb, d2b[mb]);
      stash(S.d1, mb, d1b[mb]);
    }
    if (g == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) S.dz[j * 16 + n] = (uint16_t)dzb[j];
    }
    wave_sync_lds();
    // gW2[o][i] += sum_r dh2[r][o] h1[r][i]
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 a = lds4(&S.d2[(mb * 16 + n) * 16 + 4 * g]);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) gW2[mb][nb] = mfma16(a, lds4(&S.h1[(nb * 16 + n) * 16 + 4 * g]), gW2[mb][nb]);
    }
    // gW1[o][i] += sum_r dh1[r][o] x[r][i]
    {
      const bf16x4 b = lds4(&S.x[n * 16 + 4 * g]);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) gW1[mb] = mfma16(lds4(&S.d1[(mb * 16 + n) * 16 + 4 * g]), b, gW1[mb]);
    }
    // gW3[c][i] += sum_r dz[r][c] h2[r][i]
    {
      const bf16x4 a = lds4(&S.dz[n * 16 + 4 * g]);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) gW3[nb] = mfma16(a, lds4(&S.h2[(nb * 16 + n) * 16 + 4 * g]), gW3[nb]);
    }
    wave_sync_lds();
  }
  // weight gradients: lane holds (o = 16*mb + 4*g + j, i = 16*nb + n)
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = mb * 16 + 4 * g + j;
      atomicAdd(&grad[OFF_W1 + o * MLP_IN + n], gW1[mb][j]);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) atomicAdd(&grad[OFF_W2 + o * MLP_H + nb * 16 + n], gW2[mb][nb][j]);
    }
  if (g == 0) {
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int j = 0; j < 4; ++j) atomicAdd(&grad[OFF_W3 + j * MLP_H + nb * 16 + n], gW3[nb][j]);
  }
  // bias gradients and loss: sum over the 16 row lanes of each lane group
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        gb1[mb][j] += __shfl_xor(gb1[mb][j], off);
        gb2[mb][j] += __shfl_xor(gb2[mb][j], off);
      }
#pragma unroll
    for (int j = 0; j < 4; ++j) gb3[j] += __shfl_xor(gb3[j], off);
    lsum += __shfl_xor(lsum, off);
  }
  if (n == 0) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        atomicAdd(&grad[OFF_B1 + mb * 16 + 4 * g + j], gb1[mb][j]);
        atomicAdd(&grad[OFF_B2 + mb * 16 + 4 * g + j], gb2[mb][j]);
      }
    if (g == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) atomicAdd(&grad[OFF_B3 + j], gb3[j]);
      atomicAdd(loss, lsum);
    }
  }
}

__global__ void k_fill_f32(float* p, float v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Adam with bias correction (Kingma & Ba 2015) and decoupled weight decay on the weights
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                           float wd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const bool is_w = (i < OFF_B1) || (i >= OFF_W2 && i < OFF_B2) || (i >= OFF_W3 && i < OFF_B3);
  float pi = p[i];
  if (is_w) pi -= lr * wd * pi;
  p[i] = pi - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static int mlp_grid(int64_t rows) {
  const int64_t tiles = (rows + 15) / 16;
  return (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 1024));
}

static void launch_grad(const float* X, const int* y, const int* idx, int64_t rows, const float* params,
                        const float* cw, float scale, float* grad, float* loss, bool zero_loss, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f32, dim3(8), dim3(256), 0, s, grad, 0.f, (int64_t)MLP_NPARAM);
  if (zero_loss) hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, loss, 0.f, (int64_t)1);
  if (rows > 0)
    hipLaunchKernelGGL(k_mlp_grad, dim3(std::min(mlp_grid(rows), 128)), dim3(MLP_T), 0, s, X, y, idx, rows, params,
                       cw, scale, grad, loss);
}

static void launch_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr, float beta1,
                        float beta2, float eps, float wd, hipStream_t s) {
  const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, params, grad, m, v, lr, beta1,
                     beta2, eps, c1, c2, wd);
}

}  // namespace agimp

using namespace agimp;

// Device-resident training state: dataset, parameters, Adam moments, scratch.
struct agimp_trainer {
  int device = 0;
  hipStream_t s = nullptr;
  int64_t cap = 0, rows = 0;
  float* X = nullptr;     // [cap][16]
  int* y = nullptr;       // [cap]
  int* perm = nullptr;    // [cap]
  float* Z = nullptr;     // [cap][4]
  float* p = nullptr;     // [5508] parameters, then m, v, grad (one block)
  float* m = nullptr;
  float* v = nullptr;
  float* g = nullptr;
  float* loss = nullptr;  // [1]
  float* cw = nullptr;    // [4]
  int32_t step = 0;
};

extern "C" {

int agimp_mlp_forward(const float* X, int64_t rows, const float* params, float* logits, void* stream) {
  if (!X || !params || !logits || rows < 0) return AGIMP_EINVAL;
  if (rows == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(rows)), dim3(MLP_T), 0, s, X, rows, params, logits);
  return hipGetLastError() == hipSuccess ? 0 : AGIMP_EHIP;
}

int agimp_mlp_loss_grad(const float* X, const int32_t* y, int64_t rows, const float* params,
                        const float* class_weight, float scale, float* grad, float* loss, void* stream) {
  if (!X || !y || !params || !grad || !loss || rows < 0) return AGIMP_EINVAL;
  launch_grad(X, y, nullptr, rows, params, class_weight, scale, grad, loss, true, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : AGIMP_EHIP;
}

int agimp_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr, float beta1,
                   float beta2, float eps, float weight_decay, void* stream) {
  if (!params || !grad || !m || !v || step < 1) return AGIMP_EINVAL;
  launch_adam(params, grad, m, v, step, lr, beta1, beta2, eps, weight_decay, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : AGIMP_EHIP;
}

int agimp_trainer_destroy(agimp_trainer* t) {
  if (!t) return AGIMP_EINVAL;
  if (t->s) (void)hipStreamSynchronize(t->s);
  for (void* q : {(void*)t->X, (void*)t->y, (void*)t->perm, (void*)t->Z, (void*)t->p})
    if (q) (void)hipFree(q);
  if (t->s) (void)hipStreamDestroy(t->s);
  delete t;
  return 0;
}

int agimp_trainer_create(int64_t capacity, int32_t device, agimp_trainer** out) {
  if (!out || capacity < 1) return AGIMP_EINVAL;
  *out = nullptr;
  agimp_trainer* t = new agimp_trainer();
  t->device = device;
  t->cap = capacity;
  bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&t->s, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&t->X, sizeof(float) * MLP_IN * capacity) == hipSuccess &&
            hipMalloc(&t->y, sizeof(int) * capacity) == hipSuccess &&
            hipMalloc(&t->perm, sizeof(int) * capacity) == hipSuccess &&
            hipMalloc(&t->Z, sizeof(float) * MLP_OUT * capacity) == hipSuccess &&
            hipMalloc(&t->p, sizeof(float) * (4 * MLP_NPARAM + 8)) == hipSuccess;
  if (!ok) {
    agimp_trainer_destroy(t);
    return AGIMP_ENOMEM;
  }
  t->m = t->p + MLP_NPARAM;
  t->v = t->m + MLP_NPARAM;
  t->g = t->v + MLP_NPARAM;
  t->loss = t->g + MLP_NPARAM;
  t->cw = t->loss + 4;
  const float ones[4] = {1.f, 1.f, 1.f, 1.f};
  if (hipMemsetAsync(t->p, 0, sizeof(float) * (4 * MLP_NPARAM + 8), t->s) != hipSuccess ||
      hipMemcpyAsync(t->cw, ones, sizeof(ones), hipMemcpyHostToDevice, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess) {
    agimp_trainer_destroy(t);
    return AGIMP_EHIP;
  }
  *out = t;
  return 0;
}

int agimp_trainer_set_data(agimp_trainer* t, const float* X, const int32_t* y, int64_t rows) {
  if (!t || !X || !y || rows < 0 || rows > t->cap) return AGIMP_EINVAL;
  for (int64_t i = 0; i < rows; ++i)
    if (y[i] < 0 || y[i] >= MLP_OUT) return AGIMP_EINVAL;
  if (hipMemcpyAsync(t->X, X, sizeof(float) * MLP_IN * rows, hipMemcpyHostToDevice, t->s) != hipSuccess ||
      hipMemcpyAsync(t->y, y, sizeof(int) * rows, hipMemcpyHostToDevice, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  t->rows = rows;
  return 0;
}

int agimp_trainer_set_params(agimp_trainer* t, const float* params) {
  if (!t || !params) return AGIMP_EINVAL;
  if (hipMemcpyAsync(t->p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, t->s) != hipSuccess ||
      hipMemsetAsync(t->m, 0, sizeof(float) * 2 * MLP_NPARAM, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  t->step = 0;
  return 0;
}

int agimp_trainer_get_params(agimp_trainer* t, float* params) {
  if (!t || !params) return AGIMP_EINVAL;
  if (hipMemcpyAsync(params, t->p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  return 0;
}

int agimp_trainer_epoch(agimp_trainer* t, const int32_t* perm, int64_t batch, const float* class_weight, float scale,
                        float lr, float weight_decay, double* mean_loss) {
  if (!t || !perm || batch < 1 || !class_weight) return AGIMP_EINVAL;
  const int64_t rows = t->rows;
  for (int64_t i = 0; i < rows; ++i)
    if (perm[i] < 0 || perm[i] >= rows) return AGIMP_EINVAL;
  if (hipMemcpyAsync(t->perm, perm, sizeof(int) * rows, hipMemcpyHostToDevice, t->s) != hipSuccess ||
      hipMemcpyAsync(t->cw, class_weight, sizeof(float) * 4, hipMemcpyHostToDevice, t->s) != hipSuccess)
    return AGIMP_EHIP;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, t->s, t->loss, 0.f, (int64_t)1);
  int64_t nb = 0;
  for (int64_t b0 = 0; b0 + batch <= rows || (b0 == 0 && rows > 0); b0 += batch) {
    const int64_t bs = std::min(batch, rows - b0);
    launch_grad(t->X, t->y, t->perm + b0, bs, t->p, t->cw, scale, t->g, t->loss, false, t->s);
    launch_adam(t->p, t->g, t->m, t->v, ++t->step, lr, 0.9f, 0.999f, 1e-8f, weight_decay, t->s);
    ++nb;
  }
  float l = 0.f;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&l, t->loss, sizeof(float), hipMemcpyDeviceToHost, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  if (mean_loss) *mean_loss = nb ? (double)l / (double)nb : 0.0;
  return 0;
}

int agimp_trainer_logits(agimp_trainer* t, const float* X, int64_t rows, float* Z) {
  if (!t || !X || !Z || rows < 0 || rows > t->cap) return AGIMP_EINVAL;
  if (rows == 0) return 0;
  if (hipMemcpyAsync(t->X, X, sizeof(float) * MLP_IN * rows, hipMemcpyHostToDevice, t->s) != hipSuccess) return AGIMP_EHIP;
  t->rows = 0;  // the dataset was overwritten
  hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(rows)), dim3(MLP_T), 0, t->s, t->X, rows, t->p, t->Z);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(Z, t->Z, sizeof(float) * MLP_OUT * rows, hipMemcpyDeviceToHost, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  return 0;
}

int agimp_trainer_grad(agimp_trainer* t, const float* X, const int32_t* y, int64_t rows, const float* class_weight,
                       float scale, float* grad, float* loss) {
  if (!t || !X || !y || !grad || !loss || rows < 0 || rows > t->cap) return AGIMP_EINVAL;
  int rc = agimp_trainer_set_data(t, X, y, rows);
  if (rc) return rc;
  if (class_weight && hipMemcpyAsync(t->cw, class_weight, sizeof(float) * 4, hipMemcpyHostToDevice, t->s) != hipSuccess)
    return AGIMP_EHIP;
  launch_grad(t->X, t->y, nullptr, rows, t->p, class_weight ? t->cw : nullptr, scale, t->g, t->loss, true, t->s);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(grad, t->g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, t->s) != hipSuccess ||
      hipMemcpyAsync(loss, t->loss, sizeof(float), hipMemcpyDeviceToHost, t->s) != hipSuccess ||
      hipStreamSynchronize(t->s) != hipSuccess)
    return AGIMP_EHIP;
  return 0;
}

int agimp_trainer_apply(agimp_trainer* t, const float* grad, float lr, float weight_decay) {
  if (!t || !grad) return AGIMP_EINVAL;
  if (hipMemcpyAsync(t->g, grad, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, t->s) != hipSuccess)
    return AGIMP_EHIP;
  launch_adam(t->p, t->g, t->m, t->v, ++t->step, lr, 0.9f, 0.999f, 1e-8f, weight_decay, t->s);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(t->s) != hipSuccess) return AGIMP_EHIP;
  return 0;
}

}  // extern "C"
