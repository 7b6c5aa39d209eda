// The refinement-prediction network: a 16 -> 64 -> 64 -> 4 perceptron with ReLU, evaluated
// and trained with BF16 matrix-core instructions (v_mfma_f32_16x16x16_bf16) and FP32
// accumulation.  No counterpart in the reference (it refines from the current particles only).
//
// Parameters are one FP32 array (row-major, out x in):
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]   (5508 floats)
// Rounding points (the CPU oracle follows the same ones): inputs, weights and the two
// hidden activations are rounded to BF16 before each product; biases, accumulations,
// logits and the softmax are FP32.  Backward: dz, dh2, dh1 are rounded to BF16 where they
// enter a product; bias gradients sum the unrounded FP32 values.
//
// Wavefront tiling.  One wavefront handles 16 rows.  Every layer computes the TRANSPOSED
// product  H^T[out][row] = W[out][in] * X^T[in][row]  so that the accumulator layout of one
// layer (lane l: out = 16*mb + 4*(l/16) + j, row = l%16) is exactly the B-operand layout of
// the next (lane l: k = 16*kb + 4*(l/16) + j, column = l%16): activations never leave
// registers in the forward pass.  Weight gradients sum over rows (K = rows), which needs
// rows on the K side; those operands go through a per-wavefront LDS transpose.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace agimp {

constexpr int MLP_IN = 16, MLP_H = 64, MLP_OUT = 4;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_OUT * MLP_H,
              MLP_NPARAM = OFF_B3 + MLP_OUT;
static_assert(MLP_NPARAM == 5508, "parameter count");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 round to nearest even (NaN stays NaN)
__host__ __device__ inline uint16_t f2bf(float f) {
  union { float f; uint32_t u; } c;
  c.f = f;
  uint32_t u = c.u;
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  union { float f; uint32_t u; } c;
  c.u = (uint32_t)b << 16;
  return c.f;
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Network in LDS: BF16 weights, FP32 biases (11.5 KB)
struct MlpLds {
  uint16_t w1[MLP_H * MLP_IN];
  uint16_t w2[MLP_H * MLP_H];
  uint16_t w3[MLP_OUT * MLP_H];
  float b1[MLP_H], b2[MLP_H], b3[MLP_OUT];
};

// cooperative load of the parameters by the whole workgroup (followed by __syncthreads)
__device__ __forceinline__ void mlp_load(MlpLds& W, const float* __restrict__ p) {
  for (int i = threadIdx.x; i < MLP_H * MLP_IN; i += blockDim.x) W.w1[i] = f2bf(p[OFF_W1 + i]);
  for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) W.w2[i] = f2bf(p[OFF_W2 + i]);
  for (int i = threadIdx.x; i < MLP_OUT * MLP_H; i += blockDim.x) W.w3[i] = f2bf(p[OFF_W3 + i]);
  for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) { W.b1[i] = p[OFF_B1 + i]; W.b2[i] = p[OFF_B2 + i]; }
  for (int i = threadIdx.x; i < MLP_OUT; i += blockDim.x) W.b3[i] = p[OFF_B3 + i];
}

__device__ __forceinline__ bf16x4 lds4(const uint16_t* p) {
  // 4 consecutive BF16 values (8-byte aligned by construction)
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

// Activations of one 16-row tile as held by one wavefront (see the layout note above).
struct MlpTile {
  f32x4 a1[4], a2[4];  // pre-activations of the hidden layers
  bf16x4 xb;           // input, B operand of layer 1 (k = feature)
  bf16x4 h1[4], h2[4]; // BF16 hidden activations, B operands of the next layer
  f32x4 z;             // logits: lanes 0..15 hold z[row = lane][j]
};

// Forward pass of rows r0 .. r0+15 (rows past `rows` read as 0); row r is X[idx ? idx[r] : r].
__device__ __forceinline__ void mlp_forward_tile(const MlpLds& W, const float* __restrict__ X, int64_t rows,
                                                 int64_t r0, MlpTile& t, const int* __restrict__ idx = nullptr) {
  const int lane = __lane_id();
  const int n = lane & 15, g = lane >> 4;
  const int64_t row = r0 + n;
  float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (row < rows) {
    const int64_t src = idx ? (int64_t)idx[row] : row;
    xv = *reinterpret_cast<const float4*>(X + src * MLP_IN + 4 * g);
  }
  t.xb = pack4(xv.x, xv.y, xv.z, xv.w);
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 acc = mfma16(lds4(&W.w1[(mb * 16 + n) * MLP_IN + 4 * g]), t.xb, zero);
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[j] += W.b1[mb * 16 + 4 * g + j];
      h[j] = fmaxf(acc[j], 0.f);
    }
    t.a1[mb] = acc;
    t.h1[mb] = pack4(h[0], h[1], h[2], h[3]);
  }
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 acc = zero;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma16(lds4(&W.w2[(mb * 16 + n) * MLP_H + kb * 16 + 4 * g]), t.h1[kb], acc);
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[j] += W.b2[mb * 16 + 4 * g + j];
      h[j] = fmaxf(acc[j], 0.f);
    }
    t.a2[mb] = acc;
    t.h2[mb] = pack4(h[0], h[1], h[2], h[3]);
  }
  f32x4 acc = zero;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    bf16x4 a = {0, 0, 0, 0};
    if (n < MLP_OUT) a = lds4(&W.w3[n * MLP_H + kb * 16 + 4 * g]);
    acc = mfma16(a, t.h2[kb], acc);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] += (g == 0) ? W.b3[j] : 0.f;
  t.z = acc;
}

}  // namespace agimp
