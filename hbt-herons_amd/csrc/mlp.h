// Per-particle bound-set classifier for the learned warm start: an
// 8 -> 64 -> 64 -> 1 MLP with ReLU hidden layers and a logistic output.
// Parameters are kept as FP32 master weights for the optimiser and rounded to
// BF16 for every forward/backward pass (accumulation in FP32).  Weight reads are
// wave-uniform, so they go through the scalar cache.
//
// Flat parameter layout: W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
#pragma once
#include "common.h"

namespace hbtx
{
#define MLP_IN 8
#define MLP_H 64
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H)
#define MLP_NPAR (MLP_OFF_B3 + 1)
#define MLP_BLOCK 128

HBTX_DEV float wload(const float *w, int i)
{
  typedef const __attribute__((address_space(4))) float cf;
  return ((cf *)(uintptr_t)w)[i];
}

// round-to-nearest-even FP32 -> BF16 bits
HBTX_HD uint16_t f32_to_bf16(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u)
    return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
HBTX_HD float bf16_to_f32(uint16_t b)
{
  return __uint_as_float((uint32_t)b << 16);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

HBTX_DEV __bf16 bf16_bits(uint16_t b)
{
  return __builtin_bit_cast(__bf16, b);
}

// Inference on the matrix cores (v_mfma_f32_32x32x16_bf16): one wave evaluates 32
// samples per step, neurons as matrix rows and samples as columns.
//   layer 1: H1^T[64 x 32] = W1[64 x 8] X^T, with X split into bf16 hi + lo parts
//            placed in the two K halves (so the product keeps FP32 accuracy);
//   layer 2: H2^T[64 x 32] = W2[64 x 64] relu(H1^T + b1), H1 again split hi + lo;
//   layer 3: z = w3 . relu(H2^T + b2) + b3 on the vector units.
// Lane l holds output column l % 32 and rows 8*(r/4) + 4*(l/32) + r%4 of a 32 x 32
// accumulator block (r = 0..15); a lane's layer-1 rows are exactly the K entries
// it supplies to layer 2, with the K order of both operands permuted alike.
// groups (may be null): the 32-particle groups to evaluate (particles 32 g .. 32 g + 31
// for g = groups[0 .. ngroups)); null evaluates every particle.
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ w, const uint16_t *__restrict__ wb,
                                                     const float *__restrict__ X, int64_t n,
                                                     float *__restrict__ logit, const int32_t *__restrict__ groups,
                                                     int64_t ngroups)
{
  const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  bf16x8 a1[2];
  bf16x8 a2[2][4];
  float b1v[32], b2v[32], w3v[32];
#pragma unroll
  for (int mb = 0; mb < 2; mb++)
  {
    const int row = 32 * mb + c;
#pragma unroll
    for (int s = 0; s < 8; s++)
      a1[mb][s] = bf16_bits(wb[MLP_OFF_W1 + row * MLP_IN + s]);
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int s = 0; s < 8; s++)
        a2[mb][kb][s] = bf16_bits(wb[MLP_OFF_W2 + row * MLP_H + 8 * (2 * kb + s / 4) + 4 * h + (s & 3)]);
#pragma unroll
    for (int r = 0; r < 16; r++)
    {
      const int i = 32 * mb + 8 * (r / 4) + 4 * h + (r & 3);
      b1v[16 * mb + r] = w[MLP_OFF_B1 + i];
      b2v[16 * mb + r] = w[MLP_OFF_B2 + i];
      w3v[16 * mb + r] = w[MLP_OFF_W3 + i];
    }
  }
  const float b3 = w[MLP_OFF_B3];
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / HBTX_WAVE) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / HBTX_WAVE);
  const int64_t G = groups ? ngroups : (n + 31) / 32;
  for (int64_t gi = wave; gi < G; gi += nwaves)
  {
    const int64_t g = groups ? (int64_t)groups[gi] : gi;
    const int64_t j = g * 32 + c;
    float x[8];
    if (j < n)
    {
      const float4 u0 = reinterpret_cast<const float4 *>(X + j * MLP_IN)[0];
      const float4 u1 = reinterpret_cast<const float4 *>(X + j * MLP_IN)[1];
      x[0] = u0.x, x[1] = u0.y, x[2] = u0.z, x[3] = u0.w, x[4] = u1.x, x[5] = u1.y, x[6] = u1.z, x[7] = u1.w;
    }
    else
    {
#pragma unroll
      for (int s = 0; s < 8; s++)
        x[s] = 0.f;
    }
    bf16x8 bx;
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
      const __bf16 hi = (__bf16)x[s];
      bx[s] = h ? (__bf16)(x[s] - (float)hi) : hi;
    }
    f32x16 zero;
#pragma unroll
    for (int r = 0; r < 16; r++)
      zero[r] = 0.f;
    bf16x8 bh[4], bl[4];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
    {
      const f32x16 d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[mb], bx, zero, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; r++)
      {
        const float v = fmaxf(d1[r] + b1v[16 * mb + r], 0.f);
        const int q = 4 * mb + r / 4, kb = q >> 1, s = 4 * (q & 1) + (r & 3);
        const __bf16 hi = (__bf16)v;
        bh[kb][s] = hi;
        bl[kb][s] = (__bf16)(v - (float)hi);
      }
    }
    float z = 0.f;
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
    {
      f32x16 d2 = zero;
#pragma unroll
      for (int kb = 0; kb < 4; kb++)
      {
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[mb][kb], bh[kb], d2, 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[mb][kb], bl[kb], d2, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 16; r++)
        z += w3v[16 * mb + r] * fmaxf(d2[r] + b2v[16 * mb + r], 0.f);
    }
    z += __shfl_xor(z, 32, HBTX_WAVE);
    if (h == 0 && j < n)
      logit[j] = z + b3;
  }
}

inline int mlp_forward_blocks(int64_t n)
{
  const int64_t groups = (n + 31) / 32;
  return (int)std::max<int64_t>(1, std::min<int64_t>((groups + 3) / 4, 4096));
}

// Forward + backward of one block of MLP_BLOCK samples; writes this block's
// gradient sum (already divided by the batch size) and loss sum.
// sample k of the batch is row idx[k] of X (idx may be null: row k).
__global__ __launch_bounds__(MLP_BLOCK) void k_mlp_grad(const float *__restrict__ w, const float *__restrict__ X,
                                                        const uint8_t *__restrict__ y, const int32_t *__restrict__ idx,
                                                        int64_t nbatch, float inv_batch, float *__restrict__ partial,
                                                        double *__restrict__ loss_partial)
{
  __shared__ float sx[MLP_BLOCK][MLP_IN];
  __shared__ float sh1[MLP_BLOCK][MLP_H + 1];
  __shared__ float sh2[MLP_BLOCK][MLP_H + 1]; // h2, later dh1
  __shared__ float sdh2[MLP_BLOCK][MLP_H + 1];
  __shared__ float sdz[MLP_BLOCK];
  __shared__ double sloss[MLP_BLOCK];
  const int t = threadIdx.x;
  const int64_t k = (int64_t)blockIdx.x * MLP_BLOCK + t;
  const bool valid = k < nbatch;
  const int64_t row = valid ? (idx ? (int64_t)idx[k] : k) : 0;
  float x[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++)
  {
    x[i] = valid ? X[row * MLP_IN + i] : 0.f;
    sx[t][i] = x[i];
  }
  float h1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; j++)
  {
    float a = wload(w, MLP_OFF_B1 + j);
#pragma unroll
    for (int i = 0; i < MLP_IN; i++)
      a += wload(w, MLP_OFF_W1 + j * MLP_IN + i) * x[i];
    h1[j] = fmaxf(a, 0.f);
    sh1[t][j] = h1[j];
  }
  float z = wload(w, MLP_OFF_B3);
  for (int j = 0; j < MLP_H; j++)
  {
    float a = wload(w, MLP_OFF_B2 + j);
#pragma unroll
    for (int i = 0; i < MLP_H; i++)
      a += wload(w, MLP_OFF_W2 + j * MLP_H + i) * h1[i];
    const float h2 = fmaxf(a, 0.f);
    sh2[t][j] = h2;
    z += wload(w, MLP_OFF_W3 + j) * h2;
  }
  const float yy = valid ? (float)y[row] : 0.f;
  const float sig = 1.f / (1.f + __expf(-z));
  const float dz = valid ? (sig - yy) * inv_batch : 0.f;
  sdz[t] = dz;
  // binary cross-entropy with logits: softplus(z) - y z
  sloss[t] = valid ? (double)(fmaxf(z, 0.f) + log1pf(__expf(-fabsf(z))) - yy * z) : 0.0;
  float dh1[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; i++)
    dh1[i] = 0.f;
  for (int j = 0; j < MLP_H; j++)
  {
    const float d2 = (sh2[t][j] > 0.f) ? dz * wload(w, MLP_OFF_W3 + j) : 0.f;
    sdh2[t][j] = d2;
#pragma unroll
    for (int i = 0; i < MLP_H; i++)
      dh1[i] += wload(w, MLP_OFF_W2 + j * MLP_H + i) * d2;
  }
  __syncthreads();
  float *out = partial + (int64_t)blockIdx.x * MLP_NPAR;
  // output layer and second hidden layer: sums over the block's samples
  for (int q = t; q < MLP_H * MLP_H + 2 * MLP_H + 1; q += MLP_BLOCK)
  {
    float acc = 0.f;
    if (q < MLP_H * MLP_H)
    {
      const int j = q / MLP_H, i = q % MLP_H;
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sdh2[s][j] * sh1[s][i];
      out[MLP_OFF_W2 + q] = acc;
    }
    else if (q < MLP_H * MLP_H + MLP_H)
    {
      const int j = q - MLP_H * MLP_H;
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sdh2[s][j];
      out[MLP_OFF_B2 + j] = acc;
    }
    else if (q < MLP_H * MLP_H + 2 * MLP_H)
    {
      const int j = q - MLP_H * MLP_H - MLP_H;
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sdz[s] * sh2[s][j];
      out[MLP_OFF_W3 + j] = acc;
    }
    else
    {
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sdz[s];
      out[MLP_OFF_B3] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MLP_H; i++)
    sh2[t][i] = (h1[i] > 0.f) ? dh1[i] : 0.f;
  __syncthreads();
  for (int q = t; q < MLP_H * MLP_IN + MLP_H; q += MLP_BLOCK)
  {
    float acc = 0.f;
    if (q < MLP_H * MLP_IN)
    {
      const int j = q / MLP_IN, i = q % MLP_IN;
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sh2[s][j] * sx[s][i];
      out[MLP_OFF_W1 + q] = acc;
    }
    else
    {
      const int j = q - MLP_H * MLP_IN;
      for (int s = 0; s < MLP_BLOCK; s++)
        acc += sh2[s][j];
      out[MLP_OFF_B1 + j] = acc;
    }
  }
  if (t == 0)
  {
    double l = 0.0;
    for (int s = 0; s < MLP_BLOCK; s++)
      l += sloss[s];
    loss_partial[blockIdx.x] = l;
  }
}

// Sum of the block partials (fixed order, deterministic).
__global__ void k_mlp_reduce(const float *__restrict__ partial, int nblocks, float *__restrict__ grad,
                             const double *__restrict__ loss_partial, double *__restrict__ loss)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < MLP_NPAR)
  {
    float acc = 0.f;
    for (int b = 0; b < nblocks; b++)
      acc += partial[(int64_t)b * MLP_NPAR + q];
    grad[q] = acc;
  }
  if (q == 0)
  {
    double l = 0.0;
    for (int b = 0; b < nblocks; b++)
      l += loss_partial[b];
    *loss = l;
  }
}

// Adam step on the FP32 master weights, then the BF16 working copy.
__global__ void k_mlp_adam(float *__restrict__ master, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float *__restrict__ wf, uint16_t *__restrict__ wb, float lr,
                           float b1, float b2, float eps, float c1, float c2)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NPAR)
    return;
  const float g = grad[q];
  const float mq = b1 * m[q] + (1.f - b1) * g;
  const float vq = b2 * v[q] + (1.f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  const float w = master[q] - lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  master[q] = w;
  const uint16_t b = f32_to_bf16(w);
  wb[q] = b;
  wf[q] = bf16_to_f32(b);
}

__global__ void k_mlp_round(const float *__restrict__ master, float *__restrict__ wf, uint16_t *__restrict__ wb)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NPAR)
    return;
  const uint16_t b = f32_to_bf16(master[q]);
  wb[q] = b;
  wf[q] = bf16_to_f32(b);
}

} // namespace hbtx
