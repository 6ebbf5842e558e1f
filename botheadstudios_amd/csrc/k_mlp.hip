// Learned smoothing-length seed: BF16-weight / FP32-accumulate inference and training kernels (gfx950).
//
// Inference runs on the matrix cores, 32 particles per wave, with the weights expanded from BF16 into LDS. Training
// (one particle per thread) computes the exact gradient of the mean squared error (forward, backward, then
// per-parameter wave reductions in a fixed order, so the gradient is deterministic) and applies Adam to FP32 master
// weights; the forward passes always use the BF16-rounded weights, as inference does.
#include "common.h"
#include "mlp.h"

namespace {

constexpr float kGeluK = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluC = 0.044715f;

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// GELU (tanh form) and its derivative. With u = K (z + C z^3), 0.5 z (1 + tanh u) = z s, s = 1 / (1 + e^{-2u}) =
// 1 / (1 + 2^{z (g1 + g3 z^2)}), g1 = -2 K log2 e, g3 = g1 C: a multiply, an FMA and a multiply, then the hardware
// exp2 and reciprocal (1 ulp each; an IEEE division here was a third of the inference kernel's instructions). Far
// below zero the power overflows to +inf and s = 0, far above it s = 1: the function saturates without a NaN.
constexpr float kGelu1 = -2.0f * kGeluK * 1.4426950408889634f;
constexpr float kGelu3 = kGelu1 * kGeluC;

__device__ __forceinline__ float gelu(float z, float& dg) {
  const float z2 = z * z;
  const float sg = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(kGelu3, z2, kGelu1)));
  dg = sg + 2.0f * z * sg * (1.0f - sg) * kGeluK * (1.0f + 3.0f * kGeluC * z2);
  return z * sg;
}

__device__ __forceinline__ void load_weights(float* sw, const uint16_t* __restrict__ wb) {
  for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x) sw[p] = bf16_to_f32(wb[p]);
  __syncthreads();
}

__device__ __forceinline__ float forward(const float* sw, const float x[MLP_IN]) {
  float a1[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    float z = sw[MLP_B1 + i];
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) z = fmaf(sw[MLP_W1 + i * MLP_IN + j], x[j], z);
    float dg;
    a1[i] = gelu(z, dg);
  }
  float y = sw[MLP_B3];
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    float z = sw[MLP_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z = fmaf(sw[MLP_W2 + i * MLP_H + j], a1[j], z);
    float dg;
    y = fmaf(sw[MLP_W3 + i], gelu(z, dg), y);
  }
  return y;
}

// The forward pass on the matrix cores. v_mfma_f32_32x32x2f32 takes A[i][k] from lane i + 32 k and B[k][j] from lane
// j + 32 k, and leaves D[i][j] in lane j + 32 (i / 4 mod 2), register 4 (i / 8) + i mod 4. Both layers are computed
// transposed, hidden units as rows and the 32 particles of a pass as columns, so A holds weights (the same for every
// pass: each lane keeps its 21 in registers) and B the particles' values. Layer 1 is W1, with b1 as the weight of a
// constant input 1, times the inputs: K = 8 in 4 steps. Each lane then holds 16 hidden activations of its column's
// particle, rows 8 (r / 4) + 4 half + r mod 4, which are exactly layer 2's B operands when layer 2's K index runs over
// those rows in that order (16 steps, no lane exchange); b2 enters as the weight of a 17th step's constant input 1.
// The output unit sums its lane's 16 rows and the other half's.
// FP32 inputs and accumulation, the weights as the BF16 copy holds them; only the summation order differs from
// forward().
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int W2T_STRIDE = 40;  // W2 transposed in LDS, rows of 40: a wave's two halves read 32 banks apart

struct MfmaWeights {
  float a1[4];   // layer 1's A operand for steps 0..3
  float a2[17];  // layer 2's A operand for steps 0..16
};

// sw: the parameters in order; w2t[kin * W2T_STRIDE + row] = W2[row][kin].
__device__ __forceinline__ void load_weights_mfma(float* sw, float* w2t, const uint16_t* __restrict__ wb) {
  for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x) {
    const float v = bf16_to_f32(wb[p]);
    sw[p] = v;
    const int q = p - MLP_W2;
    if (q >= 0 && q < MLP_H * MLP_H) w2t[(q % MLP_H) * W2T_STRIDE + q / MLP_H] = v;
  }
  __syncthreads();
}

__device__ __forceinline__ MfmaWeights lane_weights(const float* sw, const float* w2t, int lane) {
  const int half = lane >> 5, col = lane & 31;
  MfmaWeights w;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 2 * s + half;  // the input this lane supplies
    w.a1[s] = k < MLP_IN ? sw[MLP_W1 + col * MLP_IN + k] : (k == MLP_IN ? sw[MLP_B1 + col] : 0.f);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) w.a2[r] = w2t[((r >> 2) * 8 + half * 4 + (r & 3)) * W2T_STRIDE + col];
  w.a2[16] = half ? 0.f : sw[MLP_B2 + col];
  return w;
}

// x: the inputs of particle (lane mod 32) of the pass, in both halves; returns that particle's output in both halves.
__device__ __forceinline__ float forward_mfma(const float* sw, const MfmaWeights& w, const float x[MLP_IN], int half) {
  const float xx[8] = {x[0], x[1], x[2], x[3], x[4], x[5], 1.0f, 0.0f};
  f32x16 z1 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s)
    z1 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.a1[s], half ? xx[2 * s + 1] : xx[2 * s], z1, 0, 0, 0);
  f32x16 z2 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float dg;
    z2 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.a2[r], gelu(z1[r], dg), z2, 0, 0, 0);
  }
  z2 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.a2[16], half ? 0.f : 1.f, z2, 0, 0, 0);
  float part = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float dg;
    part = fmaf(sw[MLP_W3 + (r >> 2) * 8 + half * 4 + (r & 3)], gelu(z2[r], dg), part);
  }
  return sw[MLP_B3] + (part + __shfl_xor(part, 32, 64));
}

// A wave's 64 particles in two passes of 32: lane l holds particle l's inputs and receives particle l's output.
__device__ __forceinline__ float forward_wave(const float* sw, const MfmaWeights& w, const float x[MLP_IN], int lane) {
  const int half = lane >> 5, col = lane & 31;
  float xa[MLP_IN], xb[MLP_IN];
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) {
    xa[j] = __shfl(x[j], col, 64);
    xb[j] = __shfl(x[j], col + 32, 64);
  }
  const float ya = forward_mfma(sw, w, xa, half);
  const float yb = forward_mfma(sw, w, xb, half);
  return half ? yb : ya;
}

__device__ __forceinline__ void features(int64_t t, const MasterBufs& D, const EosDev* __restrict__ eos, double dt,
                                         float eta, float x[MLP_IN], float& h) {
  h = D.h[t];
  const float rho = fmaxf(D.rho[t], 1.0e-30f);
  const float4 f = D.feat[t];
  const uint32_t mat = D.mat[t];
  const float rho0 = eos[mat].rho0;
  x[0] = f.y * (float)dt * MLP_YSCALE;
  x[1] = __logf(rho / rho0);
  x[2] = (f.x - 47.0f) * (1.0f / 47.0f);
  x[3] = f.z * h / rho;
  x[4] = __logf(h / (eta * cbrtf(D.m[t] / rho0)));
  x[5] = (float)mat;
}

// Waves stride over tiles of 64 particles, one particle per lane (forward_wave); the grid is capped so that a wave
// loads its weights once for several tiles. The loop bound is uniform over the wave: the matrix instructions need
// every lane, and lanes past the end repeat the tile's first particle without writing.
constexpr int MFMA_BLOCKS = 1024;

__global__ __launch_bounds__(256) void k_mlp_seed(int64_t n, MasterBufs D, float* __restrict__ hs,
                                                  const EosDev* __restrict__ eos, SegView sv, float eta,
                                                  const uint16_t* __restrict__ wb, float margin) {
  __shared__ float sw[MLP_NPARAM];
  __shared__ float w2t[MLP_H * W2T_STRIDE];
  load_weights_mfma(sw, w2t, wb);
  const int lane = threadIdx.x & 63;
  const MfmaWeights w = lane_weights(sw, w2t, lane);
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * 64;
  for (int64_t tile = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; tile < n; tile += stride) {
    const int64_t t = tile + lane < n ? tile + lane : tile;
    const uint32_t sg = sv.of(t);
    const float hcap = sv.hcap[sg];
    float x[MLP_IN], h;
    features(t, D, eos, sv.clk[sg].dt, eta, x, h);
    const float y = forward_wave(sw, w, x, lane) + MLP_PHYS * x[0];
    if (tile + lane >= n) continue;
    float seed = h * __expf(fminf(fmaxf(y, -50.f), 50.f) * (1.0f / MLP_YSCALE));
    seed = fminf(fmaxf(seed, h * (1.0f / MLP_CLAMP)), h * MLP_CLAMP);
    seed = fminf(seed, fmaxf(hcap, h));
    D.h[t] = seed;
    hs[t] = fminf(seed * margin, fmaxf(hcap, seed));
  }
}

__global__ __launch_bounds__(256) void k_mlp_features(int64_t n, MasterBufs D, const EosDev* __restrict__ eos,
                                                      SegView sv, float eta, float* __restrict__ xbuf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float x[MLP_IN], h;
  features(t, D, eos, sv.clk[sv.of(t)].dt, eta, x, h);
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) xbuf[t * MLP_SAMPLE + j] = x[j];
  xbuf[t * MLP_SAMPLE + 6] = h;  // the previous h, to form the target after the solve
}

__global__ __launch_bounds__(256) void k_mlp_targets(int64_t n, const float* __restrict__ h_new,
                                                     const float* __restrict__ xbuf, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) out[t * MLP_SAMPLE + j] = xbuf[t * MLP_SAMPLE + j];
  out[t * MLP_SAMPLE + 6] = logf(h_new[t] / xbuf[t * MLP_SAMPLE + 6]) * MLP_YSCALE - MLP_PHYS * xbuf[t * MLP_SAMPLE];
  out[t * MLP_SAMPLE + 7] = 0.f;
}

// One wave per block: every lane takes samples, per-parameter gradients are wave-reduced in a fixed order
// and accumulated in LDS; the block writes its partial (parameters + loss).
constexpr int GRAD_BLOCKS = 512;

// Sample k of the batch is row (off + k * step) mod rows of the sample table (step coprime with rows spreads a
// batch over the whole table without a shuffle pass).
__global__ __launch_bounds__(64) void k_mlp_grad(const uint16_t* __restrict__ wb, const float* __restrict__ xs,
                                                 int64_t xstride, const float* __restrict__ ys, int64_t ystride,
                                                 int64_t n, int64_t rows, int64_t off, int64_t step,
                                                 float* __restrict__ gpart) {
  __shared__ float sw[MLP_NPARAM];
  __shared__ float sg[MLP_NPARAM + 1];
  for (int p = threadIdx.x; p < MLP_NPARAM + 1; p += 64) sg[p] = 0.f;
  load_weights(sw, wb);
  const int lane = threadIdx.x;
  const int64_t nround = (n + 64 * (int64_t)gridDim.x - 1) / (64 * (int64_t)gridDim.x);
  for (int64_t r = 0; r < nround; ++r) {
    const int64_t k = (r * gridDim.x + blockIdx.x) * 64 + lane;
    const bool ok = k < n;
    const int64_t s = ok ? (int64_t)(((uint64_t)off + (uint64_t)k * (uint64_t)step) % (uint64_t)rows) : 0;
    float x[MLP_IN];
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) x[j] = ok ? xs[s * xstride + j] : 0.f;
    const float tgt = ok ? ys[s * ystride] : 0.f;
    float a1[MLP_H], g1[MLP_H], a2[MLP_H], g2[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
      float z = sw[MLP_B1 + i];
#pragma unroll
      for (int j = 0; j < MLP_IN; ++j) z = fmaf(sw[MLP_W1 + i * MLP_IN + j], x[j], z);
      a1[i] = gelu(z, g1[i]);
    }
    float y = sw[MLP_B3];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
      float z = sw[MLP_B2 + i];
#pragma unroll
      for (int j = 0; j < MLP_H; ++j) z = fmaf(sw[MLP_W2 + i * MLP_H + j], a1[j], z);
      a2[i] = gelu(z, g2[i]);
      y = fmaf(sw[MLP_W3 + i], a2[i], y);
    }
    const float dy = ok ? (y - tgt) : 0.f;
    // dz2 overwrites g2; dz1 overwrites g1.
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) g2[i] *= dy * sw[MLP_W3 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int i = 0; i < MLP_H; ++i) acc = fmaf(sw[MLP_W2 + i * MLP_H + j], g2[i], acc);
      g1[j] *= acc;
    }
    auto add = [&](int p, float v) {
      v = wave_sum(v);
      if (lane == 0) sg[p] += v;
    };
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
#pragma unroll
      for (int j = 0; j < MLP_IN; ++j) add(MLP_W1 + i * MLP_IN + j, g1[i] * x[j]);
      add(MLP_B1 + i, g1[i]);
#pragma unroll
      for (int j = 0; j < MLP_H; ++j) add(MLP_W2 + i * MLP_H + j, g2[i] * a1[j]);
      add(MLP_B2 + i, g2[i]);
      add(MLP_W3 + i, dy * a2[i]);
    }
    add(MLP_B3, dy);
    add(MLP_NPARAM, 0.5f * dy * dy);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < MLP_NPARAM + 1; p += 64) gpart[(int64_t)blockIdx.x * (MLP_NPARAM + 1) + p] = sg[p];
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ gpart, int nblk, int64_t n,
                                                    float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > MLP_NPARAM) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += gpart[(int64_t)b * (MLP_NPARAM + 1) + p];
  grad[p] = (float)(s / (double)(n > 0 ? n : 1));
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ w, uint16_t* __restrict__ wb,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ grad, float lr, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = grad[p];
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  const float nw = w[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
  w[p] = nw;
  wb[p] = f32_to_bf16(nw);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NPARAM) wb[p] = f32_to_bf16(w[p]);
}

__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* __restrict__ wb, const float* __restrict__ x,
                                                     int64_t n, float* __restrict__ out) {
  __shared__ float sw[MLP_NPARAM];
  __shared__ float w2t[MLP_H * W2T_STRIDE];
  load_weights_mfma(sw, w2t, wb);
  const int lane = threadIdx.x & 63;
  const MfmaWeights w = lane_weights(sw, w2t, lane);
  const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6) * 64;
  for (int64_t tile = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; tile < n; tile += stride) {
    const int64_t t = tile + lane < n ? tile + lane : tile;
    float xx[MLP_IN];
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) xx[j] = x[t * MLP_IN + j];
    const float y = forward_wave(sw, w, xx, lane);
    if (tile + lane < n) out[t] = y;
  }
}

inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }
// Blocks of 4 waves, 64 particles per wave and tile, at most MFMA_BLOCKS blocks.
inline unsigned gmfma(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, MFMA_BLOCKS); }

}  // namespace

int mlp_alloc(MlpDev& m, int64_t cap) {
  m.cap = cap;
  m.gblocks = GRAD_BLOCKS;
  if (hipMalloc(&m.w, MLP_NPARAM * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&m.wb, MLP_NPARAM * sizeof(uint16_t)) != hipSuccess) return -1;
  if (hipMalloc(&m.adam_m, MLP_NPARAM * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&m.adam_v, MLP_NPARAM * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&m.grad, (MLP_NPARAM + 1) * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&m.gpart, (size_t)GRAD_BLOCKS * (MLP_NPARAM + 1) * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&m.xbuf, (size_t)cap * MLP_SAMPLE * sizeof(float)) != hipSuccess) return -1;
  hipMemset(m.w, 0, MLP_NPARAM * sizeof(float));
  hipMemset(m.wb, 0, MLP_NPARAM * sizeof(uint16_t));
  hipMemset(m.adam_m, 0, MLP_NPARAM * sizeof(float));
  hipMemset(m.adam_v, 0, MLP_NPARAM * sizeof(float));
  return 0;
}

void mlp_free(MlpDev& m) {
  for (void* p : {(void*)m.w, (void*)m.wb, (void*)m.adam_m, (void*)m.adam_v, (void*)m.grad, (void*)m.gpart,
                  (void*)m.xbuf, (void*)m.samples})
    if (p) hipFree(p);
  m = MlpDev{};
}

void mlp_seed_h(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D, float* hs, const void* eos,
                const SegView& sv, float eta, float margin) {
  if (!m.trained) return;
  k_mlp_seed<<<gmfma(n), 256, 0, s>>>(n, D, hs, (const EosDev*)eos, sv, eta, m.wb, margin);
}

void mlp_record_features(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D, const void* eos, const SegView& sv,
                         float eta) {
  k_mlp_features<<<g256(n), 256, 0, s>>>(n, D, (const EosDev*)eos, sv, eta, m.xbuf);
}

void mlp_record_targets(hipStream_t s, MlpDev& m, int64_t n, const MasterBufs& D) {
  if (!m.samples || m.nsamples + n > m.sample_cap) return;
  k_mlp_targets<<<g256(n), 256, 0, s>>>(n, D.h, m.xbuf, m.samples + m.nsamples * MLP_SAMPLE);
  m.nsamples += n;
}

void mlp_gradient(hipStream_t s, MlpDev& m, const float* x, int64_t xstride, const float* y, int64_t ystride,
                  int64_t n, int64_t rows, int64_t off, int64_t step) {
  k_mlp_grad<<<GRAD_BLOCKS, 64, 0, s>>>(m.wb, x, xstride, y, ystride, n, rows, off, step, m.gpart);
  k_mlp_reduce<<<g256(MLP_NPARAM + 1), 256, 0, s>>>(m.gpart, GRAD_BLOCKS, n, m.grad);
}

void mlp_adam(hipStream_t s, MlpDev& m, float lr) {
  m.adam_t += 1;
  const float bc1 = 1.0f - powf(0.9f, (float)m.adam_t), bc2 = 1.0f - powf(0.999f, (float)m.adam_t);
  k_mlp_adam<<<g256(MLP_NPARAM), 256, 0, s>>>(m.w, m.wb, m.adam_m, m.adam_v, m.grad, lr, bc1, bc2);
}

void mlp_forward(hipStream_t s, MlpDev& m, const float* x, int64_t n, float* out) {
  k_mlp_forward<<<gmfma(n), 256, 0, s>>>(m.wb, x, n, out);
}

void mlp_sync_bf16(hipStream_t s, MlpDev& m) { k_mlp_to_bf16<<<g256(MLP_NPARAM), 256, 0, s>>>(m.w, m.wb); }
