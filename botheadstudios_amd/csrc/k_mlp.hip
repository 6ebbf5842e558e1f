// Learned smoothing-length seed: BF16-weight / FP32-accumulate inference and training kernels (gfx950).
//
// Inference runs one particle per thread with the weights expanded from BF16 into LDS. Training computes the
// exact gradient of the mean squared error (forward, backward, then per-parameter wave reductions in a fixed
// order, so the gradient is deterministic) and applies Adam to FP32 master weights; the forward passes always
// use the BF16-rounded weights, exactly as inference does.
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

// GELU (tanh form) and its derivative. tanh u = 1 - 2 / (e^{2u} + 1), with the hardware reciprocal (1 ulp; the
// IEEE division was a third of the inference kernel's instructions); it saturates to +-1 as e^{2u} leaves range.
__device__ __forceinline__ float gelu(float z, float& dg) {
  const float u = kGeluK * (z + kGeluC * z * z * z);
  const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * u) + 1.0f);
  dg = 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluK * (1.0f + 3.0f * kGeluC * z * z);
  return 0.5f * z * (1.0f + t);
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

__global__ __launch_bounds__(256) void k_mlp_seed(int64_t n, MasterBufs D, float* __restrict__ hs,
                                                  const EosDev* __restrict__ eos, SegView sv, float eta,
                                                  const uint16_t* __restrict__ wb, float margin) {
  __shared__ float sw[MLP_NPARAM];
  load_weights(sw, wb);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t sg = sv.of(t);
  const float hcap = sv.hcap[sg];
  float x[MLP_IN], h;
  features(t, D, eos, sv.clk[sg].dt, eta, x, h);
  const float y = forward(sw, x) + MLP_PHYS * x[0];
  float seed = h * __expf(fminf(fmaxf(y, -50.f), 50.f) * (1.0f / MLP_YSCALE));
  seed = fminf(fmaxf(seed, h * (1.0f / MLP_CLAMP)), h * MLP_CLAMP);
  seed = fminf(seed, fmaxf(hcap, h));
  D.h[t] = seed;
  hs[t] = fminf(seed * margin, fmaxf(hcap, seed));
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
  load_weights(sw, wb);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float xx[MLP_IN];
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) xx[j] = x[t * MLP_IN + j];
  out[t] = forward(sw, xx);
}

inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

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
  k_mlp_seed<<<g256(n), 256, 0, s>>>(n, D, hs, (const EosDev*)eos, sv, eta, m.wb, margin);
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
  k_mlp_forward<<<g256(n), 256, 0, s>>>(m.wb, x, n, out);
}

void mlp_sync_bf16(hipStream_t s, MlpDev& m) { k_mlp_to_bf16<<<g256(MLP_NPARAM), 256, 0, s>>>(m.w, m.wb); }
