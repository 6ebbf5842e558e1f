// Shared device helpers and launch declarations for libihs_sph (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IHS_WAVE 64
#define IHS_NONE 0xffffffffu

constexpr float kG32 = 6.674e-11f;  // ie/shaders/bh_gravity.wgsl:11
constexpr double kG64 = 6.674e-11;  // ie/crates/engine/src/orbit.rs:18
constexpr float kPI32 = 3.14159265359f;

// ---------------------------------------------------------------------------------------------
// Order-preserving float <-> u32 keys (the reference's bbox_key / bbox_decode,
// ie/shaders/bh_gravity.wgsl:49-57): unsigned order of the key equals the float order.
// Split precision keeps each position as hi + lo: hi is the FP32 rounding of the FP64 coordinate (the working copy
// xm the passes read) and lo the remainder, held per axis as a 10-bit signed multiple of ulp(hi) / 512 in one 32-bit
// word per particle (x in bits 0-9, y in 10-19, z in 20-29). |lo| <= ulp(hi) / 2 bounds the multiple by 256, and the
// stored remainder is within ulp(hi) / 1024 of the true one (1 mm at 1e7 m).
__device__ __forceinline__ uint32_t lo_code(double x, float hi) {
  const int e = __builtin_amdgcn_frexp_expf(hi);  // hi = f 2^e, 0.5 <= |f| < 1: ulp(hi) = 2^(e - 24)
  const int q = (int)rint(ldexp(x - (double)hi, 33 - e));
  return (uint32_t)q & 0x3ffu;
}

__device__ __forceinline__ float lo_axis(uint32_t code, int shift, float hi) {
  const int q = (int)(code << (22 - shift)) >> 22;  // the signed 10-bit field at `shift`
  return __builtin_amdgcn_ldexpf((float)q, __builtin_amdgcn_frexp_expf(hi) - 33);
}

__device__ __forceinline__ float3 lo3(uint32_t code, const float4 hi) {
  return make_float3(lo_axis(code, 0, hi.x), lo_axis(code, 10, hi.y), lo_axis(code, 20, hi.z));
}

// The displacement a - b of two particles: FP32 positions (REL = false), or hi/lo positions with remainders la
// (decoded) and code cb. For neighbours the hi difference is exact (the coordinates agree in sign and lie within a
// factor of two of each other, or are both small) and the lo difference is tiny, so only the final add rounds: the
// displacement carries FP32 precision of its own size, not of the absolute position.
template <bool REL>
__device__ __forceinline__ float3 pair_d(const float4 a, const float3 la, const float4 b, uint32_t cb) {
  if (!REL) return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
  const float3 lb = lo3(cb, b);
  return make_float3((a.x - b.x) + (la.x - lb.x), (a.y - b.y) + (la.y - lb.y), (a.z - b.z) + (la.z - lb.z));
}

__device__ __forceinline__ uint32_t f2key(float f) {
  uint32_t k = __float_as_uint(f);
  return (k >> 31) ? ~k : (k ^ 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  return (k >> 31) ? __uint_as_float(k ^ 0x80000000u) : __uint_as_float(~k);
}

// Lanes below this lane whose bit is set in `mask` (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Orders this wave's LDS traffic (a wave executes its LDS ops in order; this stops the compiler
// from reordering across the point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave-uniform value moved to a scalar register (the first active lane's copy).
__device__ __forceinline__ float sgpr_f(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
// Exclusive prefix sum over the 64 lanes of a wave (lane order).
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  return x - v;
}

// Exclusive scan of 256 values held one per thread (block of 256); returns this thread's prefix.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_tmp /*[4]*/, uint32_t* total) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
#pragma unroll
  for (int ww = 0; ww < 4; ++ww) {
    const uint32_t t = s_tmp[ww];
    if (ww < w) wbase += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return wbase + x - v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Blocks are dispatched round-robin over the 8 XCDs; remap so each XCD works a contiguous range of
// logical blocks (spatially adjacent groups share tree nodes and particles in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, slot = b >> 3;
  const uint32_t per = nb >> 3, extra = nb & 7u;
  const uint32_t base = xcd * per + (xcd < extra ? xcd : extra);
  return base + slot;
}

// Cubic spline of ie/shaders/sph_step.wgsl:64-77 (support r < h), FP32.
__device__ __forceinline__ float sph_w32(float r, float h) {
  const float q = r / h;
  const float sig = 8.0f / (kPI32 * h * h * h);
  if (q < 0.5f) return sig * (1.0f - 6.0f * q * q + 6.0f * q * q * q);
  if (q < 1.0f) {
    const float t = 1.0f - q;
    return sig * 2.0f * t * t * t;
  }
  return 0.0f;
}
__device__ __forceinline__ float sph_dw32(float r, float h) {
  const float q = r / h;
  const float sig = 8.0f / (kPI32 * h * h * h);
  if (q < 0.5f) return sig * (-12.0f * q + 18.0f * q * q) / h;
  if (q < 1.0f) {
    const float t = 1.0f - q;
    return sig * (-6.0f * t * t) / h;
  }
  return 0.0f;
}

// Tillotson EOS record as stored on device (SphEos, gpu_sph.rs:39-55).
struct EosDev {
  float rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta, p0, p1;
};

// ie/shaders/sph_step.wgsl:78-89 (FP32).
__device__ __forceinline__ float till_pressure(const EosDev& e, float rho, float u) {
  const float r = fmaxf(rho, 1.0e-9f);
  const float eta = r / e.rho0;
  const float mu = eta - 1.0f;
  const float omega = u / (e.e0 * eta * eta) + 1.0f;
  const float p_c = (e.a + e.b / omega) * r * u + e.cap_a * mu + e.cap_b * mu * mu;
  if (eta >= 1.0f || u <= e.e_iv) return p_c;
  const float z = e.rho0 / r - 1.0f;
  const float p_e = e.a * r * u + (e.b * r * u / omega + e.cap_a * mu * __expf(-e.beta * z)) * __expf(-e.alpha * z * z);
  if (u >= e.e_cv) return p_e;
  return ((u - e.e_iv) * p_e + (e.e_cv - u) * p_c) / (e.e_cv - e.e_iv);
}
// ie/shaders/sph_step.wgsl:90-100 (central differences, steps 1e-3).
__device__ __forceinline__ float till_sound_speed(const EosDev& e, float rho, float u, float p) {
  const float r = fmaxf(rho, 1.0e-9f);
  const float dr = r * 1.0e-3f;
  const float dp = (till_pressure(e, r + dr, u) - till_pressure(e, r - dr, u)) / (2.0f * dr);
  const float du = fabsf(u) * 1.0e-3f + 1.0f;
  const float dpu = (till_pressure(e, r, u + du) - till_pressure(e, r, u - du)) / (2.0f * du);
  return sqrtf(fmaxf(dp + p / (r * r) * dpu, 0.0f));
}

// Kahan-compensated FP32 accumulator (used for per-particle force sums in split precision).
struct Kahan3 {
  float s[3], c[3];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int q = 0; q < 3; ++q) s[q] = c[q] = 0.f;
  }
  __device__ __forceinline__ void add(float x, int q) {
    const float y = x - c[q];
    const float t = s[q] + y;
    c[q] = (t - s[q]) - y;
    s[q] = t;
  }
};

// Blocks for a persistent launch: every CU filled to the kernel's occupancy (a multiple of the 8 XCDs).
inline unsigned persistent_blocks(const void* kernel, int block) {
  int dev = 0, cus = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0) != hipSuccess) occ = 0;
  const unsigned nb = (unsigned)((cus > 0 ? cus : 256) * (occ > 0 ? occ : 1));
  return (nb + 7u) & ~7u;
}
