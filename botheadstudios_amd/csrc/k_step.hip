// State permutation, leapfrog integration and FP64 reductions (gfx950).
//
// Integration follows the reference's KDK (ie/shaders/sph_step.wgsl:213-237; ie/crates/engine/src/
// hydrostatic.rs:345-359): v += a dt/2; u = max(u + du dt/2, 0); x += v dt; ...; v += a dt/2; u = ....
// In split precision the position and internal-energy updates run in FP64 (PT = double); otherwise
// the whole update is FP32 as in the reference. The Courant dt (ie/crates/engine/src/hydrostatic.rs:
// 366-374; impact-run:345-356) is formed on the device from the density pass's min-signal key, so a
// step needs no host round trip.
#include "common.h"
#include "launch.h"

namespace {

inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

__global__ void k_fe_init(uint32_t* flags, SegView sv) {
  if (threadIdx.x == 0) {
    flags[FLAG_H_RETRY] = 0u;
    flags[FLAG_TIE_RUNS] = 0u;
  }
  for (int sg = threadIdx.x; sg < sv.S; sg += blockDim.x) sv.clk[sg].signal = 0xffffffffu;
}

// Wider-radius pass of a seeded evaluation (runs only when the tight pass stopped some solve at its radius).
// It also re-arms the upsweep's arrival counters over the kept topology (ready for the L - 1 internal nodes, the
// pending count) and clears the h-solve's pending lists, for the second upsweep and solve.
__global__ __launch_bounds__(256) void k_retry_hs(int64_t n, const float* __restrict__ h, float* __restrict__ hs,
                                                  float margin, SegView sv, uint32_t* __restrict__ flags, int64_t L,
                                                  uint32_t* __restrict__ ready, uint32_t* __restrict__ npend,
                                                  uint32_t* __restrict__ hs_npend, int n_hs_npend) {
  if (flags[FLAG_H_RETRY] == 0u) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    atomicAdd(&flags[FLAG_RETRIES], 1u);
    *npend = 0u;
  }
  if (t < n_hs_npend) hs_npend[t] = 0u;
  if (t < L - 1) ready[t] = 0u;
  if (t >= n) return;
  const float hh = h[t];
  hs[t] = fminf(hh * margin, fmaxf(sv.hcap[sv.of(t)], hh));
}

template <class PT>
__global__ __launch_bounds__(256) void k_gather(const uint32_t* __restrict__ perm, int64_t n, MasterBufs S, MasterBufs D,
                                                float4* __restrict__ xm, float4* __restrict__ vh,
                                                float* __restrict__ hs, float margin, SegView sv,
                                                uint32_t* __restrict__ xlo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t s = perm[t];
  const PT x = ((const PT*)S.px)[s], y = ((const PT*)S.py)[s], z = ((const PT*)S.pz)[s];
  ((PT*)D.px)[t] = x;
  ((PT*)D.py)[t] = y;
  ((PT*)D.pz)[t] = z;
  const float vx = S.vx[s], vy = S.vy[s], vz = S.vz[s];
  D.vx[t] = vx;
  D.vy[t] = vy;
  D.vz[t] = vz;
  ((PT*)D.u)[t] = ((const PT*)S.u)[s];
  const float h = S.h[s];
  D.h[t] = h;
  const float m = S.m[s];
  D.m[t] = m;
  D.mat[t] = S.mat[s];
  D.id[t] = S.id[s];
  D.prov[t] = S.prov[s];
  D.rho[t] = S.rho[s];
  D.feat[t] = S.feat[s];
  D.hiter[t] = S.hiter[s];
  const float hx = (float)x, hy = (float)y, hz = (float)z;
  xm[t] = make_float4(hx, hy, hz, m);
  if (xlo) xlo[t] = lo_code((double)x, hx) | (lo_code((double)y, hy) << 10) | (lo_code((double)z, hz) << 20);
  vh[t] = make_float4(vx, vy, vz, h);
  hs[t] = fminf(h * margin, fmaxf(sv.hcap[sv.of(t)], h));
}

// An evaluation at positions the previous evaluation sorted, none of which has moved since: a sort would return the
// identity, so the state stays in its buffers and only the working copies are written, as k_gather writes them.
// Block 0 also performs the evaluation's reset (k_bbox's, in a sorting evaluation), and the threads below L - 1
// re-arm the internal nodes' arrival counters for the upsweep over the kept topology (k_karras's, otherwise).
template <class PT>
__global__ __launch_bounds__(256) void k_stage(int64_t n, MasterBufs S, float4* __restrict__ xm,
                                               float4* __restrict__ vh, float* __restrict__ hs, float margin,
                                               SegView sv, uint32_t* __restrict__ xlo, EvalInit fe, int64_t L,
                                               uint32_t* __restrict__ ready, uint32_t* __restrict__ npend) {
  if (blockIdx.x == 0) eval_init(fe, sv.S);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *npend = 0u;
  if (t < L - 1) ready[t] = 0u;
  if (t >= n) return;
  const PT x = ((const PT*)S.px)[t], y = ((const PT*)S.py)[t], z = ((const PT*)S.pz)[t];
  const float h = S.h[t];
  const float hx = (float)x, hy = (float)y, hz = (float)z;
  xm[t] = make_float4(hx, hy, hz, S.m[t]);
  if (xlo) xlo[t] = lo_code((double)x, hx) | (lo_code((double)y, hy) << 10) | (lo_code((double)z, hz) << 20);
  vh[t] = make_float4(S.vx[t], S.vy[t], S.vz[t], h);
  hs[t] = fminf(h * margin, fmaxf(sv.hcap[sv.of(t)], h));
}

// The step of segment sg: fixed, or the Courant dt from the segment's signal of this evaluation.
__device__ __forceinline__ double step_dt(const StepCtl* ctl, const SegView& sv, uint32_t sg) {
  return ctl->dt_fixed > 0.0 ? ctl->dt_fixed : ctl->cfl * (double)key2f(sv.clk[sg].signal);
}

// The segment's first storage position advances its clock.
__device__ __forceinline__ void clock_tick(const SegView& sv, uint32_t sg, int64_t t, double dt) {
  if (t == (int64_t)sv.start[sg]) {
    sv.clk[sg].dt = dt;
    sv.clk[sg].t += dt;
    sv.clk[sg].steps += 1;
  }
}

template <class PT>
__device__ __forceinline__ PT advance(PT x, float v, double dt) {
  return (PT)(x + (PT)v * (PT)dt);
}

template <class PT>
__global__ __launch_bounds__(256) void k_kick_drift(int64_t n, MasterBufs S, const float4* __restrict__ asph,
                                                    const float4* __restrict__ accg, const StepCtl* __restrict__ ctl,
                                                    SegView sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t sg = sv.of(t);
  const double dt = step_dt(ctl, sv, sg);
  clock_tick(sv, sg, t, dt);
  const float4 a = asph[t], g = accg[t];
  const float hdt = (float)(0.5 * dt);
  const float vx = S.vx[t] + (a.x + g.x) * hdt;
  const float vy = S.vy[t] + (a.y + g.y) * hdt;
  const float vz = S.vz[t] + (a.z + g.z) * hdt;
  S.vx[t] = vx;
  S.vy[t] = vy;
  S.vz[t] = vz;
  PT* u = (PT*)S.u;
  const PT un = u[t] + (PT)a.w * (PT)(0.5 * dt);
  u[t] = un > (PT)0 ? un : (PT)0;
  ((PT*)S.px)[t] = advance<PT>(((PT*)S.px)[t], vx, dt);
  ((PT*)S.py)[t] = advance<PT>(((PT*)S.py)[t], vy, dt);
  ((PT*)S.pz)[t] = advance<PT>(((PT*)S.pz)[t], vz, dt);
}

template <class PT>
__global__ __launch_bounds__(256) void k_kick(int64_t n, MasterBufs S, const float4* __restrict__ asph,
                                              const float4* __restrict__ accg, SegView sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double dt = sv.clk[sv.of(t)].dt;
  const float4 a = asph[t], g = accg[t];
  const float hdt = (float)(0.5 * dt);
  S.vx[t] += (a.x + g.x) * hdt;
  S.vy[t] += (a.y + g.y) * hdt;
  S.vz[t] += (a.z + g.z) * hdt;
  PT* u = (PT*)S.u;
  const PT un = u[t] + (PT)a.w * (PT)(0.5 * dt);
  u[t] = un > (PT)0 ? un : (PT)0;
}

// ie/shaders/sph_step.wgsl:244-257 (damped relaxation in a frame rotating at omega about +z).
template <class PT>
__global__ __launch_bounds__(256) void k_relax(int64_t n, MasterBufs S, const float4* __restrict__ asph,
                                               const float4* __restrict__ accg, const StepCtl* __restrict__ ctl,
                                               SegView sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double dt = ctl->dt_fixed;
  clock_tick(sv, sv.of(t), t, dt);
  const float4 a = asph[t], g = accg[t];
  const float w2 = ctl->omega * ctl->omega;
  const float fdt = (float)dt, damp = ctl->damp;
  PT* px = (PT*)S.px;
  PT* py = (PT*)S.py;
  PT* pz = (PT*)S.pz;
  const float x = (float)px[t], y = (float)py[t];
  const float vx = (S.vx[t] + (a.x + g.x + w2 * x) * fdt) * damp;
  const float vy = (S.vy[t] + (a.y + g.y + w2 * y) * fdt) * damp;
  const float vz = (S.vz[t] + (a.z + g.z) * fdt) * damp;
  S.vx[t] = vx;
  S.vy[t] = vy;
  S.vz[t] = vz;
  px[t] = advance<PT>(px[t], vx, dt);
  py[t] = advance<PT>(py[t], vy, dt);
  pz[t] = advance<PT>(pz[t], vz, dt);
}

constexpr int EN_BLOCKS = 1024;

// Per-block FP64 partials of KE, IE, PE (= 1/2 m phi), momentum and angular momentum about the origin.
template <class PT>
__global__ __launch_bounds__(256) void k_energy_part(int64_t n, MasterBufs S, const float4* __restrict__ accg,
                                                     double* __restrict__ part) {
  double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double m = S.m[t];
    const double vx = S.vx[t], vy = S.vy[t], vz = S.vz[t];
    const double x = ((const PT*)S.px)[t], y = ((const PT*)S.py)[t], z = ((const PT*)S.pz)[t];
    q[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    q[1] += m * (double)((const PT*)S.u)[t];
    q[2] += 0.5 * m * (double)accg[t].w;
    q[3] += m * vx;
    q[4] += m * vy;
    q[5] += m * vz;
    q[6] += m * (y * vz - z * vy);
    q[7] += m * (z * vx - x * vz);
    q[8] += m * (x * vy - y * vx);
  }
  __shared__ double s[4][9];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = wave_sum(q[k]);
  if (lane == 0)
    for (int k = 0; k < 9; ++k) s[w][k] = q[k];
  __syncthreads();
  if (threadIdx.x < 9) part[(int64_t)blockIdx.x * 9 + threadIdx.x] = s[0][threadIdx.x] + s[1][threadIdx.x] +
                                                                     s[2][threadIdx.x] + s[3][threadIdx.x];
}

// Fixed-order final sum (deterministic): 256 threads take strided block partials, then a fixed tree.
__global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
  __shared__ double s[9][256];
  double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 9; ++k) q[k] += part[(int64_t)b * 9 + k];
  for (int k = 0; k < 9; ++k) s[k][threadIdx.x] = q[k];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 9; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 9) out[threadIdx.x] = s[threadIdx.x][0];
}

template <class PT>
__global__ __launch_bounds__(256) void k_unpack(int64_t n, const SphParticle48* __restrict__ in, MasterBufs S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SphParticle48 p = in[i];
  ((PT*)S.px)[i] = (PT)p.pos[0];
  ((PT*)S.py)[i] = (PT)p.pos[1];
  ((PT*)S.pz)[i] = (PT)p.pos[2];
  S.vx[i] = p.vel[0];
  S.vy[i] = p.vel[1];
  S.vz[i] = p.vel[2];
  ((PT*)S.u)[i] = (PT)p.u;
  S.h[i] = p.h;
  S.m[i] = p.mass;
  S.mat[i] = (uint8_t)min(p.mat, 15u);
  S.id[i] = (uint32_t)i;
  S.prov[i] = p.prov;
  S.rho[i] = p.rho;
  S.feat[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  S.hiter[i] = 0;
}

template <class PT>
__global__ __launch_bounds__(256) void k_pack(int64_t n, MasterBufs S, SphParticle48* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  SphParticle48 p;
  p.pos[0] = (float)((const PT*)S.px)[t];
  p.pos[1] = (float)((const PT*)S.py)[t];
  p.pos[2] = (float)((const PT*)S.pz)[t];
  p.h = S.h[t];
  p.vel[0] = S.vx[t];
  p.vel[1] = S.vy[t];
  p.vel[2] = S.vz[t];
  p.u = (float)((const PT*)S.u)[t];
  p.mass = S.m[t];
  p.mat = S.mat[t];
  p.rho = S.rho[t];
  p.prov = S.prov[t];
  out[S.id[t]] = p;
}

template <class T>
__global__ __launch_bounds__(256) void k_scatter(int64_t n, const uint32_t* __restrict__ id, const T* __restrict__ src,
                                                 T* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[id[t]] = src[t];
}

template <class PT>
__global__ __launch_bounds__(256) void k_pos_f64(int64_t n, MasterBufs S, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t i = S.id[t];
  out[3 * (int64_t)i] = (double)((const PT*)S.px)[t];
  out[3 * (int64_t)i + 1] = (double)((const PT*)S.py)[t];
  out[3 * (int64_t)i + 2] = (double)((const PT*)S.pz)[t];
}

// assemble_from_relaxed_n (gpu_sph.rs:318-339) in place over storage positions [t0, t1): body k = prov is
// recentred on its robust centre, moved to its offset and set moving at its bulk velocity plus spin x (r - centre).
// pl holds 12 doubles per body: centre, offset, velocity, spin. Positions and velocities are rounded to FP32, as
// the reference's SphParticle output is; minh receives the order-preserving key of the smallest h.
template <class PT>
__global__ __launch_bounds__(256) void k_assemble(int64_t t0, int64_t t1, MasterBufs S, const double* __restrict__ pl,
                                                  uint32_t* __restrict__ minh) {
  const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // The smallest h of the wave first, so one lane per wave meets the single address.
  const uint32_t hk = wave_min(t < t1 ? f2key(S.h[t]) : 0xffffffffu);
  if ((threadIdx.x & 63) == 0 && hk != 0xffffffffu) atomicMin(minh, hk);
  if (t >= t1) return;
  const double* q = pl + 12 * S.prov[t];
  PT* px = (PT*)S.px;
  PT* py = (PT*)S.py;
  PT* pz = (PT*)S.pz;
  const double l0 = (double)(float)px[t] - q[0], l1 = (double)(float)py[t] - q[1], l2 = (double)(float)pz[t] - q[2];
  px[t] = (PT)(float)(l0 + q[3]);
  py[t] = (PT)(float)(l1 + q[4]);
  pz[t] = (PT)(float)(l2 + q[5]);
  // glam DVec3::cross: (s.y l.z - l.y s.z, s.z l.x - l.z s.x, s.x l.y - l.x s.y)
  S.vx[t] = (float)(q[6] + (q[10] * l2 - l1 * q[11]));
  S.vy[t] = (float)(q[7] + (q[11] * l0 - l2 * q[9]));
  S.vz[t] = (float)(q[8] + (q[9] * l1 - l0 * q[10]));
}

// Largest provenance over storage positions [t0, t1).
__global__ __launch_bounds__(256) void k_prov_max(int64_t t0, int64_t t1, const uint32_t* __restrict__ prov,
                                                  uint32_t* __restrict__ out) {
  const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = t < t1 ? prov[t] : 0u;
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0 && v) atomicMax(out, v);
}

}  // namespace

void launch_fe_init(hipStream_t s, uint32_t* flags, const SegView& sv) { k_fe_init<<<1, 256, 0, s>>>(flags, sv); }

void launch_retry_hs(hipStream_t s, int64_t n, const float* h, float* hs, float margin, const SegView& sv,
                     uint32_t* flags, int64_t L, uint32_t* ready, uint32_t* npend, uint32_t* hs_npend,
                     int n_hs_npend) {
  k_retry_hs<<<std::max(1u, g256(n)), 256, 0, s>>>(n, h, hs, margin, sv, flags, L, ready, npend, hs_npend,
                                                   n_hs_npend);
}

void launch_gather(hipStream_t s, int split, int64_t n, const uint32_t* perm, const MasterBufs& src,
                   const MasterBufs& dst, float4* xm, float4* vh, float* hs, float h_margin, const SegView& sv,
                   uint32_t* xlo) {
  if (split) k_gather<double><<<g256(n), 256, 0, s>>>(perm, n, src, dst, xm, vh, hs, h_margin, sv, xlo);
  else k_gather<float><<<g256(n), 256, 0, s>>>(perm, n, src, dst, xm, vh, hs, h_margin, sv, nullptr);
}

void launch_stage(hipStream_t s, int split, int64_t n, const MasterBufs& st, float4* xm, float4* vh, float* hs,
                  float h_margin, const SegView& sv, uint32_t* xlo, const EvalInit& fe, int64_t L, uint32_t* ready,
                  uint32_t* npend) {
  if (split) k_stage<double><<<g256(n), 256, 0, s>>>(n, st, xm, vh, hs, h_margin, sv, xlo, fe, L, ready, npend);
  else k_stage<float><<<g256(n), 256, 0, s>>>(n, st, xm, vh, hs, h_margin, sv, nullptr, fe, L, ready, npend);
}

void launch_kick_drift(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph,
                       const float4* accg, const StepCtl* ctl, const SegView& sv) {
  const unsigned g = std::max(1u, g256(n));
  if (split) k_kick_drift<double><<<g, 256, 0, s>>>(n, st, asph, accg, ctl, sv);
  else k_kick_drift<float><<<g, 256, 0, s>>>(n, st, asph, accg, ctl, sv);
}

void launch_kick(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph, const float4* accg,
                 const SegView& sv) {
  if (split) k_kick<double><<<g256(n), 256, 0, s>>>(n, st, asph, accg, sv);
  else k_kick<float><<<g256(n), 256, 0, s>>>(n, st, asph, accg, sv);
}

void launch_relax(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph, const float4* accg,
                  const StepCtl* ctl, const SegView& sv) {
  const unsigned g = std::max(1u, g256(n));
  if (split) k_relax<double><<<g, 256, 0, s>>>(n, st, asph, accg, ctl, sv);
  else k_relax<float><<<g, 256, 0, s>>>(n, st, asph, accg, ctl, sv);
}

void launch_prov_max(hipStream_t s, int64_t t0, int64_t t1, const MasterBufs& st, uint32_t* out) {
  k_prov_max<<<std::max(1u, g256(t1 - t0)), 256, 0, s>>>(t0, t1, st.prov, out);
}

void launch_assemble(hipStream_t s, int split, int64_t t0, int64_t t1, const MasterBufs& st, const double* pl,
                     uint32_t* minh) {
  const unsigned g = std::max(1u, g256(t1 - t0));
  if (split) k_assemble<double><<<g, 256, 0, s>>>(t0, t1, st, pl, minh);
  else k_assemble<float><<<g, 256, 0, s>>>(t0, t1, st, pl, minh);
}

void launch_energy(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* accg, double* part,
                   double* out) {
  const int nb = (int)std::min<int64_t>(EN_BLOCKS, std::max<int64_t>(1, (n + 255) / 256));
  if (split) k_energy_part<double><<<nb, 256, 0, s>>>(n, st, accg, part);
  else k_energy_part<float><<<nb, 256, 0, s>>>(n, st, accg, part);
  k_energy_final<<<1, 256, 0, s>>>(part, nb, out);
}

void launch_unpack(hipStream_t s, int split, int64_t n, const SphParticle48* in, const MasterBufs& st) {
  if (split) k_unpack<double><<<g256(n), 256, 0, s>>>(n, in, st);
  else k_unpack<float><<<g256(n), 256, 0, s>>>(n, in, st);
}

void launch_pack(hipStream_t s, int split, int64_t n, const MasterBufs& st, SphParticle48* out) {
  if (split) k_pack<double><<<g256(n), 256, 0, s>>>(n, st, out);
  else k_pack<float><<<g256(n), 256, 0, s>>>(n, st, out);
}

void launch_scatter_by_id(hipStream_t s, int64_t n, const uint32_t* id, const void* src, void* dst, int elem) {
  if (elem == 16) k_scatter<float4><<<g256(n), 256, 0, s>>>(n, id, (const float4*)src, (float4*)dst);
  else k_scatter<uint32_t><<<g256(n), 256, 0, s>>>(n, id, (const uint32_t*)src, (uint32_t*)dst);
}

void launch_pos_f64(hipStream_t s, int split, int64_t n, const MasterBufs& st, double* out3) {
  if (split) k_pos_f64<double><<<g256(n), 256, 0, s>>>(n, st, out3);
  else k_pos_f64<float><<<g256(n), 256, 0, s>>>(n, st, out3);
}
