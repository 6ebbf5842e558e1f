// Disk / remnant measurement of an impact snapshot and friends-of-friends clumps, on the GPU.
//
// Restates, for the particles of one scenario as the readback hands them over (48-byte records, upload order):
//   ie/tools/impact-run/src/main.rs:195-246   sum_oi, com, perigee
//   ie/tools/impact-run/src/main.rs:712-878   Measure, measure, moon_candidate
//   ie/crates/engine/src/gpu_sph.rs:399-640   disk_stats_json, moonlet_bodies, largest_moonlet_orbit, disk_moonlets
//   ie/crates/engine/src/accretion.rs:73-163  find_clumps, central_density
//   ie/crates/engine/src/orbit.rs:91-116      is_bound, perigee
//
// Arithmetic: this file is compiled with -ffp-contract=off and every per-element expression keeps the
// reference's operation order, so radii, perigees, link tests and pair potentials round exactly as the
// reference's f64 code does. Sums are the one deliberate difference: the reference adds sequentially (or sorts
// by magnitude and Kahan-sums, sum_oi); here every sum is double-double (TwoSum, ~2^-104 relative) in a fixed
// order, so a result matches sum_oi to within an ulp and a sequential sum to within that sum's own rounding, and
// it does not depend on the order in which the particles are stored.
//
// Phase 1 (frame): double-double totals of m, m x, m v and h; radii |x - com| as f64 bit patterns, sorted by a
// stable LSD radix sort (low word, then high word), a double-double prefix of the sorted masses for the 85%
// crossing; orbit classification with category mass sums; disk particles compacted in index order into SoA
// arrays. Phase 2 (clumps): a uniform grid whose cell side is at least the linking length, sorted cell keys
// (x-fastest, so each row of three neighbour cells is one key range), a lock-free union-find over linked pairs
// whose roots are the lowest member index, clump segments by a stable sort on the root, per-clump sums (one
// wave per clump), and the pair potential in 256-row work items so that a large clump spreads over many CUs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "common.h"
#include "launch.h"
#include "measure.h"

namespace {

constexpr int MB_THREADS = 256;
constexpr int MB_MAXBLK = 1024;   // blocks of the fixed-order reductions and scans
constexpr int MB_TILE = 2048;     // sorted-mass prefix tile (256 threads x 8)
constexpr int PE_ROWS = 256;      // rows of a pair-potential work item
constexpr double kPI = 3.14159265358979323846;
constexpr double kFourThirdsPi = 4.0 / 3.0 * kPI;  // accretion.rs:26

// ---------------------------------------------------------------------------------------------
// Double-double sums.
struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  const double s = a.hi + b.hi;
  const double v = s - a.hi;
  const double e = (a.hi - (s - v)) + (b.hi - v);  // exact rounding error of s
  const double t = e + (a.lo + b.lo);
  const double hi = s + t;
  return dd{hi, t - (hi - s)};
}
__device__ __forceinline__ dd dd_add1(dd a, double b) {
  const double s = a.hi + b;
  const double v = s - a.hi;
  const double e = (a.hi - (s - v)) + (b - v);
  const double t = e + a.lo;
  const double hi = s + t;
  return dd{hi, t - (hi - s)};
}
__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const dd w{__shfl_xor(v.hi, o, 64), __shfl_xor(v.lo, o, 64)};
    v = dd_add(v, w);
  }
  return v;
}

// Sums NS double-double slots over a 256-thread block; thread s < NS receives slot s's total.
template <int NS>
__device__ __forceinline__ dd block_sum_dd(const dd (&a)[NS]) {
  __shared__ double s_hi[4][NS], s_lo[4][NS];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const dd v = wave_sum_dd(a[s]);
    if (lane == 0) {
      s_hi[w][s] = v.hi;
      s_lo[w][s] = v.lo;
    }
  }
  __syncthreads();
  dd t{0.0, 0.0};
  if (threadIdx.x < NS) {
    t = dd{s_hi[0][threadIdx.x], s_lo[0][threadIdx.x]};
    for (int ww = 1; ww < 4; ++ww) t = dd_add(t, dd{s_hi[ww][threadIdx.x], s_lo[ww][threadIdx.x]});
  }
  __syncthreads();
  return t;
}

__device__ __forceinline__ uint64_t d2key(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key2d(uint64_t k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

// Contiguous per-block ranges for the order-sensitive passes (compaction, scans).
struct Chunks {
  int nblk;
  int64_t chunk;
};
Chunks chunks_for(int64_t n) {
  int64_t nb = std::max<int64_t>(1, std::min<int64_t>(MB_MAXBLK, (n + 2047) / 2048));
  int64_t ch = (n + nb - 1) / nb;
  ch = std::max<int64_t>(MB_THREADS, (ch + MB_THREADS - 1) / MB_THREADS * MB_THREADS);
  nb = std::max<int64_t>(1, (n + ch - 1) / ch);
  return Chunks{(int)nb, ch};
}
inline unsigned g256(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// ---------------------------------------------------------------------------------------------
// Exclusive scans of u32 values over [0, n) in contiguous chunks.
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n, int64_t chunk,
                                                     uint32_t* __restrict__ bsum) {
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  uint32_t s = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += MB_THREADS) s += in[i];
  s = wave_sum(s);
  __shared__ uint32_t w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

// One block: exclusive scan of nblk <= 1024 block sums in place; *total receives the sum.
__global__ __launch_bounds__(256) void k_scan_top(uint32_t* __restrict__ bsum, int nblk, uint32_t* __restrict__ total) {
  __shared__ uint32_t s_tmp[4];
  uint32_t carry = 0;
  for (int c0 = 0; c0 < nblk; c0 += 256) {
    const int c = c0 + threadIdx.x;
    const uint32_t v = c < nblk ? bsum[c] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan_256(v, s_tmp, &tot);
    if (c < nblk) bsum[c] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in, int64_t n, int64_t chunk,
                                                    const uint32_t* __restrict__ boff, uint32_t* __restrict__ out) {
  __shared__ uint32_t s_tmp[4];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  uint32_t run = boff[blockIdx.x];
  for (int64_t base = b0; base < b1; base += MB_THREADS) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < b1 ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan_256(v, s_tmp, &tot);
    if (i < b1) out[i] = run + ex;
    run += tot;
  }
}

void launch_exscan(hipStream_t s, const uint32_t* in, int64_t n, uint32_t* out, uint32_t* bsum, uint32_t* total) {
  const Chunks c = chunks_for(n);
  k_scan_reduce<<<c.nblk, MB_THREADS, 0, s>>>(in, n, c.chunk, bsum);
  k_scan_top<<<1, MB_THREADS, 0, s>>>(bsum, c.nblk, total);
  k_scan_apply<<<c.nblk, MB_THREADS, 0, s>>>(in, n, c.chunk, bsum, out);
}

// ---------------------------------------------------------------------------------------------
// Phase 1: frame.

// Block partial sums of m, m x (3), m v (3), h over a grid-strided range (impact-run main.rs:207-215 com,
// main.rs:733-737; gpu_sph.rs:406-410).
__global__ __launch_bounds__(256) void k_frame_sums(const SphParticle48* __restrict__ P, int64_t n,
                                                    dd* __restrict__ part) {
  dd a[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) a[s] = dd{0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * MB_THREADS) {
    const SphParticle48 p = P[i];
    const double m = (double)p.mass;
    a[0] = dd_add1(a[0], m);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      a[1 + q] = dd_add1(a[1 + q], (double)p.pos[q] * m);  // f32 x f32 products are exact in f64
      a[4 + q] = dd_add1(a[4 + q], (double)p.vel[q] * m);
    }
    a[7] = dd_add1(a[7], (double)p.h);
  }
  const dd t = block_sum_dd<8>(a);
  if (threadIdx.x < 8) part[(int64_t)blockIdx.x * 8 + threadIdx.x] = t;
}

// Sums nslot-slot block partials (fixed order) into tot[0..nslot).
template <int NS>
__device__ void sum_partials(const dd* __restrict__ part, int nblk, dd* tot) {
  dd a[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) a[s] = dd{0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += MB_THREADS)
#pragma unroll
    for (int s = 0; s < NS; ++s) a[s] = dd_add(a[s], part[(int64_t)b * NS + s]);
  const dd t = block_sum_dd<NS>(a);
  __shared__ dd s_t[NS];
  if (threadIdx.x < NS) s_t[threadIdx.x] = t;
  __syncthreads();
  for (int s = 0; s < NS; ++s) tot[s] = s_t[s];
}

__global__ __launch_bounds__(256) void k_frame_final(const dd* __restrict__ part, int nblk, MeasureFrame* fr) {
  dd t[8];
  sum_partials<8>(part, nblk, t);
  if (threadIdx.x == 0) {
    const double m = t[0].hi;
    fr->m_total = m;
    for (int q = 0; q < 3; ++q) {
      fr->com[q] = t[1 + q].hi / m;
      fr->v_com[q] = t[4 + q].hi / m;
    }
    fr->sum_h = t[7].hi;
    fr->r_remnant = 0.0;
    fr->m_remnant = m;
    fr->n_disk = 0;
    fr->n_clumps = 0;
    fr->n_items = 0;
  }
}

// |x - com| (main.rs:739; gpu_sph.rs:412: sqrt(dx dx + dy dy + dz dz)).
__device__ __forceinline__ double radius_of(const SphParticle48& p, const MeasureFrame* fr) {
  const double dx = (double)p.pos[0] - fr->com[0];
  const double dy = (double)p.pos[1] - fr->com[1];
  const double dz = (double)p.pos[2] - fr->com[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__global__ __launch_bounds__(256) void k_radius_keys(const SphParticle48* __restrict__ P, int64_t n,
                                                     const MeasureFrame* __restrict__ fr, uint32_t* __restrict__ klo,
                                                     uint32_t* __restrict__ khi, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i >= n) return;
  const uint64_t b = (uint64_t)__double_as_longlong(radius_of(P[i], fr));  // r >= 0: bit order = value order
  klo[i] = (uint32_t)b;
  khi[i] = (uint32_t)(b >> 32);
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_gather_u32(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ src,
                                                    int64_t n, uint32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// Per tile of 2048 sorted particles: double-double mass sum (thread order, 8 consecutive per thread).
__global__ __launch_bounds__(256) void k_mass_tiles(const SphParticle48* __restrict__ P, const uint32_t* __restrict__ perm,
                                                    int64_t n, dd* __restrict__ tsum) {
  const int64_t base = (int64_t)blockIdx.x * MB_TILE + threadIdx.x * 8;
  dd a[1] = {dd{0.0, 0.0}};
  for (int j = 0; j < 8; ++j)
    if (base + j < n) a[0] = dd_add1(a[0], (double)P[perm[base + j]].mass);
  const dd t = block_sum_dd<1>(a);
  if (threadIdx.x == 0) tsum[blockIdx.x] = t;
}

// One block: the first sorted particle whose running mass reaches frac m_total (0.85: main.rs:738-747,
// gpu_sph.rs:411-420; 0.98: body_bulk, gpu_sph.rs:293-304). r_remnant = its radius, m_remnant = the running mass
// there; without a crossing the largest radius and m_total, as the reference initialises them.
__global__ __launch_bounds__(256) void k_remnant(const SphParticle48* __restrict__ P, const uint32_t* __restrict__ perm,
                                                 int64_t n, const dd* __restrict__ tsum, int ntile, double frac,
                                                 MeasureFrame* __restrict__ fr) {
  __shared__ int s_tile;
  __shared__ double s_base_hi, s_base_lo;
  __shared__ double s_hi[256], s_lo[256];
  __shared__ int64_t s_best;
  const double thr = frac * fr->m_total;
  // Inclusive prefix over threads of the tile sums each thread owns, Hillis-Steele in LDS.
  auto block_inclusive = [&](dd v) -> dd {
    s_hi[threadIdx.x] = v.hi;
    s_lo[threadIdx.x] = v.lo;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      dd w{s_hi[threadIdx.x], s_lo[threadIdx.x]};
      if ((int)threadIdx.x >= o) w = dd_add(dd{s_hi[threadIdx.x - o], s_lo[threadIdx.x - o]}, w);
      __syncthreads();
      s_hi[threadIdx.x] = w.hi;
      s_lo[threadIdx.x] = w.lo;
      __syncthreads();
    }
    dd ex{0.0, 0.0};
    if (threadIdx.x > 0) ex = dd{s_hi[threadIdx.x - 1], s_lo[threadIdx.x - 1]};
    __syncthreads();
    return ex;  // exclusive prefix of this thread
  };
  {
    if (threadIdx.x == 0) {
      s_tile = INT32_MAX;
      s_best = INT64_MAX;
    }
    const int per = (ntile + 255) / 256;
    const int q0 = min(ntile, (int)threadIdx.x * per), q1 = min(ntile, q0 + per);
    dd own{0.0, 0.0};
    for (int q = q0; q < q1; ++q) own = dd_add(own, tsum[q]);
    dd c = block_inclusive(own);
    int found = INT32_MAX;
    dd fbase{0.0, 0.0};
    for (int q = q0; q < q1; ++q) {
      const dd nx = dd_add(c, tsum[q]);
      if (nx.hi >= thr) {
        found = q;
        fbase = c;
        break;
      }
      c = nx;
    }
    if (found != INT32_MAX) atomicMin(&s_tile, found);
    __syncthreads();
    if (found != INT32_MAX && found == s_tile) {
      s_base_hi = fbase.hi;
      s_base_lo = fbase.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_tile == INT32_MAX) s_tile = -1;
    __syncthreads();
  }
  const int tile = s_tile;
  if (tile < 0) {
    if (threadIdx.x == 0) {
      fr->r_remnant = radius_of(P[perm[n - 1]], fr);
      fr->m_remnant = fr->m_total;
    }
    return;
  }
  // Inside the tile: the same prefix over each thread's 8 sorted masses, then each thread walks its 8.
  const int64_t base = (int64_t)tile * MB_TILE + threadIdx.x * 8;
  dd own{0.0, 0.0};
  for (int j = 0; j < 8; ++j)
    if (base + j < n) own = dd_add1(own, (double)P[perm[base + j]].mass);
  const dd ex = block_inclusive(own);
  dd c = dd_add(dd{s_base_hi, s_base_lo}, ex);
  int64_t found = INT64_MAX;
  double mfound = 0.0;
  for (int j = 0; j < 8 && base + j < n; ++j) {
    c = dd_add1(c, (double)P[perm[base + j]].mass);
    if (c.hi >= thr) {
      found = base + j;
      mfound = c.hi;
      break;
    }
  }
  if (found != INT64_MAX) atomicMin((unsigned long long*)&s_best, (unsigned long long)found);
  __syncthreads();
  if (found != INT64_MAX && found == s_best) {
    fr->r_remnant = radius_of(P[perm[found]], fr);
    fr->m_remnant = mfound;
  }
  if (threadIdx.x == MB_THREADS - 1 && s_best == INT64_MAX) {
    // The in-tile association fell an ulp short of the tile sums': take the tile's last particle (this thread
    // walked the tile's last elements, so c is the running mass there).
    const int64_t last = min(n, (int64_t)(tile + 1) * MB_TILE) - 1;
    fr->r_remnant = radius_of(P[perm[last]], fr);
    fr->m_remnant = c.hi;
  }
}

// impact-run main.rs:232-247: None when unbound.
__device__ __forceinline__ bool perigee_ir(const double p[3], const double v[3], double mu, double* pg) {
  const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (r == 0.0) {
    *pg = 0.0;
    return true;
  }
  const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double energy = 0.5 * v2 - mu / r;
  if (energy >= 0.0) return false;
  const double a = -mu / (2.0 * energy);
  const double h0 = p[1] * v[2] - p[2] * v[1];
  const double h1 = p[2] * v[0] - p[0] * v[2];
  const double h2v = p[0] * v[1] - p[1] * v[0];
  const double h2 = h0 * h0 + h1 * h1 + h2v * h2v;
  const double e = sqrt(fmax(1.0 + 2.0 * energy * h2 / (mu * mu), 0.0));
  *pg = a * (1.0 - e);
  return true;
}

// orbit.rs:91-94.
__device__ __forceinline__ bool is_bound_orbit(const double p[3], const double v[3], double mu) {
  const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  return r > 0.0 && 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - mu / r < 0.0;
}

// orbit.rs:96-116: hyperbolic passes have a periapsis too.
__device__ __forceinline__ bool perigee_orbit(const double p[3], const double v[3], double mu, double* pg) {
  const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (r == 0.0) {
    *pg = 0.0;
    return true;
  }
  const double energy = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - mu / r;
  const double c0 = p[1] * v[2] - v[1] * p[2];
  const double c1 = p[2] * v[0] - v[2] * p[0];
  const double c2 = p[0] * v[1] - v[0] * p[1];
  const double h = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  const double e = sqrt(fmax(1.0 + 2.0 * energy * h * h / (mu * mu), 0.0));
  if (energy >= 0.0) {
    if (e <= 1.0 || h == 0.0) return false;
    *pg = h * h / mu / (1.0 + e);
    return true;
  }
  const double a = -mu / (2.0 * energy);
  *pg = a * (1.0 - e);
  return true;
}

// Orbit classes.
enum { CLS_REM = 0, CLS_DISK = 1, CLS_ESC = 2 };

template <int MODE>
__device__ __forceinline__ int classify(const SphParticle48& p, int64_t i, int64_t n_earth, const MeasureFrame* fr,
                                        bool* earth) {
  double rp[3], rv[3];
  for (int q = 0; q < 3; ++q) {
    rp[q] = (double)p.pos[q] - fr->com[q];
    rv[q] = (double)p.vel[q] - fr->v_com[q];
  }
  const double mu = kG64 * fr->m_remnant;
  const double r_rem = fr->r_remnant;
  double pg = 0.0;
  if (MODE == MM_IMPACT_RUN) {
    *earth = i < n_earth;  // main.rs:755
    if (!perigee_ir(rp, rv, mu, &pg)) return CLS_ESC;
    return pg > r_rem ? CLS_DISK : CLS_REM;
  }
  *earth = p.prov == 0u;  // gpu_sph.rs:434
  const bool bound = is_bound_orbit(rp, rv, mu);
  if (MODE == MM_ENGINE_BOUND && !bound) return CLS_ESC;  // gpu_sph.rs:429-432
  if (perigee_orbit(rp, rv, mu, &pg) && pg > r_rem) return CLS_DISK;
  return bound ? CLS_REM : CLS_ESC;
}

// Contiguous chunks: per-block category sums, disk flags.
template <int MODE>
__global__ __launch_bounds__(256) void k_classify(const SphParticle48* __restrict__ P, int64_t n, int64_t n_earth,
                                                  int64_t chunk, const MeasureFrame* __restrict__ fr,
                                                  uint32_t* __restrict__ flag, dd* __restrict__ part) {
  dd a[CAT_COUNT];
#pragma unroll
  for (int s = 0; s < CAT_COUNT; ++s) a[s] = dd{0.0, 0.0};
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  for (int64_t i = b0 + threadIdx.x; i < b1; i += MB_THREADS) {
    const SphParticle48 p = P[i];
    bool earth = false;
    const int c = classify<MODE>(p, i, n_earth, fr, &earth);
    const double m = (double)p.mass;
    const int slot = (c == CLS_DISK ? CAT_DISK_E : c == CLS_REM ? CAT_REM_E : CAT_ESC_E) + (earth ? 0 : 1);
#pragma unroll
    for (int s = 0; s < CAT_ESC_ALL; ++s)
      if (s == slot) a[s] = dd_add1(a[s], m);
    if (c == CLS_ESC) a[CAT_ESC_ALL] = dd_add1(a[CAT_ESC_ALL], m);
    if (c == CLS_DISK) a[CAT_DISK_H] = dd_add1(a[CAT_DISK_H], (double)p.h);
    flag[i] = c == CLS_DISK ? 1u : 0u;
  }
  const dd t = block_sum_dd<CAT_COUNT>(a);
  if (threadIdx.x < CAT_COUNT) part[(int64_t)blockIdx.x * CAT_COUNT + threadIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_class_final(const dd* __restrict__ part, int nblk, MeasureFrame* fr) {
  dd t[CAT_COUNT];
  sum_partials<CAT_COUNT>(part, nblk, t);
  if (threadIdx.x == 0)
    for (int s = 0; s < CAT_COUNT; ++s) fr->cat[s] = t[s].hi;
}

// Disk particles into the body-set arrays at their compacted index (index order = the reference's disk_idx /
// dp order). impact-run keeps FP32 positions for its FP32 differences and the raw density; the engine's disk
// floors the density at 1 (gpu_sph.rs:444, rho.max(1.0)).
template <int MODE>
__global__ __launch_bounds__(256) void k_disk_gather(const SphParticle48* __restrict__ P, int64_t n, int64_t n_earth,
                                                     const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                                     MeasureDev B) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const SphParticle48 p = P[i];
  const uint32_t j = pos[i];
  B.ax[j] = (double)p.pos[0];
  B.ay[j] = (double)p.pos[1];
  B.az[j] = (double)p.pos[2];
  B.axf[j] = p.pos[0];
  B.ayf[j] = p.pos[1];
  B.azf[j] = p.pos[2];
  B.avx[j] = (double)p.vel[0];
  B.avy[j] = (double)p.vel[1];
  B.avz[j] = (double)p.vel[2];
  B.am[j] = (double)p.mass;
  B.arho[j] = MODE == MM_IMPACT_RUN ? (double)p.rho : (double)fmaxf(p.rho, 1.0f);
  B.aearth[j] = MODE == MM_IMPACT_RUN ? (i < n_earth ? 1u : 0u) : (p.prov == 0u ? 1u : 0u);
  B.asrc[j] = (uint32_t)i;
}

// ---------------------------------------------------------------------------------------------
// Phase 2: friends-of-friends.

__global__ void k_bbox64_init(uint64_t* bb) {
  if (threadIdx.x < 6) bb[threadIdx.x] = threadIdx.x < 3 ? ~0ull : 0ull;
}

__global__ __launch_bounds__(256) void k_bbox64(const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, int64_t k, uint64_t* __restrict__ bb) {
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
  for (int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x; i < k; i += (int64_t)gridDim.x * MB_THREADS) {
    const uint64_t c[3] = {d2key(x[i]), d2key(y[i]), d2key(z[i])};
    for (int q = 0; q < 3; ++q) {
      lo[q] = min(lo[q], c[q]);
      hi[q] = max(hi[q], c[q]);
    }
  }
  __shared__ uint64_t s_b[4][6];
  const int w = threadIdx.x >> 6;
  for (int q = 0; q < 3; ++q) {
    lo[q] = wave_min(lo[q]);
    hi[q] = wave_max(hi[q]);
  }
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 3; ++q) {
      s_b[w][q] = lo[q];
      s_b[w][3 + q] = hi[q];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    uint64_t v = s_b[0][q];
    for (int ww = 1; ww < 4; ++ww) v = q < 3 ? min(v, s_b[ww][q]) : max(v, s_b[ww][q]);
    if (q < 3) atomicMin((unsigned long long*)&bb[q], (unsigned long long)v);
    else atomicMax((unsigned long long*)&bb[q], (unsigned long long)v);
  }
}

struct Grid {
  double lo[3];
  double cs;
  uint32_t nc[3];
  uint32_t bits[3];
  __device__ __forceinline__ uint64_t key(uint32_t cx, uint32_t cy, uint32_t cz) const {
    return (uint64_t)cx | ((uint64_t)cy << bits[0]) | ((uint64_t)cz << (bits[0] + bits[1]));
  }
  __device__ __forceinline__ uint32_t cell(double v, int q) const {
    const double c = floor((v - lo[q]) / cs);
    const double cmax = (double)(nc[q] - 1);
    return (uint32_t)fmin(fmax(c, 0.0), cmax);
  }
};

__global__ __launch_bounds__(256) void k_fof_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, int64_t k, Grid g,
                                                  uint64_t* __restrict__ key64, uint32_t* __restrict__ klo,
                                                  uint32_t* __restrict__ khi, uint32_t* __restrict__ vals,
                                                  uint32_t* __restrict__ parent) {
  const int64_t j = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (j >= k) return;
  const uint64_t key = g.key(g.cell(x[j], 0), g.cell(y[j], 1), g.cell(z[j], 2));
  key64[j] = key;
  klo[j] = (uint32_t)key;
  khi[j] = (uint32_t)(key >> 32);
  vals[j] = (uint32_t)j;
  parent[j] = (uint32_t)j;
}

__global__ __launch_bounds__(256) void k_gather_u64(const uint32_t* __restrict__ perm, const uint64_t* __restrict__ src,
                                                    int64_t n, uint64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// Union-find with parents that only ever decrease (p[x] <= x): a root is hooked under a smaller root, so the
// final root of a component is its lowest index. Parents are read at device scope (bypassing the CU cache);
// a stale parent is still an ancestor, so path halving by atomicMin is always valid.
__device__ __forceinline__ uint32_t uf_load(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t uf_find(uint32_t* par, uint32_t x) {
  while (true) {
    const uint32_t px = uf_load(par + x);
    if (px == x) return x;
    const uint32_t gx = uf_load(par + px);
    if (gx == px) return px;
    atomicMin(par + x, gx);
    x = gx;
  }
}
__device__ __forceinline__ void uf_unite(uint32_t* par, uint32_t a, uint32_t b) {
  while (true) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(par + a, a, b);
    if (old == a) return;
    a = old;  // a was hooked meanwhile; continue from its new parent
  }
}

// First sorted position u with skey[u] >= key.
__device__ __forceinline__ int64_t lower_bound64(const uint64_t* __restrict__ s, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (s[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Positions in cell-sorted order (the clump-order arrays are free until the clumps are formed), so that lanes
// scanning the same cells read the same addresses.
__global__ __launch_bounds__(256) void k_sorted_pos(int64_t k, const uint32_t* __restrict__ perm, MeasureDev B) {
  const int64_t t = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (t >= k) return;
  const uint32_t j = perm[t];
  B.bx[t] = B.ax[j];
  B.by[t] = B.ay[j];
  B.bz[t] = B.az[j];
  B.bxf[t] = B.axf[j];
  B.byf[t] = B.ayf[j];
  B.bzf[t] = B.azf[j];
}

// One wave per body (sorted position t): calls f(u) on the lanes for every body at sorted position u in the 3x3x3
// cells around it whose link test passes, d2 <= link^2 (neighbors.rs:69-84 / main.rs:826-835). Lanes 0-8 find
// the nine rows of three cells (one key range each, x fastest) and the wave strides over each row, so a body
// in a dense clump spreads its thousands of candidates over 64 lanes. FP32D: impact-run's FP32 position
// differences widened to f64 (main.rs:829); otherwise f64 differences (accretion.rs:107). The test is symmetric
// bit for bit (a - b = -(b - a)), so both bodies of a pair see the same answer.
template <bool FP32D, class F>
__device__ __forceinline__ void for_links_wave(int64_t k, const uint64_t* __restrict__ skey, int64_t t,
                                               const Grid& g, const MeasureDev& B, double ll2, F&& f) {
  const uint32_t lane = lane_id();
  const uint64_t key = skey[t];
  const uint32_t mx = (1u << g.bits[0]) - 1u, my = (1u << g.bits[1]) - 1u;
  const uint32_t cx = (uint32_t)(key & mx), cy = (uint32_t)((key >> g.bits[0]) & my),
                 cz = (uint32_t)(key >> (g.bits[0] + g.bits[1]));
  int64_t rs = 0, rn = 0;
  if (lane < 9) {
    const int64_t zz = (int64_t)cz + (int64_t)(lane / 3) - 1, yy = (int64_t)cy + (int64_t)(lane % 3) - 1;
    if (zz >= 0 && zz < g.nc[2] && yy >= 0 && yy < g.nc[1]) {
      const uint32_t x0 = cx > 0 ? cx - 1 : 0, x1 = min(cx + 1, g.nc[0] - 1);
      rs = lower_bound64(skey, k, g.key(x0, (uint32_t)yy, (uint32_t)zz));
      rn = lower_bound64(skey, k, g.key(x1, (uint32_t)yy, (uint32_t)zz) + 1) - rs;
    }
  }
  const double xj = B.bx[t], yj = B.by[t], zj = B.bz[t];
  const float xfj = B.bxf[t], yfj = B.byf[t], zfj = B.bzf[t];
  for (int r = 0; r < 9; ++r) {
    const int64_t s0 = __shfl(rs, r, 64), n0 = __shfl(rn, r, 64);
    for (int64_t i = lane; i < n0; i += 64) {
      const int64_t u = s0 + i;
      double d2;
      if (FP32D) {
        const double dx = (double)(xfj - B.bxf[u]), dy2 = (double)(yfj - B.byf[u]), dz2 = (double)(zfj - B.bzf[u]);
        d2 = dx * dx + dy2 * dy2 + dz2 * dz2;
      } else {
        const double dx = xj - B.bx[u], dy2 = yj - B.by[u], dz2 = zj - B.bz[u];
        d2 = dx * dx + dy2 * dy2 + dz2 * dz2;
      }
      if (d2 <= ll2) f(u);
    }
  }
}

// Pass 1 (plain stores): each body's parent is its lowest-index linked body (itself if none is lower). The
// parents form a forest inside the components (par[j] <= j).
template <bool FP32D>
__global__ __launch_bounds__(256) void k_fof_minlink(int64_t k, const uint64_t* __restrict__ skey,
                                                     const uint32_t* __restrict__ perm, Grid g, MeasureDev B,
                                                     double ll2, uint32_t* __restrict__ par) {
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= k) return;
  const uint32_t j = perm[t];
  uint32_t m = j;
  for_links_wave<FP32D>(k, skey, t, g, B, ll2, [&](int64_t u) { m = min(m, perm[u]); });
  m = wave_min(m);
  if (lane_id() == 0) par[j] = m;
}

// Pass 2: the forest root of every body (a plain array the next pass reads through the cache).
__global__ __launch_bounds__(256) void k_fof_froot(int64_t k, uint32_t* __restrict__ par, uint32_t* __restrict__ fr) {
  const int64_t j = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (j < k) fr[j] = uf_find(par, (uint32_t)j);
}

// Pass 3: links between different forest trees are united (lock-free union-find on the same parents). Inside
// a dense clump almost every link joins bodies of one tree and ends at a cached load.
template <bool FP32D>
__global__ __launch_bounds__(256) void k_fof_link(int64_t k, const uint64_t* __restrict__ skey,
                                                  const uint32_t* __restrict__ perm, Grid g, MeasureDev B, double ll2,
                                                  const uint32_t* __restrict__ fr, uint32_t* __restrict__ par) {
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= k) return;
  const uint32_t j = perm[t];
  const uint32_t rj = fr[j];
  for_links_wave<FP32D>(k, skey, t, g, B, ll2, [&](int64_t u) {
    const uint32_t l = perm[u];
    if (l > j && fr[l] != rj) uf_unite(par, rj, fr[l]);
  });
}

__global__ __launch_bounds__(256) void k_fof_roots(int64_t k, uint32_t* __restrict__ par, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (j >= k) return;
  keys[j] = uf_find(par, (uint32_t)j);
  vals[j] = (uint32_t)j;
}

__global__ __launch_bounds__(256) void k_heads(int64_t k, const uint32_t* __restrict__ sl, uint32_t* __restrict__ head) {
  const int64_t t = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (t < k) head[t] = (t == 0 || sl[t] != sl[t - 1]) ? 1u : 0u;
}

// Clump starts, each body's clump index (in input order), and the clump-ordered body copy.
__global__ __launch_bounds__(256) void k_clump_index(int64_t k, const uint32_t* __restrict__ head,
                                                     const uint32_t* __restrict__ ex, const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ nclump, MeasureDev B) {
  const int64_t t = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (t >= k) return;
  const uint32_t c = ex[t] + head[t] - 1u;
  if (head[t]) B.cstart[c] = (uint32_t)t;
  if (t == 0) B.cstart[*nclump] = (uint32_t)k;
  const uint32_t j = perm[t];
  B.label[j] = c;
  B.bx[t] = B.ax[j];
  B.by[t] = B.ay[j];
  B.bz[t] = B.az[j];
  B.bxf[t] = B.axf[j];
  B.byf[t] = B.ayf[j];
  B.bzf[t] = B.azf[j];
  B.bvx[t] = B.avx[j];
  B.bvy[t] = B.avy[j];
  B.bvz[t] = B.avz[j];
  B.bm[t] = B.am[j];
  B.brho[t] = B.arho[j];
  B.bearth[t] = B.aearth[j];
  B.bsrc[t] = B.asrc[j];
}

// Pair-potential work items per clump of >= 2 members: the T (T + 1) / 2 tiles (row tile <= column tile) of
// 256 x 256 members, T = ceil(n / 256).
__global__ __launch_bounds__(256) void k_clump_items(const uint32_t* __restrict__ cstart,
                                                     const uint32_t* __restrict__ nclump, uint32_t* __restrict__ items) {
  const int64_t c = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (c >= *nclump) return;
  const uint32_t n = cstart[c + 1] - cstart[c];
  const uint32_t T = (n + PE_ROWS - 1) / PE_ROWS;
  items[c] = n >= 2 ? T * (T + 1) / 2 : 0u;
}

__global__ void k_items_end(const uint32_t* __restrict__ nclump, const uint32_t* __restrict__ total,
                            uint32_t* __restrict__ ioff) {
  if (threadIdx.x == 0) ioff[*nclump] = *total;
}

// One wave per clump: mass, centre, velocity, volume, target mass, then the internal kinetic energy about the
// clump's velocity (accretion.rs:126-139; main.rs:843-852). Members are in index order.
template <bool IR>
__global__ __launch_bounds__(256) void k_clump_sums(int64_t C, MeasureDev B) {
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= C) return;
  const uint32_t s0 = B.cstart[c], s1 = B.cstart[c + 1];
  dd a[9];
#pragma unroll
  for (int s = 0; s < 9; ++s) a[s] = dd{0.0, 0.0};
  for (uint32_t t = s0 + lane; t < s1; t += 64) {
    const double m = B.bm[t];
    a[0] = dd_add1(a[0], m);
    a[1] = dd_add1(a[1], B.bx[t] * m);
    a[2] = dd_add1(a[2], B.by[t] * m);
    a[3] = dd_add1(a[3], B.bz[t] * m);
    a[4] = dd_add1(a[4], B.bvx[t] * m);
    a[5] = dd_add1(a[5], B.bvy[t] * m);
    a[6] = dd_add1(a[6], B.bvz[t] * m);
    a[7] = dd_add1(a[7], m / B.brho[t]);
    if (B.bearth[t]) a[8] = dd_add1(a[8], m);
  }
#pragma unroll
  for (int s = 0; s < 9; ++s) a[s] = wave_sum_dd(a[s]);
  const double mass = a[0].hi;
  const double cv[3] = {a[4].hi / mass, a[5].hi / mass, a[6].hi / mass};
  dd ke{0.0, 0.0};
  for (uint32_t t = s0 + lane; t < s1; t += 64) {
    const double dx = B.bvx[t] - cv[0], dy = B.bvy[t] - cv[1], dz = B.bvz[t] - cv[2];
    ke = dd_add1(ke, 0.5 * B.bm[t] * (dx * dx + dy * dy + dz * dz));
  }
  ke = wave_sum_dd(ke);
  if (lane == 0) {
    ihs_clump r;
    r.first = (int64_t)B.bsrc[s0];
    r.n_members = (int64_t)(s1 - s0);
    r.mass = mass;
    r.com_pos[0] = a[1].hi / mass;
    r.com_pos[1] = a[2].hi / mass;
    r.com_pos[2] = a[3].hi / mass;
    r.com_vel[0] = cv[0];
    r.com_vel[1] = cv[1];
    r.com_vel[2] = cv[2];
    const double vol = a[7].hi;
    r.rho = vol > 0.0 ? mass / vol : (IR ? 2700.0 : B.brho[s0]);
    r.radius = cbrt(mass / (kFourThirdsPi * r.rho));
    r.internal_ke = ke.hi;
    r.self_pe = 0.0;
    r.earth_mass = a[8].hi;
    r.bound = 0;
    r.outside_roche = 0;
    B.crec[c] = r;
  }
}

// Softened (accretion.rs:141-148) or floored (main.rs:853-860) pair potential over the rows [rt*256, rt*256+256)
// of one clump against every later member, member tiles staged in LDS. Writes the item's double-double sum of
// G m_a m_b / r.
template <bool IR>
__global__ __launch_bounds__(256) void k_clump_pe(int64_t C, int64_t nitems, MeasureDev B, double g, double soft2) {
  __shared__ double sx[PE_ROWS], sy[PE_ROWS], sz[PE_ROWS], sm[PE_ROWS];
  __shared__ float sxf[PE_ROWS], syf[PE_ROWS], szf[PE_ROWS];
  const int64_t item = blockIdx.x;
  // Clump of this item: the last c with ioff[c] <= item.
  int64_t lo = 0, hi = C;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (B.ioff[mid] <= (uint32_t)item) lo = mid;
    else hi = mid;
  }
  const int64_t c = lo;
  const uint32_t s0 = B.cstart[c];
  const int64_t nc = (int64_t)(B.cstart[c + 1] - s0);
  // (row tile, column tile) of this item, rows first: row tile r holds T - r items.
  const int64_t T = (nc + PE_ROWS - 1) / PE_ROWS;
  int64_t idx = item - B.ioff[c], rt = 0;
  while (idx >= T - rt) {
    idx -= T - rt;
    ++rt;
  }
  const int64_t ct = rt + idx;
  const int64_t a = rt * PE_ROWS + threadIdx.x;
  double xa = 0.0, ya = 0.0, za = 0.0, ma = 0.0;
  float xfa = 0.f, yfa = 0.f, zfa = 0.f;
  if (a < nc) {
    xa = B.bx[s0 + a];
    ya = B.by[s0 + a];
    za = B.bz[s0 + a];
    ma = B.bm[s0 + a];
    xfa = B.bxf[s0 + a];
    yfa = B.byf[s0 + a];
    zfa = B.bzf[s0 + a];
  }
  dd acc{0.0, 0.0};
  {
    const int64_t b = ct * PE_ROWS + threadIdx.x;
    if (b < nc) {
      sx[threadIdx.x] = B.bx[s0 + b];
      sy[threadIdx.x] = B.by[s0 + b];
      sz[threadIdx.x] = B.bz[s0 + b];
      sm[threadIdx.x] = B.bm[s0 + b];
      sxf[threadIdx.x] = B.bxf[s0 + b];
      syf[threadIdx.x] = B.byf[s0 + b];
      szf[threadIdx.x] = B.bzf[s0 + b];
    }
    __syncthreads();
    const int lim = (int)min((int64_t)PE_ROWS, nc - ct * PE_ROWS);
    if (a < nc) {
      const int j0 = ct == rt ? (int)threadIdx.x + 1 : 0;
      for (int jj = j0; jj < lim; ++jj) {
        double r;
        if (IR) {
          const double dx = (double)(xfa - sxf[jj]), dy = (double)(yfa - syf[jj]), dz = (double)(zfa - szf[jj]);
          r = fmax(sqrt(dx * dx + dy * dy + dz * dz), 1.0);
        } else {
          const double dx = xa - sx[jj], dy = ya - sy[jj], dz = za - sz[jj];
          r = sqrt(dx * dx + dy * dy + dz * dz + soft2);
        }
        acc = dd_add1(acc, g * ma * sm[jj] / r);
      }
    }
  }
  dd v[1] = {acc};
  const dd t = block_sum_dd<1>(v);
  if (threadIdx.x == 0) ((dd*)B.pe_part)[item] = t;
}

// Per clump: the potential from its items, boundness and the Roche gate (accretion.rs:149-152, 158-160;
// main.rs:838-866). rho_c = central mass / (4/3 pi R^3).
__global__ __launch_bounds__(256) void k_clump_final(int64_t C, MeasureDev B, double rho_c, double cr, double cx,
                                                     double cy, double cz) {
  const int64_t c = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (c >= C) return;
  ihs_clump r = B.crec[c];
  dd s{0.0, 0.0};
  for (uint32_t it = B.ioff[c]; it < B.ioff[c + 1]; ++it) s = dd_add(s, ((const dd*)B.pe_part)[it]);
  r.self_pe = 0.0 - s.hi;
  r.bound = r.internal_ke + r.self_pe < 0.0 ? 1 : 0;
  const double d_roche = 2.44 * cr * cbrt(rho_c / r.rho);
  const double dx = r.com_pos[0] - cx, dy = r.com_pos[1] - cy, dz = r.com_pos[2] - cz;
  r.outside_roche = sqrt(dx * dx + dy * dy + dz * dz) > d_roche ? 1 : 0;
  B.crec[c] = r;
}

// Counts over clumps of >= 2 members and the most massive bound clump outside Roche (main.rs:838-877,
// Clump::accretes), with integer atomics so the result does not depend on scheduling: the largest mass first
// (positive doubles order as their bit patterns), then the lowest clump index among clumps of that mass.
__global__ void k_best_init(ClumpSummary* out) {
  if (threadIdx.x == 0) {
    out->n2 = 0;
    out->n_bound = 0;
    out->best = 0x7fffffff;
    out->pad = 0;
    out->best_mass = 0ull;
  }
}

__device__ __forceinline__ bool accretes(const ihs_clump& r) {
  return r.n_members >= 2 && r.bound && r.outside_roche && r.mass > 0.0;
}

__global__ __launch_bounds__(256) void k_best_mass(int64_t C, const ihs_clump* __restrict__ rec,
                                                   ClumpSummary* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  uint32_t n2 = 0, nb = 0;
  unsigned long long mb = 0ull;
  if (c < C) {
    const ihs_clump& r = rec[c];
    n2 = r.n_members >= 2 ? 1u : 0u;
    nb = (r.n_members >= 2 && r.bound) ? 1u : 0u;
    if (accretes(r)) mb = (unsigned long long)__double_as_longlong(r.mass);
  }
  n2 = wave_sum(n2);
  nb = wave_sum(nb);
  mb = wave_max(mb);
  if ((threadIdx.x & 63) == 0) {
    if (n2) atomicAdd(&out->n2, n2);
    if (nb) atomicAdd(&out->n_bound, nb);
    if (mb) atomicMax(&out->best_mass, mb);
  }
}

__global__ __launch_bounds__(256) void k_best_index(int64_t C, const ihs_clump* __restrict__ rec,
                                                    ClumpSummary* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (c >= C) return;
  const ihs_clump& r = rec[c];
  if (out->best_mass != 0ull && accretes(r) && (unsigned long long)__double_as_longlong(r.mass) == out->best_mass)
    atomicMin(&out->best, (int32_t)c);
}

// body_bulk filters: the records of one provenance, then those within `rmax` of a centre.
__global__ __launch_bounds__(256) void k_flag_prov(const SphParticle48* __restrict__ P, int64_t n, uint32_t prov,
                                                   uint32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i < n) flag[i] = P[i].prov == prov ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_flag_within(const SphParticle48* __restrict__ P, int64_t n, double cx,
                                                     double cy, double cz, double rmax, uint32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i >= n) return;
  const double dx = (double)P[i].pos[0] - cx, dy = (double)P[i].pos[1] - cy, dz = (double)P[i].pos[2] - cz;
  flag[i] = sqrt(dx * dx + dy * dy + dz * dz) <= rmax ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_compact_rec(const SphParticle48* __restrict__ P, int64_t n,
                                                     const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                                     SphParticle48* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * MB_THREADS + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = P[i];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side.

void measure_release(MeasureBufs& B) {
  for (void* p : B.allocs) (void)hipFree(p);
  if (B.pe_part) (void)hipFree(B.pe_part);
  B = MeasureBufs{};
}

hipError_t measure_reserve(MeasureBufs& B, int64_t cap) {
  if (cap <= B.cap) return hipSuccess;
  measure_release(B);
  const int64_t c = std::max<int64_t>(cap, 1);
  hipError_t err = hipSuccess;
  auto get = [&](auto** p, size_t count) {
    if (err != hipSuccess) return;
    void* q = nullptr;
    err = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(**p));
    if (err == hipSuccess) {
      B.allocs.push_back(q);
      *p = (std::remove_reference_t<decltype(*p)>)q;
    }
  };
  for (int i = 0; i < 2; ++i) {
    get(&B.rk[i], c);
    get(&B.rv[i], c);
  }
  get(&B.khi, c);
  get(&B.hist, (size_t)256 * radix_tiles(c));
  get(&B.dtot, 256);
  get(&B.key64, c);
  get(&B.skey, c);
  dd* part = nullptr;
  get(&part, (size_t)MB_MAXBLK * CAT_COUNT);
  B.part = part;
  dd* tsum = nullptr;
  get(&tsum, (size_t)(c + MB_TILE - 1) / MB_TILE);
  B.tsum = tsum;
  get(&B.flag, c);
  get(&B.scan, c);
  get(&B.bsum, MB_MAXBLK + 1);
  get(&B.frame, 1);
  get(&B.summary, 1);
  get(&B.bbox, 6);
  for (double** p : {&B.ax, &B.ay, &B.az, &B.avx, &B.avy, &B.avz, &B.am, &B.arho, &B.bx, &B.by, &B.bz, &B.bvx,
                     &B.bvy, &B.bvz, &B.bm, &B.brho})
    get(p, c);
  for (float** p : {&B.axf, &B.ayf, &B.azf, &B.bxf, &B.byf, &B.bzf}) get(p, c);
  for (uint32_t** p : {&B.aearth, &B.asrc, &B.bearth, &B.bsrc, &B.parent, &B.label}) get(p, c);
  get(&B.cstart, c + 1);
  get(&B.ioff, c + 1);
  get(&B.crec, c);
  get(&B.rec[0], c);
  get(&B.rec[1], c);
  if (err != hipSuccess) {
    measure_release(B);
    return err;
  }
  B.cap = cap;
  return hipSuccess;
}

hipError_t measure_frame(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, int mode, int64_t n_earth,
                         MeasureFrame* out) {
  dd* part = (dd*)B.part;
  // Totals.
  const int nsum = (int)std::min<int64_t>(MB_MAXBLK, std::max<int64_t>(1, (n + 2047) / 2048));
  k_frame_sums<<<nsum, MB_THREADS, 0, s>>>(P, n, part);
  k_frame_final<<<1, MB_THREADS, 0, s>>>(part, nsum, B.frame);
  // Radii sorted by value, stable in index (low word, then high word of the f64 bit pattern).
  k_radius_keys<<<g256(n), MB_THREADS, 0, s>>>(P, n, B.frame, B.rk[0], B.khi, B.rv[0]);
  int cur = launch_radix_pairs(s, n, 32, B.rk, B.rv, 0, B.hist, B.dtot);
  k_gather_u32<<<g256(n), MB_THREADS, 0, s>>>(B.rv[cur], B.khi, n, B.rk[cur]);
  cur = launch_radix_pairs(s, n, 32, B.rk, B.rv, cur, B.hist, B.dtot);
  const uint32_t* perm = B.rv[cur];
  const int ntile = (int)((n + MB_TILE - 1) / MB_TILE);
  k_mass_tiles<<<ntile, MB_THREADS, 0, s>>>(P, perm, n, (dd*)B.tsum);
  k_remnant<<<1, MB_THREADS, 0, s>>>(P, perm, n, (const dd*)B.tsum, ntile, 0.85, B.frame);
  // Classification and the disk set.
  const Chunks ch = chunks_for(n);
  switch (mode) {
    case MM_IMPACT_RUN:
      k_classify<MM_IMPACT_RUN><<<ch.nblk, MB_THREADS, 0, s>>>(P, n, n_earth, ch.chunk, B.frame, B.flag, part);
      break;
    case MM_ENGINE_BOUND:
      k_classify<MM_ENGINE_BOUND><<<ch.nblk, MB_THREADS, 0, s>>>(P, n, n_earth, ch.chunk, B.frame, B.flag, part);
      break;
    default:
      k_classify<MM_ENGINE_ANY><<<ch.nblk, MB_THREADS, 0, s>>>(P, n, n_earth, ch.chunk, B.frame, B.flag, part);
      break;
  }
  k_class_final<<<1, MB_THREADS, 0, s>>>(part, ch.nblk, B.frame);
  launch_exscan(s, B.flag, n, B.scan, B.bsum, &B.frame->n_disk);
  if (mode == MM_IMPACT_RUN)
    k_disk_gather<MM_IMPACT_RUN><<<g256(n), MB_THREADS, 0, s>>>(P, n, n_earth, B.flag, B.scan, B);
  else
    k_disk_gather<MM_ENGINE_ANY><<<g256(n), MB_THREADS, 0, s>>>(P, n, n_earth, B.flag, B.scan, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(out, B.frame, sizeof *out, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

namespace {
int bits_for(uint64_t ncells) {
  int b = 0;
  while (b < 63 && (1ull << b) < ncells) ++b;
  return b;
}
int sort_bits_u32(uint64_t maxval) { return std::max(8, (bits_for(maxval + 1) + 7) / 8 * 8); }
}  // namespace

hipError_t clump_set(hipStream_t s, MeasureBufs& B, int64_t k, const ClumpQuery& q, std::vector<ihs_clump>* all,
                     std::vector<int32_t>* labels, ClumpSummary* sum, ihs_clump* best) {
  hipError_t e;
  // Work items of one clump (T (T + 1) / 2 tiles) are counted in 32 bits.
  if (k < 1 || k > (int64_t)1 << 24) return hipErrorInvalidValue;
  // Grid: cell side >= link (1 + 2^-20), so that rounding in (x - lo) / cs can never separate a linked pair by
  // two cells, and at most 2^20 cells per axis.
  k_bbox64_init<<<1, 64, 0, s>>>(B.bbox);
  k_bbox64<<<(unsigned)std::min<int64_t>(1024, g256(k)), MB_THREADS, 0, s>>>(B.ax, B.ay, B.az, k, B.bbox);
  uint64_t bbk[6];
  if ((e = hipMemcpyAsync(bbk, B.bbox, sizeof bbk, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  auto k2d = [](uint64_t kk) {
    const uint64_t u = (kk >> 63) ? (kk & 0x7fffffffffffffffull) : ~kk;
    double d;
    std::memcpy(&d, &u, sizeof d);
    return d;
  };
  Grid g;
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = k2d(bbk[a]);
    ext = std::max(ext, k2d(bbk[3 + a]) - g.lo[a]);
  }
  if (!std::isfinite(ext)) return hipErrorInvalidValue;
  double cs = q.link * (1.0 + std::ldexp(1.0, -20));
  cs = std::max(cs, ext * std::ldexp(1.0, -20));
  if (!(cs > 0.0)) cs = 1.0;
  g.cs = cs;
  int tb = 0;
  for (int a = 0; a < 3; ++a) {
    const double span = k2d(bbk[3 + a]) - g.lo[a];
    const uint64_t nc = (uint64_t)std::floor(span / cs) + 2;  // + 1 for the end cell, + 1 for rounding
    g.nc[a] = (uint32_t)nc;
    g.bits[a] = (uint32_t)bits_for(nc);
    tb += (int)g.bits[a];
  }
  const double ll2 = q.link * q.link;
  // Cell keys, sorted (stable in body index).
  k_fof_keys<<<g256(k), MB_THREADS, 0, s>>>(B.ax, B.ay, B.az, k, g, B.key64, B.rk[0], B.khi, B.rv[0], B.parent);
  int cur = launch_radix_pairs(s, k, std::min(tb, 32), B.rk, B.rv, 0, B.hist, B.dtot);
  if (tb > 32) {
    k_gather_u32<<<g256(k), MB_THREADS, 0, s>>>(B.rv[cur], B.khi, k, B.rk[cur]);
    cur = launch_radix_pairs(s, k, tb - 32, B.rk, B.rv, cur, B.hist, B.dtot);
  }
  k_gather_u64<<<g256(k), MB_THREADS, 0, s>>>(B.rv[cur], B.key64, k, B.skey);
  k_sorted_pos<<<g256(k), MB_THREADS, 0, s>>>(k, B.rv[cur], B);
  // Links and roots: one wave per body.
  const unsigned wgrid = (unsigned)((k + 3) / 4);
  if (q.impact_run) {
    k_fof_minlink<true><<<wgrid, MB_THREADS, 0, s>>>(k, B.skey, B.rv[cur], g, B, ll2, B.parent);
    k_fof_froot<<<g256(k), MB_THREADS, 0, s>>>(k, B.parent, B.label);
    k_fof_link<true><<<wgrid, MB_THREADS, 0, s>>>(k, B.skey, B.rv[cur], g, B, ll2, B.label, B.parent);
  } else {
    k_fof_minlink<false><<<wgrid, MB_THREADS, 0, s>>>(k, B.skey, B.rv[cur], g, B, ll2, B.parent);
    k_fof_froot<<<g256(k), MB_THREADS, 0, s>>>(k, B.parent, B.label);
    k_fof_link<false><<<wgrid, MB_THREADS, 0, s>>>(k, B.skey, B.rv[cur], g, B, ll2, B.label, B.parent);
  }
  k_fof_roots<<<g256(k), MB_THREADS, 0, s>>>(k, B.parent, B.rk[0], B.rv[0]);
  cur = launch_radix_pairs(s, k, sort_bits_u32((uint64_t)(k - 1)), B.rk, B.rv, 0, B.hist, B.dtot);
  // Clump segments.
  k_heads<<<g256(k), MB_THREADS, 0, s>>>(k, B.rk[cur], B.flag);
  launch_exscan(s, B.flag, k, B.scan, B.bsum, &B.frame->n_clumps);
  k_clump_index<<<g256(k), MB_THREADS, 0, s>>>(k, B.flag, B.scan, B.rv[cur], &B.frame->n_clumps, B);
  uint32_t C = 0;
  if ((e = hipMemcpyAsync(&C, &B.frame->n_clumps, sizeof C, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  // Pair-potential items.
  k_clump_items<<<g256(C), MB_THREADS, 0, s>>>(B.cstart, &B.frame->n_clumps, B.flag);
  launch_exscan(s, B.flag, C, B.ioff, B.bsum, &B.frame->n_items);
  k_items_end<<<1, 64, 0, s>>>(&B.frame->n_clumps, &B.frame->n_items, B.ioff);
  uint32_t T = 0;
  if ((e = hipMemcpyAsync(&T, &B.frame->n_items, sizeof T, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if ((int64_t)T > B.pe_cap) {
    if (B.pe_part) (void)hipFree(B.pe_part);
    B.pe_part = nullptr;
    B.pe_cap = 0;
    const int64_t want = std::max<int64_t>((int64_t)T, 2 * B.cap);
    if ((e = hipMalloc(&B.pe_part, (size_t)want * sizeof(dd))) != hipSuccess) return e;
    B.pe_cap = want;
  }
  if (q.impact_run) k_clump_sums<true><<<(unsigned)((C + 3) / 4), MB_THREADS, 0, s>>>(C, B);
  else k_clump_sums<false><<<(unsigned)((C + 3) / 4), MB_THREADS, 0, s>>>(C, B);
  if (T > 0) {
    if (q.impact_run) k_clump_pe<true><<<T, MB_THREADS, 0, s>>>(C, T, B, q.g, q.soft2);
    else k_clump_pe<false><<<T, MB_THREADS, 0, s>>>(C, T, B, q.g, q.soft2);
  }
  const double rho_c = q.central_mass / (kFourThirdsPi * (q.central_radius * q.central_radius * q.central_radius));
  k_clump_final<<<g256(C), MB_THREADS, 0, s>>>(C, B, rho_c, q.central_radius, q.central[0], q.central[1],
                                                q.central[2]);
  k_best_init<<<1, 64, 0, s>>>(B.summary);
  k_best_mass<<<g256(C), MB_THREADS, 0, s>>>(C, B.crec, B.summary);
  k_best_index<<<g256(C), MB_THREADS, 0, s>>>(C, B.crec, B.summary);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  ClumpSummary cs_h{};
  if ((e = hipMemcpyAsync(&cs_h, B.summary, sizeof cs_h, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if (all) {
    all->resize(C);
    if ((e = hipMemcpyAsync(all->data(), B.crec, C * sizeof(ihs_clump), hipMemcpyDeviceToHost, s)) != hipSuccess)
      return e;
  }
  std::vector<uint32_t> lab;
  if (labels) {
    lab.resize(k);
    if ((e = hipMemcpyAsync(lab.data(), B.label, k * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess)
      return e;
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (cs_h.best == 0x7fffffff) cs_h.best = -1;
  if (best) {
    *best = ihs_clump{};
    if (cs_h.best >= 0) {
      if ((e = hipMemcpy(best, B.crec + cs_h.best, sizeof *best, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    }
  }
  if (labels) labels->assign(lab.begin(), lab.end());
  if (sum) *sum = cs_h;
  return hipSuccess;
}

namespace {
// Centre and mass of P[0, n) and the radius about that centre enclosing `frac` of the mass, into the frame.
hipError_t centre_and_radius(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, double frac,
                             MeasureFrame* out) {
  dd* part = (dd*)B.part;
  const int nsum = (int)std::min<int64_t>(MB_MAXBLK, std::max<int64_t>(1, (n + 2047) / 2048));
  k_frame_sums<<<nsum, MB_THREADS, 0, s>>>(P, n, part);
  k_frame_final<<<1, MB_THREADS, 0, s>>>(part, nsum, B.frame);
  k_radius_keys<<<g256(n), MB_THREADS, 0, s>>>(P, n, B.frame, B.rk[0], B.khi, B.rv[0]);
  int cur = launch_radix_pairs(s, n, 32, B.rk, B.rv, 0, B.hist, B.dtot);
  k_gather_u32<<<g256(n), MB_THREADS, 0, s>>>(B.rv[cur], B.khi, n, B.rk[cur]);
  cur = launch_radix_pairs(s, n, 32, B.rk, B.rv, cur, B.hist, B.dtot);
  const int ntile = (int)((n + MB_TILE - 1) / MB_TILE);
  k_mass_tiles<<<ntile, MB_THREADS, 0, s>>>(P, B.rv[cur], n, (dd*)B.tsum);
  k_remnant<<<1, MB_THREADS, 0, s>>>(P, B.rv[cur], n, (const dd*)B.tsum, ntile, frac, B.frame);
  hipError_t e = hipMemcpyAsync(out, B.frame, sizeof *out, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

// The records of P[0, n) whose flag is set, in order, into dst; returns their count.
hipError_t compact_records(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, SphParticle48* dst,
                           uint32_t* count) {
  launch_exscan(s, B.flag, n, B.scan, B.bsum, &B.frame->n_sel);
  k_compact_rec<<<g256(n), MB_THREADS, 0, s>>>(P, n, B.flag, B.scan, dst);
  hipError_t e = hipMemcpyAsync(count, &B.frame->n_sel, sizeof *count, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}
}  // namespace

// gpu_sph.rs:279-316: centre of mass, then the radius enclosing 98% of the mass about it; twice, re-centre on the
// bodies within 1.5 radii of the centre (always filtering the whole body) and re-measure. The mass is the whole
// body's.
hipError_t body_bulk(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, uint32_t prov, double out[5]) {
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  k_flag_prov<<<g256(n), MB_THREADS, 0, s>>>(P, n, prov, B.flag);
  uint32_t m0 = 0;
  hipError_t e = compact_records(s, B, P, n, B.rec[0], &m0);
  if (e != hipSuccess || m0 == 0) return e;
  MeasureFrame fr;
  if ((e = centre_and_radius(s, B, B.rec[0], m0, 0.98, &fr)) != hipSuccess) return e;
  const double mass = fr.m_total;
  double c[3] = {fr.com[0], fr.com[1], fr.com[2]}, r = fr.r_remnant;
  for (int pass = 0; pass < 2; ++pass) {
    k_flag_within<<<g256(m0), MB_THREADS, 0, s>>>(B.rec[0], m0, c[0], c[1], c[2], 1.5 * r, B.flag);
    uint32_t kk = 0;
    if ((e = compact_records(s, B, B.rec[0], m0, B.rec[1], &kk)) != hipSuccess) return e;
    if (kk == 0) break;
    if ((e = centre_and_radius(s, B, B.rec[1], kk, 0.98, &fr)) != hipSuccess) return e;
    for (int q = 0; q < 3; ++q) c[q] = fr.com[q];
    r = fr.r_remnant;
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = mass;
  out[4] = r;
  return hipSuccess;
}

hipError_t clump_upload(hipStream_t s, MeasureBufs& B, int64_t n, const double* pos3, const double* vel3,
                        const double* mass, const double* rho) {
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
  std::vector<float> xf(n), yf(n), zf(n);
  std::vector<uint32_t> zero(n, 0u), src(n);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = pos3[3 * i];
    y[i] = pos3[3 * i + 1];
    z[i] = pos3[3 * i + 2];
    xf[i] = (float)x[i];
    yf[i] = (float)y[i];
    zf[i] = (float)z[i];
    vx[i] = vel3[3 * i];
    vy[i] = vel3[3 * i + 1];
    vz[i] = vel3[3 * i + 2];
    src[i] = (uint32_t)i;
  }
  hipError_t e = hipSuccess;
  auto put = [&](void* d, const void* h, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s);
  };
  put(B.ax, x.data(), n * 8);
  put(B.ay, y.data(), n * 8);
  put(B.az, z.data(), n * 8);
  put(B.axf, xf.data(), n * 4);
  put(B.ayf, yf.data(), n * 4);
  put(B.azf, zf.data(), n * 4);
  put(B.avx, vx.data(), n * 8);
  put(B.avy, vy.data(), n * 8);
  put(B.avz, vz.data(), n * 8);
  put(B.am, mass, n * 8);
  put(B.arho, rho, n * 8);
  put(B.aearth, zero.data(), n * 4);
  put(B.asrc, src.data(), n * 4);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}
