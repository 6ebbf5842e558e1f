// SPH passes over Morton-sorted particles (gfx950, wave64):
//   k_nbr_build  — per-group LBVH walk (box vs box + search radius), LDS-staged leaf tiles, per-particle
//                  neighbour lists in column-major HBM layout (lane-coalesced reads later);
//   k_hsolve1/2  — Newton smoothing-length solve to a density-consistency tolerance (new: the reference
//                  keeps h fixed, ie/crates/engine/src/hydrostatic.rs:58-61), the unconverged particles of
//                  the first pass compacted into full waves for the second;
//   k_density    — rho_i with the pairwise h_ij = (h_i+h_j)/2 (ie/shaders/sph_step.wgsl:118-149), then
//                  the Tillotson P and c per particle (sph_step.wgsl:78-100) and the Courant signal
//                  (sph_step.wgsl:262-269), fused;
//   k_forces     — pressure + Monaghan AV + du/dt (ie/shaders/sph_step.wgsl:170-208), gather form.
// The one-thread-per-particle kernels that gather neighbours map their blocks to particle ranges XCD by XCD
// (xcd_block): blocks go round-robin to the eight XCDs, each with its own L2, and a contiguous Morton range per
// XCD keeps a particle's neighbours in the L2 that reads them.
#include "common.h"
#include "launch.h"

namespace {

constexpr int FWIN_HALO = 128;                  // force pass: particles staged on each side of the block
constexpr int FWIN = 256 + 2 * FWIN_HALO;        // window size (blocks of 256)
constexpr int DWIN_HALO = 256;                  // density and first h-solve pass
constexpr int DWIN = 256 + 2 * DWIN_HALO;

// A block's Morton neighbourhood of positions staged in LDS: neighbours inside [lo, lo + n) are read there, the
// others from global memory (n == 0: no window). The compiler makes the two-way read one flat load on a selected
// address; lanes inside the window are served by LDS and issue no L2 request. (Forcing a separate LDS read and
// global load per branch was slower: the branches then serialize and fewer loads are in flight.)
struct XWin {
  const float4* w;
  const uint32_t* wl;  // REL: the window's remainder codes (split precision)
  uint32_t lo, n;
};

// A neighbour's working position and, in split precision (REL), its remainder code (pair_d in common.h).
struct PosLo {
  float4 q;
  uint32_t c;
};

template <bool REL>
__device__ __forceinline__ PosLo win_xm(const XWin& W, const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                        uint32_t j) {
  const uint32_t o = j - W.lo;
  const bool inwin = o < W.n;
  PosLo p;
  p.q = inwin ? W.w[o] : xm[j];
  p.c = REL ? (inwin ? W.wl[o] : xlo[j]) : 0u;
  return p;
}

// This particle's remainder (zero without split precision).
template <bool REL>
__device__ __forceinline__ float3 own_lo(const uint32_t* __restrict__ xlo, int64_t t, const float4 me) {
  return REL ? lo3(xlo[t], me) : make_float3(0.f, 0.f, 0.f);
}

// Window bounds of block t0 (blocks of 256) with `halo` particles on each side, clipped to [0, n).
__device__ __forceinline__ void win_bounds(int64_t t0, int64_t n, int halo, int64_t& lo, int64_t& hi) {
  lo = t0 > halo ? t0 - halo : 0;
  hi = min(t0 + 256 + (int64_t)halo, n);
}


constexpr int NB_WAVES = 4;
constexpr int NB_STACK = 512;
constexpr int NB_SMALL = 8;     // a subtree of at most this many leaves that passes the box test is taken whole
constexpr int NB_RING = 1024;   // candidate leaves: < 64 pending + up to 64 x NB_SMALL per walk step

struct NbrLds {
  uint32_t stack[NB_STACK];
  uint32_t leaf[NB_RING];
  float4 part[64];  // candidate x, y, z and its searched radius squared
  uint32_t pidx[64];
};

// Neighbour list of one group (one wave): LBVH walk with box-vs-box tests against the larger search radius,
// candidate leaves staged through LDS 64 particles at a time, each lane testing r < max(hs_i, hs_j).
// The walk stops at subtrees of at most NB_SMALL leaves: one that passes the box test hands all its leaves (a
// contiguous index range) to the candidate ring instead of being opened level by level. The lists then come out
// in runs of consecutive indices, which the neighbour loops of the later passes gather with better locality.
__device__ __forceinline__ void nbr_group(NbrLds& S, const int4 grp, uint32_t lane, const float4* __restrict__ xm,
                                          const float* __restrict__ hs, int64_t n, int K, int64_t L,
                                          const int2* __restrict__ child, const int2* __restrict__ range,
                                          const float4* __restrict__ nbmin,
                                          const float4* __restrict__ nbmax, uint4* __restrict__ nbr4,
                                          uint32_t* __restrict__ ncount, int cap, int64_t stride,
                                          uint32_t* __restrict__ flags, uint64_t& ncand, uint64_t& nnodes,
                                          uint64_t& npairs) {
  const int64_t t0 = grp.x;
  const int64_t t = t0 + lane;
  const bool valid = lane < (uint32_t)grp.y;
  const float4 me = xm[valid ? t : t0];
  const float myhs = valid ? hs[t] : 0.f;
  const uint32_t myhs2u = __float_as_uint(myhs * myhs);
  const float gx0 = wave_min(me.x), gy0 = wave_min(me.y), gz0 = wave_min(me.z);
  const float gx1 = wave_max(me.x), gy1 = wave_max(me.y), gz1 = wave_max(me.z);
  const float ghs = wave_max(myhs);
  const uint64_t nvalid = (uint64_t)__popcll(__ballot(valid));
  const uint32_t leaf0 = (uint32_t)(L - 1);
  const uint32_t lpc = 64u / (uint32_t)K;
  uint32_t sp = 1, lhead = 0, ltail = 0, cnt = 0;
  wave_sync();
  if (lane == 0) S.stack[0] = (uint32_t)grp.w;  // the group's segment root
  wave_sync();

  uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, nb = 0, nchunk = 0;
  const uint32_t cap4 = (uint32_t)cap >> 2;
  const uint32_t tt = (uint32_t)t;
  auto flush = [&](uint32_t nl) {
    const uint32_t li = lane / (uint32_t)K;
    const bool have = li < nl;
    const uint32_t leaf = have ? S.leaf[(lhead + li) & (NB_RING - 1)] : 0u;
    const int64_t src = (int64_t)leaf * K + (lane % (uint32_t)K);
    const bool ok = have && src < n;
    wave_sync();
    // Empty slots sit at 1e30 with a zero radius: they can never pass the distance test.
    const float ph = ok ? hs[src] : 0.f;
    const float4 pp = ok ? xm[src] : make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.f);
    S.part[lane] = make_float4(pp.x, pp.y, pp.z, ph * ph);
    S.pidx[lane] = ok ? (uint32_t)src : IHS_NONE;
    wave_sync();
    const uint32_t m = (min(64u, nl * (uint32_t)K) + 3u) & ~3u;
    ncand += nvalid * m;
    // Branch-free test of every candidate into a per-lane hit mask (r < hs_i or r < hs_j), then visit only the
    // hits: a divergent branch inside the candidate loop would serialise the LDS loads. The loop is unrolled
    // so that each candidate's mask bit is a constant; the particle itself is dropped among the hits.
    uint32_t hlo = 0, hhi = 0;
#pragma unroll
    for (uint32_t k0 = 0; k0 < 64; k0 += 4) {
      if (k0 < m) {
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
          const uint32_t k = k0 + e;
          const float4 p = S.part[k];
          const float dx = p.x - me.x, dy = p.y - me.y, dz = p.z - me.z;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          // r < max(hs_i, hs_j): non-negative floats order as their bit patterns, so the maximum and the
          // comparison are integer ones (no NaN canonicalisation).
          const bool hit = __float_as_uint(r2) < max(myhs2u, __float_as_uint(p.w));
          if (k < 32) hlo |= hit ? (1u << k) : 0u;
          else hhi |= hit ? (1u << (k - 32)) : 0u;
        }
      }
    }
    uint64_t hits = valid ? ((uint64_t)hhi << 32) | hlo : 0ull;
    while (hits) {
      const uint32_t k = (uint32_t)__builtin_ctzll(hits);
      hits &= hits - 1;
      const uint32_t j = S.pidx[k];
      if (j == tt) continue;
      // Neighbours are staged four at a time (shifting in at b3) and written as one 16-byte chunk.
      b0 = b1;
      b1 = b2;
      b2 = b3;
      b3 = j;
      ++cnt;
      if (++nb == 4) {
        if (nchunk < cap4) nbr4[(int64_t)nchunk * stride + t] = make_uint4(b0, b1, b2, b3);
        ++nchunk;
        nb = 0;
      }
    }
    lhead += nl;
    wave_sync();
  };

  while (sp > 0) {
    const uint32_t take = min(sp, 64u);
    const uint32_t base = sp - take;
    const uint32_t node = lane < take ? S.stack[base + lane] : IHS_NONE;
    sp = base;
    nnodes += take;
    wave_sync();
    bool pass = false;
    int2 rg = make_int2(0, 0), ch = make_int2(0, 0);
    if (node != IHS_NONE) {
      const float4 c0 = nbmin[node], c1 = nbmax[node];
      if (node < leaf0) {  // fetched with the box: no second dependent load for an opened node
        rg = range[node];
        ch = child[node];
      }
      const float dx = fmaxf(fmaxf(c0.x - gx1, gx0 - c1.x), 0.f);
      const float dy = fmaxf(fmaxf(c0.y - gy1, gy0 - c1.y), 0.f);
      const float dz = fmaxf(fmaxf(c0.z - gz1, gz0 - c1.z), 0.f);
      const float R = fmaxf(ghs, c0.w);
      pass = dx * dx + dy * dy + dz * dz < R * R;
    }
    const uint32_t nleaf = node >= leaf0 ? 1u : (uint32_t)(rg.y - rg.x + 1);
    const bool whole = pass && nleaf <= (uint32_t)NB_SMALL;
    const bool open = pass && !whole;
    const uint32_t first = node >= leaf0 ? node - leaf0 : (uint32_t)rg.x;
    const uint32_t take_l = whole ? nleaf : 0u;
    const uint32_t off = wave_excl_scan(take_l);
    for (uint32_t k = 0; k < take_l; ++k) S.leaf[(ltail + off + k) & (NB_RING - 1)] = first + k;
    ltail += wave_sum(take_l);
    const uint64_t mo = __ballot(open);
    const uint32_t nopen = (uint32_t)__popcll(mo);
    if (sp + 2u * nopen > (uint32_t)NB_STACK) {
      if (lane == 0) atomicAdd(&flags[FLAG_WALK_OVERFLOW], 1u);
    } else {
      if (open) {
        const uint32_t pos = sp + 2u * lanes_below(mo);
        S.stack[pos] = (uint32_t)ch.x;
        S.stack[pos + 1] = (uint32_t)ch.y;
      }
      sp += 2u * nopen;
    }
    wave_sync();
    while (ltail - lhead >= lpc) flush(lpc);
  }
  while (ltail != lhead) flush(min(lpc, ltail - lhead));
  // A partial last chunk holds its nb entries in b[4-nb..3]: move them to the front, padding with the last.
  if (valid && nb > 0 && nchunk < cap4)
    nbr4[(int64_t)nchunk * stride + t] = nb == 1 ? make_uint4(b3, b3, b3, b3)
                                      : nb == 2 ? make_uint4(b2, b3, b3, b3) : make_uint4(b1, b2, b3, b3);
  const uint32_t capn = cap4 * 4u;
  if (valid) ncount[t] = min(cnt, capn);
  const uint64_t over = __ballot(valid && cnt > capn);
  npairs += wave_sum((unsigned long long)(valid ? min(cnt, capn) : 0u));
  if (lane == 0 && over) atomicAdd(&flags[FLAG_NBR_OVERFLOW], (uint32_t)__popcll(over));
}

// Persistent: waves pull groups from per-XCD queues (wq[8..16)), then help the other queues.
__global__ __launch_bounds__(NB_WAVES * 64) void k_nbr_build(const float4* __restrict__ xm, const float* __restrict__ hs,
                                                             int64_t n, int K, int64_t L, const int2* __restrict__ child,
                                                             const int2* __restrict__ range,
                                                             const float4* __restrict__ nbmin,
                                                             const float4* __restrict__ nbmax,
                                                             const int4* __restrict__ groups,
                                                             const uint32_t* __restrict__ ngroups,
                                                             uint32_t* __restrict__ wq, uint4* __restrict__ nbr4,
                                                             uint32_t* __restrict__ ncount, int cap, int64_t stride,
                                                             uint32_t* __restrict__ flags,
                                                             unsigned long long* __restrict__ counters,
                                                             const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  __shared__ NbrLds lds[NB_WAVES];
  const int w = threadIdx.x >> 6;
  const uint32_t lane = threadIdx.x & 63;
  NbrLds& S = lds[w];
  const uint32_t ng = *ngroups;
  const uint32_t home = blockIdx.x & 7u;
  uint64_t ncand = 0, nnodes = 0, npairs = 0;
  for (uint32_t r = 0; r < 8; ++r) {
    const uint32_t q = (home + r) & 7u;
    const uint32_t lo = (uint32_t)((uint64_t)ng * q / 8u), hi = (uint32_t)((uint64_t)ng * (q + 1u) / 8u);
    for (;;) {
      uint32_t g = 0;
      if (lane == 0) g = atomicAdd(&wq[8 + q], 1u);
      g = (uint32_t)__builtin_amdgcn_readfirstlane((int)g) + lo;
      if (g >= hi) break;
      nbr_group(S, groups[g], lane, xm, hs, n, K, L, child, range, nbmin, nbmax, nbr4, ncount, cap, stride, flags,
                ncand, nnodes, npairs);
    }
  }
  if (lane == 0) {
    cnt_add(counters, CNT_PAIRS, npairs);
    cnt_add(counters, CNT_NBR_CAND, ncand);
    cnt_add(counters, CNT_WALK_NODES, nnodes);
  }
}

// Visits the neighbour list of particle t: 16-byte chunks of four indices, chunk c of particle t at
// c * stride + t (lanes of a wave read consecutive chunks). Every slot of a written chunk holds a valid
// particle index (slots past the count repeat earlier entries, or 0), so the four sources of a chunk are
// fetched together, before any of them is used, and the next chunk's indices are prefetched: the loop keeps
// several loads in flight instead of waiting on one neighbour at a time. Slots past the count are masked.
template <class Fetch, class Use>
__device__ __forceinline__ void for_neighbours(const uint4* __restrict__ nbr4, int64_t stride, int64_t t, uint32_t cnt,
                                               Fetch&& fetch, Use&& use) {
  if (cnt == 0) return;
  uint4 q = nbr4[t];
  for (uint32_t k = 0; k < cnt; k += 4) {
    const uint4 qn = k + 4 < cnt ? nbr4[(int64_t)((k >> 2) + 1) * stride + t] : q;
    const auto a = fetch(q.x);
    const auto b = fetch(q.y);
    const auto c = fetch(q.z);
    const auto d = fetch(q.w);
    use(q.x, a, true);
    use(q.y, b, k + 1 < cnt);
    use(q.z, c, k + 2 < cnt);
    use(q.w, d, k + 3 < cnt);
    q = qn;
  }
}

// As for_neighbours, visiting only the entries whose bit is set in the particle's support mask (bit k of word
// k / 32 at (k / 32) * stride + t): the sources of the other entries are not fetched.
template <class Fetch, class Use>
__device__ __forceinline__ void for_support(const uint4* __restrict__ nbr4, const uint32_t* __restrict__ mask,
                                            int64_t stride, int64_t t, uint32_t cnt, Fetch&& fetch, Use&& use) {
  if (cnt == 0) return;
  uint4 q = nbr4[t];
  uint32_t mw = mask[t];
  for (uint32_t k = 0; k < cnt; k += 4) {
    const uint32_t kn = k + 4;
    const uint4 qn = kn < cnt ? nbr4[(int64_t)(kn >> 2) * stride + t] : q;
    const uint32_t mwn = (kn & 31u) == 0u && kn < cnt ? mask[(int64_t)(kn >> 5) * stride + t] : mw;
    const uint32_t bits = mw >> (k & 31u);
    using S = decltype(fetch(0u));
    S a{}, b{}, c{}, d{};
    if (bits & 1u) a = fetch(q.x);
    if (bits & 2u) b = fetch(q.y);
    if (bits & 4u) c = fetch(q.z);
    if (bits & 8u) d = fetch(q.w);
    use(q.x, a, (bits & 1u) != 0u);
    use(q.y, b, (bits & 2u) != 0u);
    use(q.z, c, (bits & 4u) != 0u);
    use(q.w, d, (bits & 8u) != 0u);
    q = qn;
    mw = mwn;
  }
}

// Cubic spline with support r < h (sph_step.wgsl:64-77) from q = r/h and 1/h: W and dW/dr.
__device__ __forceinline__ void spline_wdw(float q, float hinv, float& w, float& dw) {
  const float sig = (8.0f / kPI32) * hinv * hinv * hinv;
  if (q < 0.5f) {
    w = sig * (1.0f - 6.0f * q * q + 6.0f * q * q * q);
    dw = sig * (-12.0f * q + 18.0f * q * q) * hinv;
  } else if (q < 1.0f) {
    const float tq = 1.0f - q;
    w = sig * 2.0f * tq * tq * tq;
    dw = sig * (-6.0f * tq * tq) * hinv;
  } else {
    w = 0.f;
    dw = 0.f;
  }
}

template <bool REL>
__device__ __forceinline__ float own_density(const float4 me, const float3 melo, float hh,
                                             const uint4* __restrict__ nbr4, int64_t stride, int64_t t, uint32_t cnt,
                                             const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                             const XWin& W, float& hdrho, uint32_t& nngb) {
  const float hinv = 1.0f / hh;
  const float sig = (8.0f / kPI32) * hinv * hinv * hinv;
  float rho = me.w * sig;
  float hd = -3.0f * rho;
  const float h2 = hh * hh;
  uint32_t c = 0;
  for_neighbours(nbr4, stride, t, cnt, [&](uint32_t j) { return win_xm<REL>(W, xm, xlo, j); },
                 [&](uint32_t, const PosLo p, bool ok) {
    const float4 q = p.q;
    const float3 d = pair_d<REL>(me, melo, q, p.c);
    const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (ok && r2 < h2) {
      const float qq = __builtin_amdgcn_sqrtf(r2) * hinv;
      float wq, dq;
      if (qq < 0.5f) {
        wq = 1.0f - 6.0f * qq * qq + 6.0f * qq * qq * qq;
        dq = -12.0f * qq + 18.0f * qq * qq;
      } else {
        const float tt = 1.0f - qq;
        wq = 2.0f * tt * tt * tt;
        dq = -6.0f * tt * tt;
      }
      const float wv = sig * wq;
      rho += q.w * wv;
      hd -= q.w * (3.0f * wv + qq * sig * dq);
      ++c;
    }
  });
  hdrho = hd;
  nngb = c;
  return rho;
}

// Newton iteration on f(h) = rho~(h) - m (eta/h)^3 from hh with the bracket (lo, hi): stops when
// |f| <= tol rho (DONE), when the searched radius stops a growing h (CAPPED), after max_iter updates
// (UNCONVERGED), or after max_evals density sums with an update pending (PENDING: resume from (h, lo, hi)).
enum : uint32_t { HS_PENDING = 0, HS_DONE = 1, HS_CAPPED = 2, HS_UNCONVERGED = 3 };

struct HsolveState {
  float h, lo, hi;
  uint32_t it, evals, state;
};

// Own-h density of particle t summed by a group of HS2_G lanes: lane `sub` takes chunks sub, sub + HS2_G, ... of
// the list, and a butterfly over the group gives every lane of it the same totals (a + b == b + a in floating point),
// so the group's lanes take the same Newton steps.
constexpr int HS2_G = 4;
template <bool REL>
__device__ __forceinline__ float own_density_group(const float4 me, const float3 melo, float hh,
                                                   const uint4* __restrict__ nbr4, int64_t stride, int64_t t,
                                                   uint32_t cnt, const float4* __restrict__ xm,
                                                   const uint32_t* __restrict__ xlo, uint32_t sub, float& hdrho) {
  const float hinv = 1.0f / hh;
  const float sig = (8.0f / kPI32) * hinv * hinv * hinv;
  const float h2 = hh * hh;
  float rho = 0.f, hd = 0.f;
  for (uint32_t c = sub; 4u * c < cnt; c += HS2_G) {
    const uint4 q = nbr4[(int64_t)c * stride + t];
    const uint32_t jj[4] = {q.x, q.y, q.z, q.w};
    float4 src[4];
    uint32_t lc[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      src[e] = xm[jj[e]];
      if (REL) lc[e] = xlo[jj[e]];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float4 qv = src[e];
      const float3 d = pair_d<REL>(me, melo, qv, lc[e]);
      const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
      if (4u * c + (uint32_t)e < cnt && r2 < h2) {
        const float qq = __builtin_amdgcn_sqrtf(r2) * hinv;
        float wq, dq;
        if (qq < 0.5f) {
          wq = 1.0f - 6.0f * qq * qq + 6.0f * qq * qq * qq;
          dq = -12.0f * qq + 18.0f * qq * qq;
        } else {
          const float tt = 1.0f - qq;
          wq = 2.0f * tt * tt * tt;
          dq = -6.0f * tt * tt;
        }
        const float wv = sig * wq;
        rho += qv.w * wv;
        hd -= qv.w * (3.0f * wv + qq * sig * dq);
      }
    }
  }
#pragma unroll
  for (int o = 1; o < HS2_G; o <<= 1) {
    rho += __shfl_xor(rho, o, 64);
    hd += __shfl_xor(hd, o, 64);
  }
  const float self = me.w * sig;
  hdrho = hd - 3.0f * self;
  return rho + self;
}

// hsolve_loop for a group of HS2_G lanes sharing one particle (every lane of the group holds the same state).
template <bool REL>
__device__ __forceinline__ void hsolve_loop_group(HsolveState& st, const float4 me, const float3 melo, float hsi,
                                                  const uint4* __restrict__ nbr4, int64_t stride, int64_t t,
                                                  uint32_t cnt, const float4* __restrict__ xm,
                                                  const uint32_t* __restrict__ xlo, uint32_t sub, float eta3,
                                                  float tol, int max_iter) {
  float hh = st.h;
  for (;;) {
    float hd;
    const float rho = own_density_group<REL>(me, melo, hh, nbr4, stride, t, cnt, xm, xlo, sub, hd);
    ++st.evals;
    const float hinv = 1.0f / hh;
    const float target = me.w * eta3 * hinv * hinv * hinv;
    const float f = rho - target;
    if (fabsf(f) <= tol * rho) {
      st.state = HS_DONE;
      break;
    }
    if ((int)st.it >= max_iter) {
      st.state = HS_UNCONVERGED;
      break;
    }
    if (f < 0.f) st.lo = hh;
    else st.hi = hh;
    if (f < 0.f && hh >= hsi) {
      st.state = HS_CAPPED;
      break;
    }
    const float hdf = hd + 3.0f * target;  // h * df/dh
    float hn = hdf > 0.f ? hh - f * hh / hdf : -1.0f;
    if (!(hn > st.lo && hn < st.hi)) hn = isinf(st.hi) ? 2.0f * hh : 0.5f * (st.lo + st.hi);
    hh = fminf(hn, hsi);
    ++st.it;
  }
  st.h = hh;
}

template <bool REL>
__device__ __forceinline__ void hsolve_loop(HsolveState& st, const float4 me, const float3 melo, float hsi,
                                            uint32_t max_evals, const uint4* __restrict__ nbr4, int64_t stride,
                                            int64_t t, uint32_t cnt, const float4* __restrict__ xm,
                                            const uint32_t* __restrict__ xlo, const XWin& W, float eta3, float tol,
                                            int max_iter) {
  float hh = st.h;
  for (;;) {
    float hd;
    uint32_t nn;
    const float rho = own_density<REL>(me, melo, hh, nbr4, stride, t, cnt, xm, xlo, W, hd, nn);
    ++st.evals;
    const float hinv = 1.0f / hh;
    const float target = me.w * eta3 * hinv * hinv * hinv;
    const float f = rho - target;
    if (fabsf(f) <= tol * rho) {
      st.state = HS_DONE;
      break;
    }
    if ((int)st.it >= max_iter) {
      st.state = HS_UNCONVERGED;
      break;
    }
    if (f < 0.f) st.lo = hh;
    else st.hi = hh;
    if (f < 0.f && hh >= hsi) {
      st.state = HS_CAPPED;
      break;
    }
    const float hdf = hd + 3.0f * target;  // h * df/dh
    float hn = hdf > 0.f ? hh - f * hh / hdf : -1.0f;
    if (!(hn > st.lo && hn < st.hi)) hn = isinf(st.hi) ? 2.0f * hh : 0.5f * (st.lo + st.hi);
    hh = fminf(hn, hsi);
    ++st.it;
    if (st.evals >= max_evals) {
      st.state = HS_PENDING;
      break;
    }
  }
  st.h = hh;
}

// Wave-aggregated flags and counters. A solve stopped by a search radius below the global cap asks for the
// wider-radius pass instead of being reported capped when `retry` is given.
__device__ __forceinline__ void hsolve_account(bool valid, const HsolveState& st, float hsi, uint32_t* __restrict__ retry,
                                               float hcap /* this particle's segment cap */, uint32_t* __restrict__ flags,
                                               unsigned long long* __restrict__ counters, bool first_pass) {
  const bool capped = valid && st.state == HS_CAPPED;
  const bool again = capped && retry && hsi < hcap;
  const uint64_t mc = __ballot(capped && !again), mr = __ballot(again);
  const uint64_t mu = __ballot(valid && st.state == HS_UNCONVERGED);
  const unsigned long long si = wave_sum((unsigned long long)(valid ? st.it : 0u));
  const unsigned long long se = wave_sum((unsigned long long)(valid ? st.evals : 0u));
  const unsigned long long sn = (unsigned long long)__popcll(__ballot(valid));
  if ((threadIdx.x & 63) == 0 && sn) {
    if (mc) atomicAdd(&flags[FLAG_H_CAPPED], (uint32_t)__popcll(mc));
    if (mu) atomicAdd(&flags[FLAG_H_UNCONVERGED], (uint32_t)__popcll(mu));
    if (mr) atomicAdd(retry, (uint32_t)__popcll(mr));
    cnt_add(counters, CNT_H_ITERS, si);
    cnt_add(counters, CNT_H_EVALS, se);
    if (first_pass) cnt_add(counters, CNT_H_SOLVED, sn);
  }
}

// First pass: one density sum per particle at its seed h. Converged particles are done; the others take one
// Newton update and join a pending list (wave-aggregated append, in particle order within a wave; block b appends to
// shard b mod PEND_SHARDS, whose entries start at shard * region) with their bracket, so that the second pass runs
// full waves of unconverged particles instead of every wave waiting for its slowest lane. One shared counter took
// an L2 atomic from every wave with a pending lane, one after another (140 us per seeded evaluation at 1M).
template <bool REL>
__global__ __launch_bounds__(256) void k_hsolve1(const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                                 const float* __restrict__ hs,
                                                 float* __restrict__ h, uint8_t* __restrict__ iters,
                                                 const uint4* __restrict__ nbr4, const uint32_t* __restrict__ ncount,
                                                 int64_t stride, int64_t n, float eta3, float tol, int max_iter,
                                                 const uint32_t* __restrict__ gate, uint32_t* __restrict__ retry,
                                                 SegView sv, uint32_t* __restrict__ pend, float2* __restrict__ bracket,
                                                 uint32_t* __restrict__ npend, uint32_t region,
                                                 uint32_t* __restrict__ flags, unsigned long long* __restrict__ counters) {
  if (gate && *gate == 0u) return;
  __shared__ float4 wx[DWIN];
  __shared__ uint32_t wl[REL ? DWIN : 1];
  const int64_t t0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  int64_t wlo, whi;
  win_bounds(t0, n, DWIN_HALO, wlo, whi);
  for (int64_t k = wlo + threadIdx.x; k < whi; k += blockDim.x) {
    wx[k - wlo] = xm[k];
    if (REL) wl[k - wlo] = xlo[k];
  }
  __syncthreads();
  const XWin W{wx, wl, (uint32_t)wlo, (uint32_t)(whi - wlo)};
  const int64_t t = t0 + threadIdx.x;
  const bool valid = t < n;
  HsolveState st{0.f, 0.f, INFINITY, 0u, 0u, HS_DONE};
  float hsi = 0.f, hcap = 0.f;
  if (valid) {
    hsi = hs[t];
    hcap = sv.hcap[sv.of(t)];
    st.h = fminf(h[t], hsi);
    const float4 me = xm[t];
    hsolve_loop<REL>(st, me, own_lo<REL>(xlo, t, me), hsi, 1u, nbr4, stride, t, ncount[t], xm, xlo, W, eta3, tol,
                     max_iter);
    h[t] = st.h;
  }
  const bool more = valid && st.state == HS_PENDING;
  const uint64_t mm = __ballot(more);
  if (mm) {
    const uint32_t shard = blockIdx.x & (PEND_SHARDS - 1);
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(&npend[shard], (uint32_t)__popcll(mm));
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base) + shard * region;
    if (more) {
      pend[base + lanes_below(mm)] = (uint32_t)t;
      bracket[t] = make_float2(st.lo, st.hi);
    }
  }
  if (valid && !more) iters[t] = (uint8_t)min(st.it, 255u);
  hsolve_account(valid, st, hsi, retry, hcap, flags, counters, true);
}

// Second pass over the pending lists (persistent grid-stride loop over their concatenation): resumes each solve
// from its first update.
template <bool REL>
__global__ __launch_bounds__(256) void k_hsolve2(const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                                 const float* __restrict__ hs,
                                                 float* __restrict__ h, uint8_t* __restrict__ iters,
                                                 const uint4* __restrict__ nbr4, const uint32_t* __restrict__ ncount,
                                                 int64_t stride, float eta3, float tol, int max_iter,
                                                 const uint32_t* __restrict__ gate, uint32_t* __restrict__ retry,
                                                 SegView sv, const uint32_t* __restrict__ pend,
                                                 const float2* __restrict__ bracket, const uint32_t* __restrict__ npend,
                                                 uint32_t region, uint32_t* __restrict__ flags,
                                                 unsigned long long* __restrict__ counters) {
  if (gate && *gate == 0u) return;
  // Offsets of the shards in the concatenation: soff[r] = sum of the lengths of shards below r.
  __shared__ uint32_t soff[PEND_SHARDS + 1];
  if (threadIdx.x < 64) {
    static_assert(PEND_SHARDS == 64, "one wave scans the shard lengths");
    uint32_t c = npend[threadIdx.x];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(c, o, 64);
      if ((int)threadIdx.x >= o) c += y;
    }
    soff[threadIdx.x + 1] = c;
    if (threadIdx.x == 0) soff[0] = 0u;
  }
  __syncthreads();
  const uint32_t np = soff[PEND_SHARDS];
  // HS2_G lanes per pending particle: 64 / HS2_G particles per wave and round. Every lane of a wave runs the same
  // number of rounds so the wave-wide accounting stays converged.
  constexpr uint32_t PER_WAVE = 64u / HS2_G;
  const uint32_t lane = threadIdx.x & 63u, sub = lane % HS2_G;
  const uint32_t wave = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) / 64u;
  const uint32_t nwave = gridDim.x * blockDim.x / 64u;
  for (uint32_t b = wave * PER_WAVE; b < np; b += nwave * PER_WAVE) {
    const uint32_t k = b + lane / HS2_G;
    const bool valid = k < np;
    HsolveState st{0.f, 0.f, INFINITY, 1u, 0u, HS_DONE};
    float hsi = 0.f, hcap = 0.f;
    int64_t t = 0;
    if (valid) {
      uint32_t lo = 0, hi = PEND_SHARDS;  // the shard r with soff[r] <= k < soff[r + 1]
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (soff[mid] <= k) lo = mid;
        else hi = mid;
      }
      t = pend[lo * region + (k - soff[lo])];
      hsi = hs[t];
      hcap = sv.hcap[sv.of(t)];
      const float2 br = bracket[t];
      st.h = h[t];
      st.lo = br.x;
      st.hi = br.y;
      const float4 me = xm[t];
      hsolve_loop_group<REL>(st, me, own_lo<REL>(xlo, t, me), hsi, nbr4, stride, t, ncount[t], xm, xlo, sub, eta3,
                             tol, max_iter);
      if (sub == 0) {
        h[t] = st.h;
        iters[t] = (uint8_t)min(st.it, 255u);
      }
    }
    const bool lead = valid && sub == 0;
    st.it = lead ? st.it - 1u : 0u;  // the first update was counted by the first pass
    if (!lead) st.evals = 0u;
    hsolve_account(lead, st, hsi, retry, hcap, flags, counters, false);
  }
}

// Density with h_ij, EOS, Courant signal; optional MLP features (neighbour count within h_i,
// div v and |grad rho|, both with the h_ij kernel gradient).
// The Courant signal is one word per segment that every block would otherwise contend on: one atomic per
// block when the block lies in one segment (always, for a single scenario), else one per lane. Blocks of 256.
__device__ __forceinline__ void signal_reduce(uint32_t key, int64_t t, int64_t n, bool valid, const SegView& sv) {
  __shared__ uint32_t skey[4];
  const int64_t b0 = t - threadIdx.x;  // the block's first particle, whatever the block-to-range mapping
  const uint32_t sg0 = sv.of(b0), sg1 = sv.of(min(b0 + (int64_t)blockDim.x, n) - 1);
  if (sg0 == sg1) {
    key = wave_min(key);
    if ((threadIdx.x & 63) == 0) skey[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
      key = min(min(skey[0], skey[1]), min(skey[2], skey[3]));
      // Most blocks cannot lower the minimum: a plain read first keeps them off the one address (a stale read
      // only costs an atomic that changes nothing).
      if (key != 0xffffffffu && key < __hip_atomic_load(&sv.clk[sg0].signal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(&sv.clk[sg0].signal, key);
    }
  } else if (valid && key != 0xffffffffu) {
    atomicMin(&sv.clk[sv.of(t)].signal, key);
  }
}

// Velocity-only evaluation at unchanged positions: the density of the last evaluation stands (positions and h
// are the same), so only the EOS (u changed with the kick), the particle velocities and the Courant signal are
// refreshed before the forces.
template <class PT>
__global__ __launch_bounds__(256) void k_eos(int64_t n, const float* __restrict__ rho_in, const float* __restrict__ h,
                                             const float* __restrict__ vx, const float* __restrict__ vy,
                                             const float* __restrict__ vz, const PT* __restrict__ u,
                                             const uint8_t* __restrict__ mat, const EosDev* __restrict__ eos,
                                             float4* __restrict__ rpc, float4* __restrict__ vh,
                                             uint32_t* __restrict__ flags, SegView sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < n;
  uint32_t key = 0xffffffffu;
  if (valid) {
    const float rho = rho_in[t], hi = h[t];
    const float vxi = vx[t], vyi = vy[t], vzi = vz[t];
    const EosDev e = eos[mat[t]];
    const float uu = (float)u[t];
    const float p = till_pressure(e, rho, uu);
    const float c = till_sound_speed(e, rho, uu, p);
    rpc[t] = make_float4(rho, p / (rho * rho), c, p);
    vh[t] = make_float4(vxi, vyi, vzi, hi);
    const float sig = hi / fmaxf(c + sqrtf(vxi * vxi + vyi * vyi + vzi * vzi), 1.0f);
    key = f2key(sig);
    if (!(sig == sig)) atomicAdd(&flags[FLAG_NONFINITE], 1u);
  }
  signal_reduce(key, t, n, valid, sv);
}

// Density, EOS and the Courant signal, one thread per particle.
// vh holds (v, h_prev) from the permutation pass on entry; its h is set here to the solved h.
// FEAT: also the network's first input, the neighbour count within h (feat.x); the force pass of the same
// evaluation adds div v and |grad rho| (feat.y, feat.z) from its pair terms, so no pass here gathers velocities.
// mask (optional): the support mask of each list, bit k set when entry k is inside the kernel support
// (r < h_ij), the pairs the force pass sums; entries between h_ij and the searched radius are then not fetched
// again. The words are assembled in LDS (thread-private) and written once, after the loop.
template <class PT, bool FEAT, bool REL>
__global__ __launch_bounds__(256) void k_density(const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                                 const float* __restrict__ h,
                                                 const PT* __restrict__ u, const uint8_t* __restrict__ mat,
                                                 const EosDev* __restrict__ eos, const uint4* __restrict__ nbr4,
                                                 const uint32_t* __restrict__ ncount, int64_t stride, int64_t n,
                                                 float4* __restrict__ rpc, float4* vh, float* __restrict__ rho_out,
                                                 float4* __restrict__ feat, uint32_t* __restrict__ mask,
                                                 uint32_t* __restrict__ flags, SegView sv) {
  __shared__ uint32_t smask[NMASK_WORDS * 256];
  __shared__ float4 wx[DWIN];
  __shared__ float wh[DWIN];
  __shared__ uint32_t wl[REL ? DWIN : 1];
  const int64_t t0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  int64_t wlo, whi;
  win_bounds(t0, n, DWIN_HALO, wlo, whi);
  for (int64_t k = wlo + threadIdx.x; k < whi; k += blockDim.x) {
    wx[k - wlo] = xm[k];
    wh[k - wlo] = h[k];
    if (REL) wl[k - wlo] = xlo[k];
  }
  __syncthreads();
  const uint32_t wlb = (uint32_t)wlo, wn = (uint32_t)(whi - wlo);
  const int64_t t = t0 + threadIdx.x;
  const bool valid = t < n;
  uint32_t key = 0xffffffffu;
  if (valid) {
    const float4 me = xm[t];
    const float3 melo = own_lo<REL>(xlo, t, me);
    const float hi = h[t];
    const float4 vi = vh[t];
    const uint32_t cnt = ncount[t];
    const float hinv_i = 1.0f / hi;
    float rho = me.w * (8.0f / kPI32) * hinv_i * hinv_i * hinv_i;
    uint32_t nn = 0, ent = 0;
    const uint32_t nwords = (cnt + 31u) >> 5;
    uint32_t* my = &smask[threadIdx.x];
    if (mask)
      for (uint32_t w = 0; w < nwords; ++w) my[w * 256] = 0u;
    struct Src {
      float4 q;
      float hj;
      uint32_t c;
    };
    for_neighbours(nbr4, stride, t, cnt,
                   [&](uint32_t j) {
                     const uint32_t o = j - wlb;
                     if (o < wn) return Src{wx[o], wh[o], REL ? wl[o] : 0u};
                     return Src{xm[j], h[j], REL ? xlo[j] : 0u};
                   },
                   [&](uint32_t, const Src src, bool ok) {
      const float4 q = src.q;
      const float3 d = pair_d<REL>(me, melo, q, src.c);
      const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
      const float hij = 0.5f * (hi + src.hj);
      if (FEAT && ok && r2 < hi * hi) ++nn;
      if (ok && r2 < hij * hij) {
        const float r = __builtin_amdgcn_sqrtf(r2);
        const float hinv = __builtin_amdgcn_rcpf(hij);
        float w, dw;
        spline_wdw(r * hinv, hinv, w, dw);
        rho += q.w * w;
        if (mask) atomicOr(&my[(ent >> 5) * 256], 1u << (ent & 31u));
      }
      ++ent;
    });
    if (mask)
      for (uint32_t w = 0; w < nwords; ++w) mask[(int64_t)w * stride + t] = my[w * 256];
    const EosDev e = eos[mat[t]];
    const float uu = (float)u[t];
    const float p = till_pressure(e, rho, uu);
    const float c = till_sound_speed(e, rho, uu, p);
    rpc[t] = make_float4(rho, p / (rho * rho), c, p);
    vh[t].w = hi;
    rho_out[t] = rho;
    if (FEAT) feat[t] = make_float4((float)nn, 0.f, 0.f, 0.f);
    const float sig = hi / fmaxf(c + sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z), 1.0f);
    key = f2key(sig);
    if (!(sig == sig) || !(rho == rho)) atomicAdd(&flags[FLAG_NONFINITE], 1u);
  }
  signal_reduce(key, t, n, valid, sv);
}

// DIVV: also form the network's pair-term inputs from the same pair terms (the support pairs, r < h_ij, in list
// order): the SPH velocity divergence (feat.y) and |grad rho| (feat.z). The density pass wrote feat.x.
// The block's Morton neighbourhood, particles [t0 - FWIN_HALO, t0 + 256 + FWIN_HALO), is staged in LDS with
// coalesced loads; neighbours inside it are read from LDS and only the others are gathered from global memory
// (each gathered lane costs its own L2 request, the bound of these passes). In Morton order about two thirds of
// the pairs of the impact fall inside such a window.
template <bool KAHAN, bool DIVV, bool REL>
__global__ __launch_bounds__(256) void k_forces(const float4* __restrict__ xm, const uint32_t* __restrict__ xlo,
                                                const float4* __restrict__ vh,
                                                const float4* __restrict__ rpc, const uint4* __restrict__ nbr4,
                                                const uint32_t* __restrict__ ncount,
                                                const uint32_t* __restrict__ mask, int64_t stride, int64_t n,
                                                const float* __restrict__ av, float4* __restrict__ asph,
                                                float4* __restrict__ feat) {
  __shared__ float4 wx[FWIN], wv[FWIN], wr[FWIN];
  __shared__ uint32_t wl[REL ? FWIN : 1];
  const int64_t t0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  int64_t wlo, whi;
  win_bounds(t0, n, FWIN_HALO, wlo, whi);
  for (int64_t k = wlo + threadIdx.x; k < whi; k += blockDim.x) {
    wx[k - wlo] = xm[k];
    wv[k - wlo] = vh[k];
    wr[k - wlo] = rpc[k];
    if (REL) wl[k - wlo] = xlo[k];
  }
  __syncthreads();
  const uint32_t wlb = (uint32_t)wlo, wn = (uint32_t)(whi - wlo);
  const int64_t t = t0 + threadIdx.x;
  if (t >= n) return;
  const float alpha = av[0], beta = av[1];
  const float4 me = xm[t];
  const float3 melo = own_lo<REL>(xlo, t, me);
  const float4 vi = vh[t];
  const float4 ri = rpc[t];
  const uint32_t cnt = ncount[t];
  float ax = 0.f, ay = 0.f, az = 0.f, de = 0.f, dv = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  Kahan3 ka;
  ka.zero();
  struct Src {
    float4 q, vj, rj;
    uint32_t c;
  };
  auto fetch = [&](uint32_t j) {
    const uint32_t o = j - wlb;
    if (o < wn) return Src{wx[o], wv[o], wr[o], REL ? wl[o] : 0u};
    return Src{xm[j], vh[j], rpc[j], REL ? xlo[j] : 0u};
  };
  auto use = [&](uint32_t, const Src& src, bool ok) {
    const float4 q = src.q, vj = src.vj, rj = src.rj;
    const float3 d = pair_d<REL>(me, melo, q, src.c);
    const float dx = d.x, dy = d.y, dz = d.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float hij = 0.5f * (vi.w + vj.w);
    if (!ok || !(r2 < hij * hij) || !(r2 > 1.0e-18f)) return;
    const float r = __builtin_amdgcn_sqrtf(r2);
    const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
    const float vr = dvx * dx + dvy * dy + dvz * dz;
    float pij = 0.f;
    if (vr < 0.f) {
      // Monaghan's viscosity; the two quotients use the hardware reciprocal (1 ulp) instead of IEEE divisions,
      // which were a fifth of the pair's instructions.
      const float mu = hij * vr * __builtin_amdgcn_rcpf(r2 + 0.01f * hij * hij);
      const float cbar = 0.5f * (ri.z + rj.z);
      const float rbar = 0.5f * (ri.x + rj.x);
      pij = (-alpha * cbar * mu + beta * mu * mu) * __builtin_amdgcn_rcpf(rbar);
    }
    const float coeff = ri.y + rj.y + pij;
    const float hinv = __builtin_amdgcn_rcpf(hij);
    float w, dw;
    spline_wdw(r * hinv, hinv, w, dw);
    const float gw = dw * __builtin_amdgcn_rcpf(r);
    const float fx = dx * gw, fy = dy * gw, fz = dz * gw;
    const float s = -coeff * q.w;
    if (KAHAN) {
      ka.add(fx * s, 0);
      ka.add(fy * s, 1);
      ka.add(fz * s, 2);
    } else {
      ax = fmaf(fx, s, ax);
      ay = fmaf(fy, s, ay);
      az = fmaf(fz, s, az);
    }
    const float vdw = dvx * fx + dvy * fy + dvz * fz;
    de += 0.5f * q.w * coeff * vdw;
    if (DIVV) {
      gx += q.w * fx;
      gy += q.w * fy;
      gz += q.w * fz;
      dv += q.w * vdw;
    }
  };
  if (mask) for_support(nbr4, mask, stride, t, cnt, fetch, use);
  else for_neighbours(nbr4, stride, t, cnt, fetch, use);
  if (KAHAN) {
    ax = ka.s[0];
    ay = ka.s[1];
    az = ka.s[2];
  }
  asph[t] = make_float4(ax, ay, az, de);
  if (DIVV) {
    feat[t].y = -dv / ri.x;
    feat[t].z = sqrtf(gx * gx + gy * gy + gz * gz);
  }
}

inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_nbr_build(hipStream_t s, int64_t n, int K, const float4* xm, const float* hs, const TreeBufs& tb,
                      uint32_t* nbr, uint32_t* ncount, int cap, int64_t stride, uint32_t* flags,
                      unsigned long long* counters, const uint32_t* gate) {
  uint4* nbr4 = (uint4*)nbr;
  const int64_t L = (n + K - 1) / K;
  static unsigned nblk = 0;
  if (!nblk) nblk = persistent_blocks((const void*)k_nbr_build, NB_WAVES * 64);
  k_nbr_build<<<nblk, NB_WAVES * 64, 0, s>>>(xm, hs, n, K, L, tb.child, tb.range, tb.nbmin, tb.nbmax, tb.groups,
                                             tb.ngroups, tb.wq, nbr4, ncount, cap, stride, flags, counters, gate);
}

template <bool REL>
void hsolve_launch(hipStream_t s, int64_t n, const float4* xm, const uint32_t* xlo, const float* hs, float* h,
                   uint8_t* iters, const uint32_t* nbr, const uint32_t* ncount, int64_t stride, float eta3, float tol,
                   int max_iter, const HsolveScratch& hw, uint32_t* flags, unsigned long long* counters,
                   const SegView& sv, const uint32_t* gate, uint32_t* retry) {
  static unsigned nblk2 = 0;
  if (!nblk2) nblk2 = persistent_blocks((const void*)k_hsolve2<REL>, 256);
  // Shard r holds the pending particles of blocks r, r + PEND_SHARDS, ...: at most region entries.
  const uint32_t region = (uint32_t)((g256(n) + PEND_SHARDS - 1) / PEND_SHARDS) * 256u;
  k_hsolve1<REL><<<g256(n), 256, 0, s>>>(xm, xlo, hs, h, iters, (const uint4*)nbr, ncount, stride, n, eta3, tol,
                                         max_iter, gate, retry, sv, hw.pend, hw.bracket, hw.npend, region, flags,
                                         counters);
  k_hsolve2<REL><<<nblk2, 256, 0, s>>>(xm, xlo, hs, h, iters, (const uint4*)nbr, ncount, stride, eta3, tol, max_iter,
                                       gate, retry, sv, hw.pend, hw.bracket, hw.npend, region, flags, counters);
}

void launch_hsolve(hipStream_t s, int64_t n, const float4* xm, const uint32_t* xlo, const float* hs, float* h,
                   uint8_t* iters, const uint32_t* nbr, const uint32_t* ncount, int64_t stride, float eta3, float tol,
                   int max_iter, const HsolveScratch& hw, uint32_t* flags, unsigned long long* counters,
                   const SegView& sv, const uint32_t* gate, uint32_t* retry) {
  if (xlo)
    hsolve_launch<true>(s, n, xm, xlo, hs, h, iters, nbr, ncount, stride, eta3, tol, max_iter, hw, flags, counters, sv,
                        gate, retry);
  else
    hsolve_launch<false>(s, n, xm, xlo, hs, h, iters, nbr, ncount, stride, eta3, tol, max_iter, hw, flags, counters,
                         sv, gate, retry);
}

void launch_density(hipStream_t s, int split, int feat, int64_t n, const float4* xm, const uint32_t* xlo, const float* h,
                    const void* u, const uint8_t* mat, const void* eos, const uint32_t* nbr, const uint32_t* ncount,
                    int64_t stride, float4* rpc, float4* vh, float* rho, float4* featv, uint32_t* mask,
                    uint32_t* flags, const SegView& sv) {
  const EosDev* e = (const EosDev*)eos;
#define IHS_DENS(PT, F, R)                                                                                            \
  k_density<PT, F, R><<<g256(n), 256, 0, s>>>(xm, xlo, h, (const PT*)u, mat, e, (const uint4*)nbr, ncount, stride, \
                                              n, rpc, vh, rho, featv, mask, flags, sv)
  if (split && xlo) {
    if (feat) IHS_DENS(double, true, true);
    else IHS_DENS(double, false, true);
  } else if (split) {
    if (feat) IHS_DENS(double, true, false);
    else IHS_DENS(double, false, false);
  } else {
    if (feat) IHS_DENS(float, true, false);
    else IHS_DENS(float, false, false);
  }
#undef IHS_DENS
}

void launch_forces(hipStream_t s, int kahan, int64_t n, const float4* xm, const uint32_t* xlo, const float4* vh,
                   const float4* rpc, const uint32_t* nbr, const uint32_t* ncount, const uint32_t* mask,
                   int64_t stride, const float* av, float4* asph, float4* divv_feat) {
  const uint4* nbr4 = (const uint4*)nbr;
  const unsigned g = g256(n);
#define IHS_FORCES(K, D, R) \
  k_forces<K, D, R><<<g, 256, 0, s>>>(xm, xlo, vh, rpc, nbr4, ncount, mask, stride, n, av, asph, divv_feat)
  if (divv_feat) {
    if (kahan && xlo) IHS_FORCES(true, true, true);
    else if (kahan) IHS_FORCES(true, true, false);
    else IHS_FORCES(false, true, false);
  } else {
    if (kahan && xlo) IHS_FORCES(true, false, true);
    else if (kahan) IHS_FORCES(true, false, false);
    else IHS_FORCES(false, false, false);
  }
#undef IHS_FORCES
}

void launch_eos(hipStream_t s, int split, int64_t n, const float* rho, const float* h, const float* vx, const float* vy,
                const float* vz, const void* u, const uint8_t* mat, const void* eos, float4* rpc, float4* vh,
                uint32_t* flags, const SegView& sv) {
  const EosDev* e = (const EosDev*)eos;
  const unsigned g = std::max(1u, g256(n));
  if (split)
    k_eos<double><<<g, 256, 0, s>>>(n, rho, h, vx, vy, vz, (const double*)u, mat, e, rpc, vh, flags, sv);
  else
    k_eos<float><<<g, 256, 0, s>>>(n, rho, h, vx, vy, vz, (const float*)u, mat, e, rpc, vh, flags, sv);
}
