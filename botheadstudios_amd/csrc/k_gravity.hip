// Barnes-Hut self-gravity over the LBVH: a wavefront-coherent group walk (gfx950, wave64).
//
// One wave owns a walk group: a subtree of at most 64 consecutive Morton-sorted particles (so the group's
// box is bounded by its node's Morton cell, never stretched across a jump of the curve). It walks the tree once for the whole
// group with a short LDS stack, popping up to 64 nodes per step (one per lane). A node is accepted as a
// monopole when the distance from its centre of mass to the GROUP's bounding box (softened) is at least
// the node's critical radius |bmax-bmin|/theta + |centre-com| — the reference's robust MAC
// (ie/shaders/bh_gravity.wgsl:375-386) evaluated at the group's nearest point, so every particle of the
// group satisfies the per-particle test and the walk is never less accurate than the reference's
// per-particle walk (bh_gravity.wgsl:337-394). Leaves are never accepted: their particles are summed
// exactly (bh_gravity.wgsl:356-369). With quadrupoles, an accepted cell that also passes the test at a smaller
// angle theta_mono acts as a monopole (13 instead of 31 VALU instructions per interaction): its quadrupole term,
// of order theta_mono^2, is below the truncation error of the cells accepted at theta, of order theta^3.
// Accepted nodes and opened leaves are staged in LDS and evaluated by
// all 64 lanes against broadcast sources, so there is no divergence in the force loop.
#include "common.h"
#include "launch.h"

namespace {

constexpr int GW_WAVES = 4;
constexpr int GW_STACK = 1024;
constexpr int GW_RING = 128;

template <bool QUAD>
struct WalkLds {
  uint32_t stack[GW_STACK];
  float4 cell[GW_RING];
  float4 qa[QUAD ? GW_RING : 1];  // accepted cells' quadrupoles (xx, yy, zz, xy), (xz, yz)
  float2 qb[QUAD ? GW_RING : 1];
  float4 mono[QUAD ? GW_RING : 1];  // accepted one-particle leaves and far cells: monopoles, without the quadrupole
  uint32_t leaf[GW_RING];
  float4 part[64];
  uint32_t pidx[64];
  uint32_t phs2[64];  // with neighbour lists: the staged particles' search radius squared (float bits)
};

// A 16-byte neighbour-list chunk, stored non-temporally: the lists stream out of the walk, and outside the L2's
// normal replacement they do not evict the tree nodes that other groups still read.
__device__ __forceinline__ void store_chunk(uint4* p, uint4 v) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, reinterpret_cast<u32x4*>(p));
}

// Padding entry of the source rings: massless and far from every particle, so that 0 * 1/r stays 0 even without
// softening.
constexpr float4 kPadSource = {1.0e18f, 0.f, 0.f, 0.f};

struct Acc {
  float a[3], pot;
  Kahan3 k;
  float kpot, cpot;
};

template <bool KAHAN>
__device__ __forceinline__ void acc_add(Acc& A, float px, float py, float pz, float pp) {
  if (KAHAN) {
    A.k.add(px, 0);
    A.k.add(py, 1);
    A.k.add(pz, 2);
    const float y = pp - A.cpot;
    const float t = A.kpot + y;
    A.cpot = (t - A.kpot) - y;
    A.kpot = t;
  } else {
    A.a[0] += px;
    A.a[1] += py;
    A.a[2] += pz;
    A.pot += pp;
  }
}

template <bool POT = true>
__device__ __forceinline__ void interact(const float4 s, const float4 me, float soft2, bool self, float& ax, float& ay,
                                         float& az, float& pot) {
  const float dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
  float rinv = __builtin_amdgcn_rsqf(r2);
  rinv = self ? 0.f : rinv;
  const float mr = s.w * rinv;
  const float mr3 = mr * rinv * rinv;
  ax = fmaf(dx, mr3, ax);
  ay = fmaf(dy, mr3, ay);
  az = fmaf(dz, mr3, az);
  if (POT) pot -= mr;
}

// As interact, from the unsoftened r^2 the neighbour test has already formed (one add instead of three FMAs).
template <bool POT = true>
__device__ __forceinline__ void interact_r2(const float4 s, float dx, float dy, float dz, float r2u, float soft2,
                                            bool self, float& ax, float& ay, float& az, float& pot) {
  float rinv = __builtin_amdgcn_rsqf(r2u + soft2);
  rinv = self ? 0.f : rinv;
  const float mr = s.w * rinv;
  const float mr3 = mr * rinv * rinv;
  ax = fmaf(dx, mr3, ax);
  ay = fmaf(dy, mr3, ay);
  az = fmaf(dz, mr3, az);
  if (POT) pot -= mr;
}

// Monopole + traceless quadrupole of a cell (com, M, mass-normalised q = Q/M) at d = com - x,
// r^2 = |d|^2 + eps^2: a = M [d / r^3 - q d / r^5 + 5/2 (d.q.d) d / r^7], phi = -M [1/r + (d.q.d) / (2 r^5)]
// (G applied later). M r^-k is formed by repeated multiplication so that no intermediate leaves FP32 range.
template <bool POT>
__device__ __forceinline__ void interact_quad(const float4 s, const float4 qa, const float2 qb, const float4 me,
                                              float soft2, float& ax, float& ay, float& az, float& pot) {
  const float dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float rinv2 = rinv * rinv;
  const float mr = s.w * rinv;
  const float mr3 = mr * rinv2;
  const float mr5 = mr3 * rinv2;
  const float mr7 = mr5 * rinv2;
  const float qdx = fmaf(qa.x, dx, fmaf(qa.w, dy, qb.x * dz));
  const float qdy = fmaf(qa.w, dx, fmaf(qa.y, dy, qb.y * dz));
  const float qdz = fmaf(qb.x, dx, fmaf(qb.y, dy, qa.z * dz));
  const float dqd = fmaf(dx, qdx, fmaf(dy, qdy, dz * qdz));
  // Every product here is a power of 1/r times at most |d|^3 / r^3, so a far or padding source (mass 0 at
  // 1e18 m) yields exact zeros. Folding the r^-7 term as M r^-5 [d (r^2 + 5/2 (d.q.d) / r^2) - q d] saves one
  // multiply but forms |d| r^2, which overflows beyond ~7e12 m and turns 0 * inf into NaN.
  const float f = fmaf(2.5f * dqd, mr7, mr3);
  ax = fmaf(dx, f, fmaf(-qdx, mr5, ax));
  ay = fmaf(dy, f, fmaf(-qdy, mr5, ay));
  az = fmaf(dz, f, fmaf(-qdz, mr5, az));
  if (POT) pot -= fmaf(0.5f * dqd, mr5, mr);
}

// Walk of one group by one wave; each lane owns T targets (particles t0 + lane + 64 q, q < T), so a group holds
// up to 64 T particles, one walk serves all of them, and every source read from LDS feeds T interactions.
// SELF: compare source indices to skip the particle itself. This is needed only without softening: with
// softening the self term contributes zero force and a known potential -m/eps, which is removed at the end.
// POT: accumulate the potential. QUAD: accepted cells act with their quadrupoles too.
// NBR: also write the group's neighbour lists (one target per lane; the neighbour build's contract: every j with
// r_ij^2 < max(hs_i^2, hs_j^2), j != i, in 16-byte chunks nbr4[chunk * stride + t]). A node passing the
// neighbour search's box test is opened instead of accepted, so every candidate leaf reaches the leaf path, where
// the candidate's distance is the one its gravity term computes.
// nmac (optional): each node's (critical radius, monopole distance^2) record (k_mac_record); mac2: the opening test's
// scale, (theta / theta')^2 for a walk at theta' below the tree's theta (1: the tree's theta).
template <bool KAHAN, bool SELF, int T, bool POT, bool QUAD, bool NBR>
__device__ __forceinline__ void walk_group(WalkLds<QUAD>& S, const int4 grp, uint32_t lane, const float4* __restrict__ xm,
                                           int64_t n, int K, int64_t L, const int2* __restrict__ child,
                                           const float4* __restrict__ ncom, const float4* __restrict__ nbmax,
                                           const float4* __restrict__ nqa, const float2* __restrict__ nqb,
                                           float soft2, float4* __restrict__ accg, uint32_t* __restrict__ flags,
                                           uint64_t& n_cell, uint64_t& n_pp, uint64_t& n_nodes, const NbrIo& io,
                                           uint64_t& ncand, uint64_t& npairs, const float2* __restrict__ nmac,
                                           float mac2) {
  static_assert(!NBR || T == 1, "neighbour lists are written by walks with one target per lane");
  const int64_t t0 = grp.x;
  int64_t t[T];
  bool valid[T];
  float4 me[T];
  float bx0 = 3.0e38f, by0 = 3.0e38f, bz0 = 3.0e38f, bx1 = -3.0e38f, by1 = -3.0e38f, bz1 = -3.0e38f;
  uint32_t nv = 0;
#pragma unroll
  for (int q = 0; q < T; ++q) {
    t[q] = t0 + lane + 64 * q;
    valid[q] = lane + 64u * q < (uint32_t)grp.y;
    me[q] = xm[valid[q] ? t[q] : t0];  // invalid slots repeat the first particle: the group box is unchanged
    bx0 = fminf(bx0, me[q].x); by0 = fminf(by0, me[q].y); bz0 = fminf(bz0, me[q].z);
    bx1 = fmaxf(bx1, me[q].x); by1 = fmaxf(by1, me[q].y); bz1 = fmaxf(bz1, me[q].z);
    nv += valid[q] ? 1u : 0u;
  }
  // The group box is wave-uniform: held in scalar registers, it frees vector registers for the walk.
  bx0 = sgpr_f(wave_min(bx0)); by0 = sgpr_f(wave_min(by0)); bz0 = sgpr_f(wave_min(bz0));
  bx1 = sgpr_f(wave_max(bx1)); by1 = sgpr_f(wave_max(by1)); bz1 = sgpr_f(wave_max(bz1));
  const uint64_t nvalid = (uint64_t)wave_sum((unsigned long long)nv);

  Acc A[T];
#pragma unroll
  for (int q = 0; q < T; ++q) {
    A[q].a[0] = A[q].a[1] = A[q].a[2] = A[q].pot = 0.f;
    A[q].k.zero();
    A[q].kpot = A[q].cpot = 0.f;
  }
  // Neighbour-list state (NBR): this lane's search radius, the group's largest, the 4-entry staging register.
  const float myhs = NBR && valid[0] ? io.hs[t[0]] : 0.f;
  const uint32_t myhs2u = __float_as_uint(myhs * myhs);
  const float ghs = NBR ? sgpr_f(wave_max(myhs)) : 0.f;
  uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, nb = 0, nchunk = 0, cntn = 0;
  const uint32_t cap4 = (uint32_t)io.cap >> 2;

  const uint32_t leaf0 = (uint32_t)(L - 1);
  const uint32_t lpc = 64u / (uint32_t)K;
  uint32_t sp = 1, chead = 0, ctail = 0, lhead = 0, ltail = 0, mhead = 0, mtail = 0;
  // With one particle per leaf a leaf's quadrupole is zero: accepted leaves go to a monopole ring.
  const bool mono_leaves = QUAD && K == 1;
  wave_sync();
  if (lane == 0) S.stack[0] = (uint32_t)grp.w;  // the group's segment root
  wave_sync();

  // Accepted nodes: chead stays a multiple of 64 until the final flush, so a chunk is contiguous in the
  // 128-entry ring; cnt is a multiple of 4 (the tail is padded with massless entries).
  auto eval_cells = [&](uint32_t cnt, uint32_t real) {
    const uint32_t h0 = chead & (GW_RING - 1);
    const float4* src = &S.cell[h0];
    const float4* sqa = &S.qa[h0];
    const float2* sqb = &S.qb[h0];
    float ax[T], ay[T], az[T], pp[T];
#pragma unroll
    for (int q = 0; q < T; ++q) ax[q] = ay[q] = az[q] = pp[q] = 0.f;
    for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        const float4 c = src[k0 + e];
        if (QUAD) {
          const float4 qa = sqa[k0 + e];
          const float2 qb = sqb[k0 + e];
#pragma unroll
          for (int q = 0; q < T; ++q) interact_quad<POT>(c, qa, qb, me[q], soft2, ax[q], ay[q], az[q], pp[q]);
        } else {
#pragma unroll
          for (int q = 0; q < T; ++q) interact<POT>(c, me[q], soft2, false, ax[q], ay[q], az[q], pp[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < T; ++q) acc_add<KAHAN>(A[q], ax[q], ay[q], az[q], pp[q]);
    chead += cnt;
    n_cell += nvalid * real;
  };
  auto eval_mono = [&](uint32_t cnt, uint32_t real) {
    const float4* src = &S.mono[mhead & (GW_RING - 1)];
    float ax[T], ay[T], az[T], pp[T];
#pragma unroll
    for (int q = 0; q < T; ++q) ax[q] = ay[q] = az[q] = pp[q] = 0.f;
    for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        const float4 c = src[k0 + e];
#pragma unroll
        for (int q = 0; q < T; ++q) interact<POT>(c, me[q], soft2, false, ax[q], ay[q], az[q], pp[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < T; ++q) acc_add<KAHAN>(A[q], ax[q], ay[q], az[q], pp[q]);
    mhead += cnt;
    n_pp += nvalid * real;
  };
  auto eval_leaves = [&](uint32_t nl) {
    const uint32_t li = lane / (uint32_t)K;
    const bool have = li < nl;
    const uint32_t leaf = have ? S.leaf[(lhead + li) & (GW_RING - 1)] : 0u;
    const int64_t src = (int64_t)leaf * K + (lane % (uint32_t)K);
    const bool ok = have && src < n;
    wave_sync();
    S.part[lane] = ok ? xm[src] : kPadSource;
    if (SELF || NBR) S.pidx[lane] = ok ? (uint32_t)src : IHS_NONE;
    if (NBR) {
      const float ph = ok ? io.hs[src] : 0.f;
      S.phs2[lane] = __float_as_uint(ph * ph);
    }
    wave_sync();
    float ax[T], ay[T], az[T], pp[T];
#pragma unroll
    for (int q = 0; q < T; ++q) ax[q] = ay[q] = az[q] = pp[q] = 0.f;
    const uint32_t cnt = (min(64u, nl * (uint32_t)K) + 3u) & ~3u;  // padding slots are massless
    if (NBR) {
      // The gravity term and the neighbour test of each staged particle; candidate k's hit is bit k of a per-lane
      // mask (the loop is unrolled so the bit is a constant), and only the hits are visited afterwards.
      uint32_t hlo = 0, hhi = 0;
#pragma unroll
      for (uint32_t k0 = 0; k0 < 64; k0 += 4) {
        if (k0 < cnt) {
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t k = k0 + e;
            const float4 c = S.part[k];
            const float dx = c.x - me[0].x, dy = c.y - me[0].y, dz = c.z - me[0].z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            interact_r2<POT>(c, dx, dy, dz, r2, soft2, SELF && S.pidx[k] == (uint32_t)t[0], ax[0], ay[0], az[0],
                             pp[0]);
            // r < max(hs_i, hs_j): non-negative floats order as their bit patterns.
            const bool hit = __float_as_uint(r2) < max(myhs2u, S.phs2[k]);
            if (k < 32) hlo |= hit ? (1u << k) : 0u;
            else hhi |= hit ? (1u << (k - 32)) : 0u;
          }
        }
      }
      acc_add<KAHAN>(A[0], ax[0], ay[0], az[0], pp[0]);
      ncand += nvalid * cnt;
      uint64_t hits = valid[0] ? ((uint64_t)hhi << 32) | hlo : 0ull;
      const uint32_t tt = (uint32_t)t[0];
      while (hits) {
        const uint32_t k = (uint32_t)__builtin_ctzll(hits);
        hits &= hits - 1;
        const uint32_t j = S.pidx[k];
        if (j == tt) continue;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = j;
        ++cntn;
        if (++nb == 4) {
          if (nchunk < cap4) store_chunk(io.nbr4 + (int64_t)nchunk * io.stride + t[0], make_uint4(b0, b1, b2, b3));
          ++nchunk;
          nb = 0;
        }
      }
    } else {
      for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
          const float4 c = S.part[k0 + e];
#pragma unroll
          for (int q = 0; q < T; ++q)
            interact<POT>(c, me[q], soft2, SELF && S.pidx[k0 + e] == (uint32_t)t[q], ax[q], ay[q], az[q], pp[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < T; ++q) acc_add<KAHAN>(A[q], ax[q], ay[q], az[q], pp[q]);
    }
    lhead += nl;
    n_pp += nvalid * (uint64_t)__popcll(__ballot(ok));
    wave_sync();
  };

  while (sp > 0) {
    const uint32_t take = min(sp, 64u);
    const uint32_t base = sp - take;
    const uint32_t node = lane < take ? S.stack[base + lane] : IHS_NONE;
    sp = base;
    n_nodes += take;
    wave_sync();
    // Every node, leaves included, is tested with the MAC (the reference always sums leaves exactly,
    // bh_gravity.wgsl:356-369; with buckets of K particles a far leaf is the same monopole test as any node).
    const bool live = node != IHS_NONE;
    bool acc = false, far = false;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
      c = ncom[node];
      float rc, dfar = 3.0e38f;
      if (nmac) {
        const float2 m = nmac[node];
        rc = m.x;
        dfar = m.y;
      } else {
        rc = nbmax[node].w;
      }
      const float dx = fmaxf(fmaxf(bx0 - c.x, c.x - bx1), 0.f);
      const float dy = fmaxf(fmaxf(by0 - c.y, c.y - by1), 0.f);
      const float dz = fmaxf(fmaxf(bz0 - c.z, c.z - bz1), 0.f);
      const float d2 = dx * dx + dy * dy + dz * dz + soft2;
      // Without softening a leaf touching the group box may hold a target itself: it takes the
      // particle-particle path, which skips the self pair.
      acc = d2 >= rc * rc * mac2 && (!SELF || node < leaf0 || d2 > 0.f);
      // The monopole distance scales with the opening test: a walk at theta' (mac2 = (theta / theta')^2) takes its
      // monopole cells at theta_mono theta' / theta.
      far = d2 >= dfar * mac2;
      if (NBR && acc) {
        // The neighbour build's box test (group box against the node box grown by the larger search radius): a
        // node that passes is opened, and its leaves come to the leaf path.
        const float4 c0 = io.nbmin[node], c1 = nbmax[node];
        const float ex = fmaxf(fmaxf(c0.x - bx1, bx0 - c1.x), 0.f);
        const float ey = fmaxf(fmaxf(c0.y - by1, by0 - c1.y), 0.f);
        const float ez = fmaxf(fmaxf(c0.z - bz1, bz0 - c1.z), 0.f);
        const float R = fmaxf(ghs, c0.w);
        if (ex * ex + ey * ey + ez * ez < R * R) acc = false;
      }
    }
    const bool isleaf = live && !acc && node >= leaf0;
    const bool open = live && !acc && node < leaf0;
    // Monopole ring: one-particle leaves (zero quadrupole) and cells far enough that their quadrupole term is below
    // the error of the cells accepted at the edge of the test.
    const bool accm = acc && QUAD && ((mono_leaves && node >= leaf0) || far);
    acc = acc && !accm;
    const uint64_t ma = __ballot(acc), mo = __ballot(open), ml = __ballot(isleaf), mm = __ballot(accm);
    if (accm) S.mono[(mtail + lanes_below(mm)) & (GW_RING - 1)] = c;
    mtail += (uint32_t)__popcll(mm);
    if (acc) {
      const uint32_t slot = (ctail + lanes_below(ma)) & (GW_RING - 1);
      S.cell[slot] = c;
      if (QUAD) {
        S.qa[slot] = nqa[node];
        S.qb[slot] = nqb[node];
      }
    }
    ctail += (uint32_t)__popcll(ma);
    if (isleaf) S.leaf[(ltail + lanes_below(ml)) & (GW_RING - 1)] = node - leaf0;
    ltail += (uint32_t)__popcll(ml);
    const uint32_t nopen = (uint32_t)__popcll(mo);
    if (sp + 2u * nopen > (uint32_t)GW_STACK) {
      if (lane == 0) atomicAdd(&flags[FLAG_WALK_OVERFLOW], 1u);
    } else {
      if (open) {
        const int2 ch = child[node];
        const uint32_t pos = sp + 2u * lanes_below(mo);
        S.stack[pos] = (uint32_t)ch.x;
        S.stack[pos + 1] = (uint32_t)ch.y;
      }
      sp += 2u * nopen;
    }
    wave_sync();
    while (ctail - chead >= 64u) eval_cells(64u, 64u);
    while (mtail - mhead >= 64u) eval_mono(64u, 64u);
    while (ltail - lhead >= lpc) eval_leaves(lpc);
  }
  if (mtail != mhead) {
    const uint32_t real = mtail - mhead;
    const uint32_t pad = (4u - (real & 3u)) & 3u;
    if (lane < pad) S.mono[(mtail + lane) & (GW_RING - 1)] = kPadSource;
    wave_sync();
    mtail += pad;
    eval_mono(mtail - mhead, real);
  }
  if (ctail != chead) {
    const uint32_t real = ctail - chead;
    const uint32_t pad = (4u - (real & 3u)) & 3u;
    if (lane < pad) {
      const uint32_t slot = (ctail + lane) & (GW_RING - 1);
      S.cell[slot] = kPadSource;
      if (QUAD) {
        S.qa[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        S.qb[slot] = make_float2(0.f, 0.f);
      }
    }
    wave_sync();
    ctail += pad;
    eval_cells(ctail - chead, real);
  }
  while (ltail != lhead) eval_leaves(min(lpc, ltail - lhead));

  if (NBR) {
    // A partial last chunk holds its nb entries in b[4-nb..3]: move them to the front, padding with the last.
    if (valid[0] && nb > 0 && nchunk < cap4)
      store_chunk(io.nbr4 + (int64_t)nchunk * io.stride + t[0],
                  nb == 1 ? make_uint4(b3, b3, b3, b3) : nb == 2 ? make_uint4(b2, b3, b3, b3) : make_uint4(b1, b2, b3, b3));
    const uint32_t capn = cap4 * 4u;
    if (valid[0]) io.ncount[t[0]] = min(cntn, capn);
    const uint64_t over = __ballot(valid[0] && cntn > capn);
    npairs += wave_sum((unsigned long long)(valid[0] ? min(cntn, capn) : 0u));
    if (lane == 0 && over) atomicAdd(&flags[FLAG_NBR_OVERFLOW], (uint32_t)__popcll(over));
  }

#pragma unroll
  for (int q = 0; q < T; ++q) {
    if (!valid[q]) continue;
    float a0, a1, a2, p;
    if (KAHAN) {
      a0 = A[q].k.s[0]; a1 = A[q].k.s[1]; a2 = A[q].k.s[2]; p = A[q].kpot;
    } else {
      a0 = A[q].a[0]; a1 = A[q].a[1]; a2 = A[q].a[2]; p = A[q].pot;
    }
    if (!POT) p = 0.f;
    else if (!SELF) p += me[q].w * __builtin_amdgcn_rsqf(soft2);  // remove the self term -m/eps
    accg[t[q]] = make_float4(kG32 * a0, kG32 * a1, kG32 * a2, kG32 * p);
  }
}

// Persistent walk: every wave pulls groups from its XCD's queue (a contiguous eighth of the groups, so that
// XCD's L2 serves neighbouring groups), then helps the other queues; irregular groups therefore overlap
// instead of leaving the machine idle behind the slowest one.
template <bool KAHAN, bool SELF, int T, bool POT, bool QUAD, bool NBR>
__global__ __launch_bounds__(GW_WAVES * 64) void k_gravity_walk(const float4* __restrict__ xm, int64_t n, int K,
                                                                 int64_t L, const int2* __restrict__ child,
                                                                 const float4* __restrict__ ncom,
                                                                 const float4* __restrict__ nbmax,
                                                                 const float4* __restrict__ nqa,
                                                                 const float2* __restrict__ nqb,
                                                                 const int4* __restrict__ groups,
                                                                 const uint32_t* __restrict__ ngroups,
                                                                 uint32_t* __restrict__ wq, SegView sv,
                                                                 float4* __restrict__ accg, uint32_t* __restrict__ flags,
                                                                 unsigned long long* __restrict__ counters, NbrIo io,
                                                                 const float2* __restrict__ nmac, float mac2) {
  __shared__ WalkLds<QUAD> lds[GW_WAVES];
  const int w = threadIdx.x >> 6;
  const uint32_t lane = threadIdx.x & 63;
  WalkLds<QUAD>& S = lds[w];
  const uint32_t ng = *ngroups;
  const uint32_t home = blockIdx.x & 7u;
  uint64_t n_cell = 0, n_pp = 0, n_nodes = 0, ncand = 0, npairs = 0;
  for (uint32_t r = 0; r < 8; ++r) {
    const uint32_t q = (home + r) & 7u;
    const uint32_t lo = (uint32_t)((uint64_t)ng * q / 8u), hi = (uint32_t)((uint64_t)ng * (q + 1u) / 8u);
    for (;;) {
      uint32_t g = 0;
      if (lane == 0) g = atomicAdd(&wq[q], 1u);
      g = (uint32_t)__builtin_amdgcn_readfirstlane((int)g) + lo;
      if (g >= hi) break;
      const int4 grp = groups[g];
      const float soft2 = sv.soft2[sv.of(grp.x)];
      walk_group<KAHAN, SELF, T, POT, QUAD, NBR>(S, grp, lane, xm, n, K, L, child, ncom, nbmax, nqa, nqb, soft2,
                                                 accg, flags, n_cell, n_pp, n_nodes, io, ncand, npairs, nmac, mac2);
    }
  }
  if (lane == 0) {
    cnt_add(counters, CNT_GRAV_CELL, n_cell);
    cnt_add(counters, CNT_GRAV_PP, n_pp);
    cnt_add(counters, CNT_WALK_NODES, n_nodes);
    if (NBR) {
      cnt_add(counters, CNT_PAIRS, npairs);
      cnt_add(counters, CNT_NBR_CAND, ncand);
    }
  }
}

// Tiled exact FP32 sum: every thread one target, sources staged through LDS 256 at a time.
__global__ __launch_bounds__(256) void k_gravity_direct(const float4* __restrict__ xm, int64_t n, float soft2,
                                                        float4* __restrict__ acc4) {
  __shared__ float4 tile[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const float4 me = xm[i < n ? i : 0];
  float ax = 0.f, ay = 0.f, az = 0.f, pp = 0.f;
  for (int64_t j0 = 0; j0 < n; j0 += 256) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    tile[threadIdx.x] = j < n ? xm[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int cnt = (int)min<int64_t>(256, n - j0);
    for (int k = 0; k < cnt; ++k) interact(tile[k], me, soft2, j0 + k == i, ax, ay, az, pp);
  }
  if (i < n) acc4[i] = make_float4(kG32 * ax, kG32 * ay, kG32 * az, kG32 * pp);
}

// Standalone gravity (the GravityField backend): AoS doubles in, SoA out.
__global__ __launch_bounds__(256) void k_bodies_in(const double* __restrict__ pos3, const double* __restrict__ mass,
                                                   int64_t n, double* __restrict__ px, double* __restrict__ py,
                                                   double* __restrict__ pz, float* __restrict__ m) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  px[i] = pos3[3 * i];
  py[i] = pos3[3 * i + 1];
  pz[i] = pos3[3 * i + 2];
  m[i] = (float)mass[i];
}

__global__ __launch_bounds__(256) void k_gather_xm(const uint32_t* __restrict__ perm, int64_t n,
                                                   const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, const float* __restrict__ m,
                                                   float4* __restrict__ xm) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const uint32_t s = perm ? perm[t] : (uint32_t)t;
  xm[t] = make_float4((float)px[s], (float)py[s], (float)pz[s], m[s]);
}

__global__ __launch_bounds__(256) void k_acc_out(const uint32_t* __restrict__ perm, int64_t n,
                                                 const float4* __restrict__ acc4, double* __restrict__ acc3,
                                                 double* __restrict__ pot) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int64_t i = perm ? (int64_t)perm[t] : t;
  const float4 a = acc4[t];
  acc3[3 * i] = a.x;
  acc3[3 * i + 1] = a.y;
  acc3[3 * i + 2] = a.z;
  if (pot) pot[i] = a.w;
}

// FP64 direct sum on selected targets (the accuracy probe of the walk): DT_TILE targets per block, every
// thread strides over the sources and the block reduces in LDS. Positions in storage order (PT = float or
// double); tpos = storage index of each target.
constexpr int DT_TILE = 8;

template <class PT>
__global__ __launch_bounds__(256) void k_direct_f64(const PT* __restrict__ px, const PT* __restrict__ py,
                                                    const PT* __restrict__ pz, const float* __restrict__ m, int64_t n,
                                                    const uint32_t* __restrict__ tpos, int64_t nt,
                                                    const double* __restrict__ seg_soft2,
                                                    const uint32_t* __restrict__ segof,
                                                    const uint32_t* __restrict__ seg_start, double* __restrict__ acc3) {
  __shared__ double red[256];
  const int64_t t0 = (int64_t)blockIdx.x * DT_TILE;
  double tx[DT_TILE], ty[DT_TILE], tz[DT_TILE], ax[DT_TILE], ay[DT_TILE], az[DT_TILE], s2[DT_TILE];
  int64_t ti[DT_TILE], lo[DT_TILE], hi[DT_TILE];
  int64_t jlo = n, jhi = 0;
#pragma unroll
  for (int k = 0; k < DT_TILE; ++k) {
    const int64_t q = t0 + k < nt ? (int64_t)tpos[t0 + k] : 0;
    const uint32_t sg = segof ? segof[q] : 0u;  // sources: the target's own scenario
    ti[k] = q;
    tx[k] = (double)px[q];
    ty[k] = (double)py[q];
    tz[k] = (double)pz[q];
    s2[k] = (double)seg_soft2[sg];
    lo[k] = seg_start[sg];
    hi[k] = seg_start[sg + 1];
    jlo = min(jlo, lo[k]);
    jhi = max(jhi, hi[k]);
    ax[k] = ay[k] = az[k] = 0.0;
  }
  for (int64_t j = jlo + threadIdx.x; j < jhi; j += blockDim.x) {
    const double xj = (double)px[j], yj = (double)py[j], zj = (double)pz[j], mj = (double)m[j];
#pragma unroll
    for (int k = 0; k < DT_TILE; ++k) {
      if (j == ti[k] || j < lo[k] || j >= hi[k]) continue;
      const double dx = xj - tx[k], dy = yj - ty[k], dz = zj - tz[k];
      const double r2 = dx * dx + dy * dy + dz * dz + s2[k];
      const double rinv = 1.0 / sqrt(r2);
      const double f = mj * rinv * rinv * rinv;
      ax[k] += f * dx;
      ay[k] += f * dy;
      az[k] += f * dz;
    }
  }
  for (int k = 0; k < DT_TILE; ++k) {
    for (int c = 0; c < 3; ++c) {
      red[threadIdx.x] = c == 0 ? ax[k] : (c == 1 ? ay[k] : az[k]);
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0 && t0 + k < nt) acc3[3 * (t0 + k) + c] = kG64 * red[0];
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_storage_index(const uint32_t* __restrict__ id, int64_t n,
                                                       uint32_t* __restrict__ inv) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) inv[id[t]] = (uint32_t)t;
}

__global__ __launch_bounds__(256) void k_map_ids(const uint32_t* __restrict__ inv, const uint32_t* __restrict__ ids,
                                                 int64_t nt, uint32_t* __restrict__ tpos) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < nt) tpos[k] = inv[ids[k]];
}

}  // namespace

void launch_bodies_in(hipStream_t s, const double* pos3, const double* mass, int64_t n, double* px, double* py,
                      double* pz, float* m) {
  k_bodies_in<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos3, mass, n, px, py, pz, m);
}
void launch_gather_xm(hipStream_t s, const uint32_t* perm, int64_t n, const double* px, const double* py,
                      const double* pz, const float* m, float4* xm) {
  k_gather_xm<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(perm, n, px, py, pz, m, xm);
}
void launch_acc_out(hipStream_t s, const uint32_t* perm, int64_t n, const float4* acc4, double* acc3, double* pot) {
  k_acc_out<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(perm, n, acc4, acc3, pot);
}

// Each node's opening record for walks that let far or weak-quadrupole cells act as monopoles: (critical radius,
// smallest squared distance from a walk group's box at which the accepted cell is a monopole), one 8-byte load at
// the opening test in place of the critical radius alone. The distance: the opening test at theta_mono,
// d >= rc theta / theta_mono (far2 = (theta / theta_mono)^2; rc = b / theta + delta, so this is at least
// b / theta_mono + delta), or a quadrupole term below tol times the monopole term: with d = com - x and r >= |d|,
// |a_quad| = M |-q d / r^5 + 5/2 (d.q.d) d / r^7| <= 3.5 |q|_F M / r^4, a fraction 3.5 |q|_F / r^2 of the monopole's
// M / r^2, so d^2 >= 3.5 |q|_F / tol. One-particle leaves (K = 1) store no quadrupole and are monopoles.
__global__ __launch_bounds__(256) void k_mac_record(int64_t nodes, int64_t leaf0, int K, const float4* __restrict__ nbmax,
                                                    const float4* __restrict__ nqa, const float2* __restrict__ nqb,
                                                    float far2, float qscale, float2* __restrict__ nmac) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nodes) return;
  const float rc = nbmax[x].w;
  float d2 = 0.f;
  if (K != 1 || x < leaf0) {
    d2 = far2 > 0.f ? rc * rc * far2 : 3.0e38f;
    if (qscale > 0.f) {
      const float4 qa = nqa[x];
      const float2 qb = nqb[x];
      const float qf = __builtin_sqrtf(qa.x * qa.x + qa.y * qa.y + qa.z * qa.z +
                                       2.f * (qa.w * qa.w + qb.x * qb.x + qb.y * qb.y));
      d2 = fminf(d2, qf * qscale);
    }
  }
  nmac[x] = make_float2(rc, d2);
}

void launch_mac_record(hipStream_t s, int64_t n, int K, float theta, const TreeBufs& tb) {
  if (!tb.nmac || !tb.quadrupole) return;
  const int64_t L = (n + K - 1) / K;
  const float r = tb.theta_mono > 0.f && theta > 0.f ? theta / tb.theta_mono : 0.f;
  const float qscale = tb.quad_tol > 0.f ? 3.5f / tb.quad_tol : 0.f;
  const int64_t nodes = 2 * L - 1;
  k_mac_record<<<(unsigned)((nodes + 255) / 256), 256, 0, s>>>(nodes, L - 1, K, tb.nbmax, tb.nqa, tb.nqb, r * r,
                                                                qscale, tb.nmac);
}

void launch_gravity_walk(hipStream_t s, int64_t n, int K, int kahan, const float4* xm, const TreeBufs& tb,
                         const int4* groups, const uint32_t* ngroups, const SegView& sv, float4* accg, uint32_t* flags,
                         unsigned long long* counters, bool pot, bool self, const NbrIo* nbr, float mac2) {
  const int64_t L = (n + K - 1) / K;
  const bool two = tb.group_max > GROUP_MAX;
  const bool quad = tb.quadrupole != 0;
  const float2* nmac = quad ? tb.nmac : nullptr;
  const bool lists = nbr != nullptr && !two;
  const NbrIo io = lists ? *nbr : NbrIo{};
  static unsigned nblk[64] = {0};
  const int variant = (lists ? 32 : 0) + (quad ? 16 : 0) + (kahan ? 8 : 0) + (self ? 4 : 0) + (two ? 2 : 0) + (pot ? 1 : 0);
#define IHS_WALK(KA, SE, T, PO, QU, NB)                                                                         \
  do {                                                                                                          \
    if (!nblk[variant])                                                                                         \
      nblk[variant] = persistent_blocks((const void*)k_gravity_walk<KA, SE, T, PO, QU, NB>, GW_WAVES * 64);     \
    k_gravity_walk<KA, SE, T, PO, QU, NB><<<nblk[variant], GW_WAVES * 64, 0, s>>>(                              \
        xm, n, K, L, tb.child, tb.ncom, tb.nbmax, tb.nqa, tb.nqb, groups, ngroups, tb.wq, sv, accg,               \
        flags, counters, io, nmac, mac2);                                                                       \
  } while (0)
#define IHS_WALK_P(KA, SE, T, QU, NB)           \
  do {                                          \
    if (pot) IHS_WALK(KA, SE, T, true, QU, NB); \
    else IHS_WALK(KA, SE, T, false, QU, NB);    \
  } while (0)
#define IHS_WALK_T(KA, SE, QU)                     \
  do {                                             \
    if (two) IHS_WALK_P(KA, SE, 2, QU, false);     \
    else if (lists) IHS_WALK_P(KA, SE, 1, QU, true); \
    else IHS_WALK_P(KA, SE, 1, QU, false);         \
  } while (0)
#define IHS_WALK_Q(KA, SE)                       \
  do {                                           \
    if (quad) IHS_WALK_T(KA, SE, true);          \
    else IHS_WALK_T(KA, SE, false);              \
  } while (0)
  if (kahan) {
    if (self) IHS_WALK_Q(true, true);
    else IHS_WALK_Q(true, false);
  } else {
    if (self) IHS_WALK_Q(false, true);
    else IHS_WALK_Q(false, false);
  }
#undef IHS_WALK_Q
#undef IHS_WALK_T
#undef IHS_WALK_P
#undef IHS_WALK
}

void launch_gravity_direct(hipStream_t s, int64_t n, float soft2, const float4* xm, float4* acc4) {
  k_gravity_direct<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(xm, n, soft2, acc4);
}

void launch_direct_f64_targets(hipStream_t s, int split, const void* px, const void* py, const void* pz, const float* m,
                               const uint32_t* id, int64_t n, const uint32_t* ids, int64_t nt, const double* seg_soft2,
                               const uint32_t* segof, const uint32_t* seg_start, uint32_t* inv, uint32_t* tpos,
                               double* acc3) {
  k_storage_index<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(id, n, inv);
  k_map_ids<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(inv, ids, nt, tpos);
  const unsigned nb = (unsigned)((nt + DT_TILE - 1) / DT_TILE);
  if (split)
    k_direct_f64<double><<<nb, 256, 0, s>>>((const double*)px, (const double*)py, (const double*)pz, m, n, tpos, nt,
                                            seg_soft2, segof, seg_start, acc3);
  else
    k_direct_f64<float><<<nb, 256, 0, s>>>((const float*)px, (const float*)py, (const float*)pz, m, n, tpos, nt,
                                           seg_soft2, segof, seg_start, acc3);
}
