// Morton keys, LDS-staged stable radix sort, Karras LBVH build and ordered multipole upsweep (gfx950).
//
// Bit-exactness contract with the reference's CPU Morton leg (ie/crates/engine/src/gpu_gravity.rs:334-368):
//   - bbox: exact float min/max through order-preserving u32 keys (ie/shaders/bh_gravity.wgsl:49-78);
//   - Morton: IEEE f32 (p-lo)/ext, *1024, floor, clamp; expand_bits (bh_gravity.wgsl:82-98);
//   - order: lexicographic (code, particle id) (gpu_gravity.rs:358-363): stable LSD radix sort, then
//     runs of equal codes reordered by id;
//   - Karras topology: determineRange/findSplit with the index tie-break (bh_gravity.wgsl:114-184).
// This file is compiled without fast-math so division and floor round exactly as IEEE requires.
#include "common.h"
#include "launch.h"

namespace {

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v &= 0x000003ffu;
  v = (v * 0x00010001u) & 0xff0000ffu;
  v = (v * 0x00000101u) & 0x0f00f00fu;
  v = (v * 0x00000011u) & 0xc30c30c3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// bps blocks per segment; block b of segment s strides over the segment's storage range and writes its partial box.
// Block 0 also clears the evaluation's flags and per-scenario Courant signals, and the h-solve's pending-list lengths
// (fe, when given). Folding the partials in the block that finishes last (1,024 blocks, each with a device-scope
// release fence and an atomic on one counter) instead of k_bbox_final made the sort stage 0.135 -> 0.199 ms.
template <class PT>
__global__ __launch_bounds__(256) void k_bbox(const PT* __restrict__ px, const PT* __restrict__ py,
                                              const PT* __restrict__ pz, const uint32_t* __restrict__ start, int bps,
                                              uint32_t* __restrict__ part, int S, EvalInit fe) {
  if (blockIdx.x == 0) eval_init(fe, S);
  const int sg = blockIdx.x / bps, b = blockIdx.x % bps;
  const int64_t i0 = start[sg], i1 = start[sg + 1];
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  for (int64_t i = i0 + (int64_t)b * blockDim.x + threadIdx.x; i < i1; i += (int64_t)bps * blockDim.x) {
    const uint32_t k[3] = {f2key((float)px[i]), f2key((float)py[i]), f2key((float)pz[i])};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      lo[q] = min(lo[q], k[q]);
      hi[q] = max(hi[q], k[q]);
    }
  }
  __shared__ uint32_t s[4][6];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lo[q] = wave_min(lo[q]);
    hi[q] = wave_max(hi[q]);
  }
  if (lane == 0)
    for (int q = 0; q < 3; ++q) {
      s[w][q] = lo[q];
      s[w][3 + q] = hi[q];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    uint32_t v = s[0][q];
    for (int ww = 1; ww < (int)(blockDim.x >> 6); ++ww) v = q < 3 ? min(v, s[ww][q]) : max(v, s[ww][q]);
    part[6 * blockIdx.x + q] = v;
  }
}

// Folds the blocks' partial boxes into each segment's box (one workgroup; block b of k_bbox is segment b / bps).
// Same-address atomics from every block of k_bbox would serialize in L2; here they are LDS atomics, and a wave whose
// lanes share a segment folds its lanes first.
__global__ __launch_bounds__(1024) void k_bbox_final(const uint32_t* __restrict__ part, int nbk, int bps, int S,
                                                     uint32_t* __restrict__ bboxes) {
  __shared__ uint32_t sb[6 * IHS_MAX_SEGMENTS];
  for (int i = threadIdx.x; i < 6 * S; i += blockDim.x) sb[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  for (int t0 = 0; t0 < nbk; t0 += blockDim.x) {
    const int t = t0 + threadIdx.x;
    const bool ok = t < nbk;
    const int sg = ok ? t / bps : 0;
    const int sg0 = __shfl(sg, 0, 64);
    const bool uniform = __all(!ok || sg == sg0);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      uint32_t v = ok ? part[6 * t + q] : (q < 3 ? 0xffffffffu : 0u);
      if (uniform) {
        v = q < 3 ? wave_min(v) : wave_max(v);
        if ((threadIdx.x & 63) == 0) {
          if (q < 3) atomicMin(&sb[6 * sg0 + q], v);
          else atomicMax(&sb[6 * sg0 + q], v);
        }
      } else if (ok) {
        if (q < 3) atomicMin(&sb[6 * sg + q], v);
        else atomicMax(&sb[6 * sg + q], v);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 6 * S; i += blockDim.x) bboxes[i] = sb[i];
}

// Also clears the upsweep block-root marks of the L leaves (k_karras sets those of this evaluation's tree).
template <class PT>
__global__ __launch_bounds__(256) void k_morton(const PT* __restrict__ px, const PT* __restrict__ py,
                                                const PT* __restrict__ pz, int64_t n, const uint32_t* __restrict__ bboxes,
                                                const uint32_t* __restrict__ segof, uint32_t* __restrict__ codes,
                                                uint32_t* __restrict__ vals, uint32_t* __restrict__ codes0,
                                                uint32_t* __restrict__ bstart, int64_t L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L) bstart[i] = IHS_NONE;
  if (i >= n) return;
  const uint32_t* bbox = bboxes + 6 * (segof ? segof[i] : 0u);
  float lo[3], ext[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lo[q] = key2f(bbox[q]);
    ext[q] = fmaxf(key2f(bbox[3 + q]) - lo[q], 1.0e-30f);
  }
  const float p[3] = {(float)px[i], (float)py[i], (float)pz[i]};
  uint32_t qq[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float u = __fdiv_rn(__fsub_rn(p[q], lo[q]), ext[q]);
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    float f = floorf(__fmul_rn(u, 1024.0f));
    f = fminf(fmaxf(f, 0.0f), 1023.0f);
    qq[q] = (uint32_t)f;
  }
  const uint32_t code = expand_bits(qq[0]) * 4u + expand_bits(qq[1]) * 2u + expand_bits(qq[2]);
  codes[i] = code;
  vals[i] = (uint32_t)i;
  if (codes0) codes0[i] = code;
}

// Segment pass of a batch: keys = segment of the particle at each sorted position, payload = the permutation.
__global__ __launch_bounds__(256) void k_seg_keys(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ segof,
                                                  int64_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  keys[i] = segof[p];
  vals[i] = p;
}

__global__ __launch_bounds__(256) void k_gather_codes(const uint32_t* __restrict__ perm,
                                                      const uint32_t* __restrict__ codes0, int64_t n,
                                                      uint32_t* __restrict__ skeys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) skeys[i] = codes0[perm[i]];
}

// ---------------------------------------------------------------------------------------------
// Radix sort, 8-bit digits. Tile t covers keys [t*TILE, (t+1)*TILE); item r of thread x is
// t*TILE + r*256 + x, so (r, wave, lane) order is the key order inside the tile (stability).
__global__ __launch_bounds__(RADIX_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                              uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
#pragma unroll
  for (int r = 0; r < RADIX_ITEMS; ++r) {
    const int64_t i = base + r * RADIX_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan over the nb tile counts of that digit; dtot[d] = digit total.
__global__ __launch_bounds__(256) void k_radix_scan(uint32_t* __restrict__ hist, int nb, uint32_t* __restrict__ dtot) {
  __shared__ uint32_t s_tmp[4];
  const int d = blockIdx.x;
  uint32_t carry = 0;
  for (int c0 = 0; c0 < nb; c0 += 256) {
    const int c = c0 + threadIdx.x;
    const uint32_t v = c < nb ? hist[(int64_t)d * nb + c] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan_256(v, s_tmp, &tot);
    if (c < nb) hist[(int64_t)d * nb + c] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) dtot[d] = carry;
}

__global__ __launch_bounds__(RADIX_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                 const uint32_t* __restrict__ vin,
                                                                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                                 int64_t n, int shift, const uint32_t* __restrict__ hist,
                                                                 const uint32_t* __restrict__ dtot, int nb) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wc[4][256];
  __shared__ uint32_t s_tmp[4];
  const int tid = threadIdx.x, w = tid >> 6;
  {
    const uint32_t dbase = block_excl_scan_256(dtot[tid], s_tmp, nullptr);
    s_base[tid] = dbase + hist[(int64_t)tid * nb + blockIdx.x];
    for (int ww = 0; ww < 4; ++ww) s_wc[ww][tid] = 0;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
  for (int r = 0; r < RADIX_ITEMS; ++r) {
    const int64_t i = base + r * RADIX_THREADS + tid;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    // Lanes of this wave holding the same digit.
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = lanes_below(peers);
    if (valid && rank == 0) s_wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + rank;
      for (int ww = 0; ww < w; ++ww) pos += s_wc[ww][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    {
      uint32_t add = 0;
#pragma unroll
      for (int ww = 0; ww < 4; ++ww) {
        add += s_wc[ww][tid];
        s_wc[ww][tid] = 0;
      }
      s_base[tid] += add;
    }
    __syncthreads();
  }
}

// Runs of equal codes are ordered by storage position after the stable sort; each run is reordered by particle id
// so the final order is (code, id). Storage holds the previous evaluation's sorted order, so a run is usually short
// and nearly in order: its first thread insertion-sorts runs of up to kShortRun entries. A longer run (a dense clump
// inside one Morton cell, as when escaping particles stretch the box) is listed for k_fix_long_ties, which sorts it
// with a whole workgroup in O(r log^2 r).
constexpr int kShortRun = 64;
constexpr int kLongThreads = 1024;
constexpr int kLongLds = 4096;  // runs up to this length sort in LDS
constexpr int kLongBlocks = 64;

__global__ __launch_bounds__(256) void k_fix_ties(const uint32_t* __restrict__ skeys, uint32_t* __restrict__ perm,
                                                  const uint32_t* __restrict__ ids, int64_t n,
                                                  uint32_t* __restrict__ flags, uint32_t* __restrict__ long_runs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t c = skeys[t];
  if (t > 0 && skeys[t - 1] == c) return;          // not a run start
  if (t + 1 >= n || skeys[t + 1] != c) return;     // run of one
  const int64_t lim = min(n, t + kShortRun + 1);
  int64_t e = t + 1;
  while (e < lim && skeys[e] == c) ++e;
  if (e - t > kShortRun) {
    // At most n / (kShortRun + 1) runs are this long, the capacity of long_runs.
    long_runs[atomicAdd(&flags[FLAG_TIE_RUNS], 1u)] = (uint32_t)t;
    return;
  }
  for (int64_t a = t + 1; a < e; ++a) {
    const uint32_t pv = perm[a];
    const uint32_t idv = ids[pv];
    int64_t b = a - 1;
    while (b >= t && ids[perm[b]] > idv) {
      perm[b + 1] = perm[b];
      --b;
    }
    perm[b + 1] = pv;
  }
}

// Ascending bitonic network on r keyed entries (any r: the entries past r count as +infinity, and every comparison
// with one of them leaves the pair as it is). Each merge stage first compares mirrored pairs (i, i ^ (k - 1)), then
// half-cleans with distances k/4 ... 1, so all comparisons order the pair ascending. ACC provides key(i), swap(i, j).
template <class ACC>
__device__ void bitonic_sort_block(ACC& a, uint32_t r) {
  uint32_t P = 1;
  while (P < r) P <<= 1;
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t p = threadIdx.x; p < P / 2; p += blockDim.x) {
      const uint32_t half = k >> 1;
      const uint32_t i = (p / half) * k + (p % half);
      const uint32_t j = i ^ (k - 1);
      if (j < r && a.key(j) < a.key(i)) a.swap(i, j);
    }
    a.sync();
    for (uint32_t d = k >> 2; d >= 1; d >>= 1) {
      for (uint32_t p = threadIdx.x; p < P / 2; p += blockDim.x) {
        const uint32_t i = (p / d) * 2 * d + (p % d);
        const uint32_t j = i + d;
        if (j < r && a.key(j) < a.key(i)) a.swap(i, j);
      }
      a.sync();
    }
  }
}

struct LdsRun {
  uint32_t* k;
  uint32_t* v;
  __device__ uint32_t key(uint32_t i) const { return k[i]; }
  __device__ void swap(uint32_t i, uint32_t j) {
    const uint32_t ki = k[i], vi = v[i];
    k[i] = k[j];
    v[i] = v[j];
    k[j] = ki;
    v[j] = vi;
  }
  __device__ void sync() { __syncthreads(); }
};

// Global-memory run: (id, storage index) pairs in the sort's idle ping-pong buffers at the run's own offsets, so runs
// sorted by different workgroups never overlap.
struct GlobalRun {
  uint32_t* k;
  uint32_t* v;
  __device__ uint32_t key(uint32_t i) const { return k[i]; }
  __device__ void swap(uint32_t i, uint32_t j) {
    const uint32_t ki = k[i], vi = v[i];
    k[i] = k[j];
    v[i] = v[j];
    k[j] = ki;
    v[j] = vi;
  }
  __device__ void sync() {
    __threadfence_block();
    __syncthreads();
  }
};

// One workgroup per listed run: find the run's end, sort its (id, storage index) pairs by id, write the storage
// indices back. kLongBlocks workgroups stride over the list; with no long runs they return at once.
__global__ __launch_bounds__(kLongThreads) void k_fix_long_ties(const uint32_t* __restrict__ skeys,
                                                                uint32_t* __restrict__ perm,
                                                                const uint32_t* __restrict__ ids, int64_t n,
                                                                const uint32_t* __restrict__ flags,
                                                                const uint32_t* __restrict__ long_runs,
                                                                uint32_t* __restrict__ tkey, uint32_t* __restrict__ tval) {
  __shared__ uint32_t s_key[kLongLds], s_val[kLongLds];
  __shared__ uint32_t s_end;
  const uint32_t nl = flags[FLAG_TIE_RUNS];
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const uint32_t t = long_runs[q];
    const uint32_t c = skeys[t];
    if (threadIdx.x == 0) s_end = (uint32_t)n;
    __syncthreads();
    // The run's end: the first position past t whose code differs, found kLongThreads positions at a time.
    for (uint32_t base = t;; base += kLongThreads) {
      const uint32_t j = base + threadIdx.x;
      if (j < (uint32_t)n && skeys[j] != c) atomicMin(&s_end, j);
      __syncthreads();
      const bool done = s_end != (uint32_t)n || base + kLongThreads >= (uint32_t)n;
      __syncthreads();
      if (done) break;
    }
    const uint32_t r = s_end - t;
    if (r <= (uint32_t)kLongLds) {
      for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) {
        const uint32_t pv = perm[t + i];
        s_val[i] = pv;
        s_key[i] = ids[pv];
      }
      __syncthreads();
      LdsRun acc{s_key, s_val};
      bitonic_sort_block(acc, r);
      for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) perm[t + i] = s_val[i];
    } else {
      for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) {
        const uint32_t pv = perm[t + i];
        tval[t + i] = pv;
        tkey[t + i] = ids[pv];
      }
      __threadfence_block();
      __syncthreads();
      GlobalRun acc{tkey + t, tval + t};
      bitonic_sort_block(acc, r);
      for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) perm[t + i] = tval[t + i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// Karras 2012 over the L leaf cluster codes: the first sorted code of each bucket of K particles, skeys[c K].
// Karras' delta; leaves of different segments never share a node (-1), which splits the build into one tree
// per segment (lseg = segment of each sorted particle; leaves hold K particles of one segment).
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ cc, int64_t L, int64_t i, int64_t j,
                                      const uint32_t* __restrict__ lseg, int K) {
  if (j < 0 || j >= L) return -1;
  if (lseg && lseg[i * K] != lseg[j * K]) return -1;
  const uint32_t a = cc[i * K], b = cc[j * K];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

// Thread i builds internal node i (i < L - 1); thread c also settles leaf c when it is a whole segment (or the whole
// tree): such a leaf is its own root and upsweep block root. Every parent entry is written once: by the parent's
// thread, or NONE by the node's own thread for roots and for indices that name no node. cc = the sorted codes; leaf c's
// code is cc[c K].
__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ cc, int64_t L, int2* __restrict__ child,
                                                int2* __restrict__ range, uint32_t* __restrict__ parent,
                                                uint32_t* __restrict__ ready, uint32_t* __restrict__ npend,
                                                const uint32_t* __restrict__ lseg, int K, uint32_t* __restrict__ bstart,
                                                int64_t bk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *npend = 0u;
  if (i >= L) return;
  if ((i == 0 || (lseg && lseg[i * K] != lseg[(i - 1) * K])) && (i == L - 1 || (lseg && lseg[i * K] != lseg[(i + 1) * K]))) {
    parent[L - 1 + i] = IHS_NONE;
    bstart[i] = (uint32_t)(L - 1 + i);
  }
  if (i >= L - 1) return;
  // In a forest the index of each segment's last leaf (but the final one) names no node: its range would
  // duplicate the segment's root at the segment's first leaf.
  if (lseg && lseg[i * K] != lseg[(i + 1) * K]) {
    parent[i] = IHS_NONE;
    return;
  }
  const int64_t d = kdelta(cc, L, i, i + 1, lseg, K) > kdelta(cc, L, i, i - 1, lseg, K) ? 1 : -1;
  const int dmin = kdelta(cc, L, i, i - d, lseg, K);
  int64_t lmax = 2;
  while (kdelta(cc, L, i, i + lmax * d, lseg, K) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(cc, L, i, i + (l + t) * d, lseg, K) > dmin) l += t;
  const int64_t j = i + l * d;
  const int64_t first = i < j ? i : j, last = i < j ? j : i;
  int64_t split = first;
  const uint32_t fc = cc[first * K], lc = cc[last * K];
  if (fc == lc) {
    // All codes in [first, last] are equal: split where the index-augmented key (code, index) flips its
    // highest differing index bit, the split that is consistent with the ranges kdelta() assigns to the
    // children. The reference's (first + last) >> 1 (bh_gravity.wgsl:159-160) agrees for runs of two but
    // yields a tree with shared or orphaned nodes for longer runs of equal cluster codes.
    const int p = 31 - __clz((uint32_t)first ^ (uint32_t)last);
    split = (int64_t)((((uint32_t)first >> p) + 1u) << p) - 1;
  } else {
    const int cp = __clz(fc ^ lc);
    int64_t step = last - first;
    do {
      step = (step + 1) >> 1;
      const int64_t ns = split + step;
      if (ns < last && __clz(fc ^ cc[ns * K]) > cp) split = ns;
    } while (step > 1);
  }
  const uint32_t li = split == first ? (uint32_t)(L - 1 + split) : (uint32_t)split;
  const uint32_t ri = split + 1 == last ? (uint32_t)(L - 1 + split + 1) : (uint32_t)(split + 1);
  child[i] = make_int2((int)li, (int)ri);
  range[i] = make_int2((int)first, (int)last);
  parent[li] = (uint32_t)i;
  parent[ri] = (uint32_t)i;
  ready[i] = 0u;
  const bool root = (first == 0 || (lseg && lseg[first * K] != lseg[(first - 1) * K])) &&
                    (last == L - 1 || (lseg && lseg[last * K] != lseg[(last + 1) * K]));
  if (root) parent[i] = IHS_NONE;
  // Upsweep block roots: a child of at most bk leaves under a node of more, or a segment's root of at most bk.
  if (last - first + 1 > bk) {
    if (split - first + 1 <= bk) bstart[first] = li;
    if (last - split <= bk) bstart[split + 1] = ri;
  } else if (root) {
    bstart[first] = (uint32_t)i;
  }
}

// ---------------------------------------------------------------------------------------------
// Upsweep. Phase 1: one workgroup per upsweep block root (a subtree of at most UB_LEAVES_* leaves whose parent holds
// more; k_karras marks them) sums its leaves and climbs the subtree with the node values, topology and arrival
// counters in LDS; a node is merged by the second child to arrive. The climb that completes a block root records its
// arrival at the parent. Phase 2 (one workgroup) replays the recorded arrivals up the nodes above the block roots.

constexpr int UB_LEAVES_1 = 1024;  // block-root leaf capacity with one particle per leaf (1,024 threads, ~112 KB LDS)
constexpr int UB_LEAVES_K = 512;   // with buckets of K > 1 particles (their full leaf values in LDS, ~82 KB)
constexpr int TOP_SLOTS = 8192;                          // top-phase arrival table in LDS (node ids)
constexpr int TOP_PROBES = 32;                           // probes before a node counts in global memory

// Mass-normalised traceless quadrupole q = Q / M (units m^2: Q itself overflows FP32 in SI units) merged about
// the parent's centre of mass (parallel-axis shift): q = sum_c (m_c / M) (q_c + 3 s s^T - |s|^2 I),
// s = com_c - com. w = m_c / M.
__device__ __forceinline__ void quad_shift(float4& qa, float2& qb, const float4 cq, const float4 qca, const float2 qcb,
                                           const float4 c, float w) {
#pragma clang fp contract(off)
  const float sx = cq.x - c.x, sy = cq.y - c.y, sz = cq.z - c.z;
  const float s2 = sx * sx + sy * sy + sz * sz;
  qa.x += w * (qca.x + 3.0f * sx * sx - s2);
  qa.y += w * (qca.y + 3.0f * sy * sy - s2);
  qa.z += w * (qca.z + 3.0f * sz * sz - s2);
  qa.w += w * (qca.w + 3.0f * sx * sy);
  qb.x += w * (qcb.x + 3.0f * sx * sz);
  qb.y += w * (qcb.y + 3.0f * sy * sz);
}

// One node's moments and boxes, as the node arrays hold them.
struct NodeVals {
  float4 c, b0, b1, qa;
  float2 qb;
};

__device__ __forceinline__ NodeVals load_node(uint32_t x, const float4* ncom, const float4* nbmin, const float4* nbmax,
                                              const float4* nqa, const float2* nqb) {
  return NodeVals{ncom[x], nbmin[x], nbmax[x], nqa[x], nqb[x]};
}

__device__ __forceinline__ void store_node(uint32_t x, const NodeVals& v, float4* ncom, float4* nbmin, float4* nbmax,
                                           float4* nqa, float2* nqb) {
  ncom[x] = v.c;
  nbmin[x] = v.b0;
  nbmax[x] = v.b1;
  nqa[x] = v.qa;
  nqb[x] = v.qb;
}

// The moment arithmetic is not contracted into FMAs, so a node gets the same bits whichever kernel merges it
// (batched and single scenarios cut the leaves into blocks at different places). l and r are the left and right
// children. 1/M and the two lengths of the critical radius use the hardware reciprocal and square root (1 ulp; the
// IEEE sequences were most of a merge's instructions on the climb's critical path).
__device__ __forceinline__ NodeVals merge_vals(const NodeVals& l, const NodeVals& r, float inv_theta) {
#pragma clang fp contract(off)
  const float4 cl = l.c, cr = r.c;
  const float4 bl0 = l.b0, br0 = r.b0;
  const float4 bl1 = l.b1, br1 = r.b1;
  const float4 qla = l.qa, qra = r.qa;
  const float2 qlb = l.qb, qrb = r.qb;
  const float m = cl.w + cr.w;
  float4 c = make_float4(0.f, 0.f, 0.f, m);
  float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 qb = make_float2(0.f, 0.f);
  if (m > 0.f) {
    const float im = __builtin_amdgcn_rcpf(m);
    c.x = (cl.x * cl.w + cr.x * cr.w) * im;
    c.y = (cl.y * cl.w + cr.y * cr.w) * im;
    c.z = (cl.z * cl.w + cr.z * cr.w) * im;
    quad_shift(qa, qb, cl, qla, qlb, c, cl.w * im);
    quad_shift(qa, qb, cr, qra, qrb, c, cr.w * im);
  }
  const float4 b0 = make_float4(fminf(bl0.x, br0.x), fminf(bl0.y, br0.y), fminf(bl0.z, br0.z), fmaxf(bl0.w, br0.w));
  float4 b1 = make_float4(fmaxf(bl1.x, br1.x), fmaxf(bl1.y, br1.y), fmaxf(bl1.z, br1.z), 0.f);
  const float sx = b1.x - b0.x, sy = b1.y - b0.y, sz = b1.z - b0.z;
  const float ex = 0.5f * (b0.x + b1.x) - c.x, ey = 0.5f * (b0.y + b1.y) - c.y, ez = 0.5f * (b0.z + b1.z) - c.z;
  b1.w = __builtin_amdgcn_sqrtf(sx * sx + sy * sy + sz * sz) * inv_theta +
         __builtin_amdgcn_sqrtf(ex * ex + ey * ey + ez * ez);
  return NodeVals{c, b0, b1, qa, qb};
}

// The parent of a held node and its sibling: the operands are put in (left, right) order first, so lanes that hold
// left and right children run one merge, not both branches.
__device__ __forceinline__ NodeVals sel(bool a, const NodeVals& x, const NodeVals& y) {
  return NodeVals{a ? x.c : y.c, a ? x.b0 : y.b0, a ? x.b1 : y.b1, a ? x.qa : y.qa, a ? x.qb : y.qb};
}
__device__ __forceinline__ NodeVals merge_pair(bool held_left, const NodeVals& held, const NodeVals& sib,
                                               float inv_theta) {
  return merge_vals(sel(held_left, held, sib), sel(held_left, sib, held), inv_theta);
}

// The search-radius-only upsweep (HONLY: moments, boxes and topology unchanged since the last full upsweep, only the
// search radii hs changed) carries each node's box and largest radius, the values the walk groups are marked from,
// and stores the radius alone. min and max are exact, so the boxes and radii come out as a full merge forms them.
__device__ __forceinline__ NodeVals merge_box(const NodeVals& a, const NodeVals& b) {
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  return NodeVals{z,
                  make_float4(fminf(a.b0.x, b.b0.x), fminf(a.b0.y, b.b0.y), fminf(a.b0.z, b.b0.z), fmaxf(a.b0.w, b.b0.w)),
                  make_float4(fmaxf(a.b1.x, b.b1.x), fmaxf(a.b1.y, b.b1.y), fmaxf(a.b1.z, b.b1.z), 0.f), z,
                  make_float2(0.f, 0.f)};
}
__device__ __forceinline__ NodeVals load_box(uint32_t x, const float4* nbmin, const float4* nbmax) {
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  return NodeVals{z, nbmin[x], nbmax[x], z, make_float2(0.f, 0.f)};
}
__device__ __forceinline__ void store_hs(uint32_t x, float hm, float4* nbmin) {
  reinterpret_cast<float*>(nbmin + x)[3] = hm;
}

// A walk group's box bound: diagonal at most diag_hs times the node's largest search radius (none: no bound).
__device__ __forceinline__ bool group_box_ok(const float4 b0, const float4 b1, float diag_hs) {
  if (!(b0.w > 0.f)) return true;
  const float dx = b1.x - b0.x, dy = b1.y - b0.y, dz = b1.z - b0.z;
  return dx * dx + dy * dy + dz * dz <= (diag_hs * b0.w) * (diag_hs * b0.w);
}

// Parent p of `from`, whose values the caller holds: the sibling is loaded and the children merged in order.
__device__ __forceinline__ NodeVals merge_with_sibling(const NodeVals& cur, uint32_t from, const int2 ch,
                                                       const float4* ncom, const float4* nbmin, const float4* nbmax,
                                                       const float4* nqa, const float2* nqb, float inv_theta) {
  const bool left = (uint32_t)ch.x == from;
  const NodeVals sib = load_node(left ? (uint32_t)ch.y : (uint32_t)ch.x, ncom, nbmin, nbmax, nqa, nqb);
  return merge_pair(left, cur, sib, inv_theta);
}

// Particles of a leaf read from the sorted working copies (search radius 0 without them).
struct GlobalParticles {
  const float4* xm;
  const float* hs;
  __device__ __forceinline__ float4 x(int64_t s) const { return xm[s]; }
  __device__ __forceinline__ float h(int64_t s) const { return hs ? hs[s] : 0.f; }
};
// Moments and box of the leaf holding particles [start, end), read through src.
template <class P>
__device__ __forceinline__ NodeVals leaf_vals(const P& src, int64_t start, int64_t end, float inv_theta) {
  // The centre of mass is formed relative to the leaf's first particle: a one-particle leaf's is then exactly
  // that particle's position (m x / m is not, in FP32), so a walk never sees a particle's own leaf as a source
  // an ulp away from it.
  const float4 b0 = src.x(start);
  float m = 0.f, mx = 0.f, my = 0.f, mz = 0.f, hm = 0.f;
  float lo[3] = {1.0e30f, 1.0e30f, 1.0e30f}, hi[3] = {-1.0e30f, -1.0e30f, -1.0e30f};
  for (int64_t s = start; s < end; ++s) {
    const float4 b = src.x(s);
    m += b.w;
    mx += (b.x - b0.x) * b.w;
    my += (b.y - b0.y) * b.w;
    mz += (b.z - b0.z) * b.w;
    lo[0] = fminf(lo[0], b.x); lo[1] = fminf(lo[1], b.y); lo[2] = fminf(lo[2], b.z);
    hi[0] = fmaxf(hi[0], b.x); hi[1] = fmaxf(hi[1], b.y); hi[2] = fmaxf(hi[2], b.z);
    hm = fmaxf(hm, src.h(s));
  }
  float4 cm = make_float4(b0.x, b0.y, b0.z, m);
  if (m > 0.f) {
    cm.x += mx / m;
    cm.y += my / m;
    cm.z += mz / m;
  }
  // Second pass for the quadrupole about the leaf's centre of mass (no large-coordinate cancellation).
  float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 qb = make_float2(0.f, 0.f);
  if (end - start > 1 && m > 0.f) {
    const float im = 1.0f / m;
    for (int64_t s = start; s < end; ++s) {
      const float4 b = src.x(s);
      quad_shift(qa, qb, b, make_float4(0.f, 0.f, 0.f, 0.f), make_float2(0.f, 0.f), cm, b.w * im);
    }
  }
  const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
  const float ex = 0.5f * (lo[0] + hi[0]) - cm.x, ey = 0.5f * (lo[1] + hi[1]) - cm.y, ez = 0.5f * (lo[2] + hi[2]) - cm.z;
  return NodeVals{cm, make_float4(lo[0], lo[1], lo[2], hm),
                  make_float4(hi[0], hi[1], hi[2],
                              sqrtf(sx * sx + sy * sy + sz * sz) * inv_theta + sqrtf(ex * ex + ey * ey + ez * ez)),
                  qa, qb};
}

// The box and largest search radius of the leaf holding particles [start, end) (HONLY), as leaf_vals forms them.
template <class P>
__device__ __forceinline__ NodeVals leaf_box(const P& src, int64_t start, int64_t end) {
  float hm = 0.f;
  float lo[3] = {1.0e30f, 1.0e30f, 1.0e30f}, hi[3] = {-1.0e30f, -1.0e30f, -1.0e30f};
  for (int64_t s = start; s < end; ++s) {
    const float4 b = src.x(s);
    lo[0] = fminf(lo[0], b.x); lo[1] = fminf(lo[1], b.y); lo[2] = fminf(lo[2], b.z);
    hi[0] = fmaxf(hi[0], b.x); hi[1] = fmaxf(hi[1], b.y); hi[2] = fmaxf(hi[2], b.z);
    hm = fmaxf(hm, src.h(s));
  }
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  return NodeVals{z, make_float4(lo[0], lo[1], lo[2], hm), make_float4(hi[0], hi[1], hi[2], 0.f), z,
                  make_float2(0.f, 0.f)};
}

// LDS of one phase-1 workgroup: values of the subtree's internal nodes (node x at x - f, f = the root's first leaf;
// a subtree's internal nodes are numbered inside its leaf range), its leaves (leaf c at c - f; a one-particle leaf's
// quadrupole is zero and not stored), the internal nodes' children and parents, their arrival counters, and the
// tile's block roots.
template <int BK, bool KONE>
struct UbLds {
  float4 ic[BK], ib0[BK], ib1[BK], iqa[BK];
  float2 iqb[BK];
  float4 lc[BK], lb0[BK], lb1[BK];
  float4 lqa[KONE ? 1 : BK];
  float2 lqb[KONE ? 1 : BK];
  int2 child[BK];
  uint32_t parent[BK];
  uint32_t icount[BK];  // internal nodes' leaf counts (walk-group marking)
  uint32_t arrive[BK];
  uint32_t roots[BK];
  uint32_t wcnt[BK / 64];
};

// A leaf of K particles is within the walk-group size limit.
__device__ __forceinline__ bool hc_leaf_ok(int64_t c, int K, int64_t n, int gmax) {
  return min((c + 1) * K, n) - c * K <= (int64_t)gmax;
}

// Phase 1. Workgroup b takes the block roots whose first leaf lies in leaves [b BK, (b + 1) BK), one after another.
// For a root over leaves [f, l], thread t sums leaf f + t (stored to the node arrays as well as to LDS) and climbs:
// it publishes the node it holds in LDS, counts its arrival at the parent (LDS atomic; the fences order LDS only, so
// no wait for the global stores sits on the climb), and the second arrival merges its held values with the
// sibling's from LDS in (left, right) order. Merging is not contracted into FMAs, so every node gets the same bits
// whichever thread or kernel merges it. The thread that completes the root records the arrival at its parent for
// phase 2. HONLY: only the search radii changed since the last full upsweep; boxes and radii are carried (merge_box)
// and the nodes' radii stored, the moments in memory stand.
template <bool KONE, bool HONLY>
__global__ __launch_bounds__(KONE ? UB_LEAVES_1 : UB_LEAVES_K) void k_upsweep_blocks(
    const float4* __restrict__ xm, const float* __restrict__ hs, int64_t n, int K, int64_t L,
    const int2* __restrict__ child, const int2* __restrict__ range, const uint32_t* __restrict__ parent,
    const uint32_t* __restrict__ bstart, float4* __restrict__ ncom, float4* __restrict__ nbmin,
    float4* __restrict__ nbmax, float4* __restrict__ nqa, float2* __restrict__ nqb, float inv_theta,
    uint32_t* __restrict__ pend, uint32_t* __restrict__ npend, const uint32_t* __restrict__ gate,
    uint32_t* __restrict__ gstart, float diag_hs, int gmax) {
  constexpr int BK = KONE ? UB_LEAVES_1 : UB_LEAVES_K;
  if (gate && *gate == 0u) return;
  __shared__ UbLds<BK, KONE> S;
  const int t = (int)threadIdx.x;
  const int w = t >> 6;
  // The tile's block roots in leaf order (a root holds at most BK leaves, so every tile holds at least one start).
  const int64_t c0 = (int64_t)blockIdx.x * BK + t;
  const uint32_t rs = c0 < L ? bstart[c0] : IHS_NONE;
  const uint64_t bal = __ballot(rs != IHS_NONE);
  if ((t & 63) == 0) S.wcnt[w] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t base = 0, nroots = 0;
#pragma unroll
  for (int k = 0; k < BK / 64; ++k) {
    const uint32_t v = S.wcnt[k];
    base += k < w ? v : 0u;
    nroots += v;
  }
  if (rs != IHS_NONE) S.roots[base + lanes_below(bal)] = rs;
  __syncthreads();
  const uint32_t Lm1 = (uint32_t)(L - 1);
  for (uint32_t q = 0; q < nroots; ++q) {
    const uint32_t r = S.roots[q];
    int64_t f = (int64_t)r - (L - 1), l = f;
    if (r < Lm1) {
      const int2 rg = range[r];
      f = rg.x;
      l = rg.y;
    }
    const int64_t cnt = l - f + 1;
    NodeVals cur;
    uint32_t from = IHS_NONE, p = IHS_NONE, p0 = IHS_NONE, hc = 0;
    if (t < cnt) {
      const int64_t x = f + t;
      hc = (uint32_t)(min((x + 1) * K, n) - x * K);
      if (x < L - 1) {
        S.child[t] = child[x];
        S.parent[t] = parent[x];
      }
      S.arrive[t] = 0u;
      from = (uint32_t)(L - 1 + x);
      if (HONLY) {
        cur = leaf_box(GlobalParticles{xm, hs}, x * K, min((x + 1) * K, n));
        store_hs(from, cur.b0.w, nbmin);
      } else {
        cur = leaf_vals(GlobalParticles{xm, hs}, x * K, min((x + 1) * K, n), inv_theta);
        store_node(from, cur, ncom, nbmin, nbmax, nqa, nqb);
      }
      p = parent[from];
      p0 = p;
      if (!HONLY) S.lc[t] = cur.c;
      S.lb0[t] = cur.b0;
      S.lb1[t] = cur.b1;
      if (!KONE && !HONLY) {
        S.lqa[t] = cur.qa;
        S.lqb[t] = cur.qb;
      }
    }
    __syncthreads();
    if (t < cnt) {
      while (from != r) {
        const int pi = (int)((int64_t)p - f);
        if (from < Lm1) {
          const int fi = (int)((int64_t)from - f);
          if (!HONLY) {
            S.ic[fi] = cur.c;
            S.iqa[fi] = cur.qa;
            S.iqb[fi] = cur.qb;
          }
          S.ib0[fi] = cur.b0;
          S.ib1[fi] = cur.b1;
          S.icount[fi] = hc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        const uint32_t old = __hip_atomic_fetch_add(&S.arrive[pi], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0u) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const int2 ch = S.child[pi];
        const bool left = (uint32_t)ch.x == from;
        const uint32_t sib = left ? (uint32_t)ch.y : (uint32_t)ch.x;
        NodeVals sv;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const float2 z2 = make_float2(0.f, 0.f);
        if (sib >= Lm1) {
          const int64_t sc = (int64_t)(sib - Lm1);
          const int si = (int)(sc - f);
          sv = NodeVals{HONLY ? z4 : S.lc[si], S.lb0[si], S.lb1[si], KONE || HONLY ? z4 : S.lqa[si],
                        KONE || HONLY ? z2 : S.lqb[si]};
          hc += (uint32_t)(min((sc + 1) * K, n) - sc * K);
        } else {
          const int si = (int)((int64_t)sib - f);
          sv = NodeVals{HONLY ? z4 : S.ic[si], S.ib0[si], S.ib1[si], HONLY ? z4 : S.iqa[si], HONLY ? z2 : S.iqb[si]};
          hc += S.icount[si];
        }
        if (HONLY) {
          cur = merge_box(cur, sv);
          store_hs(p, cur.b0.w, nbmin);
        } else {
          cur = merge_pair(left, cur, sv, inv_theta);
          store_node(p, cur, ncom, nbmin, nbmax, nqa, nqb);
        }
        from = p;
        p = S.parent[pi];
      }
      if (from == r) {
        if (p != IHS_NONE) pend[atomicAdd(npend, 1u)] = p;
        if (gmax > 0 && r < Lm1) {
          const int ri = (int)((int64_t)r - f);
          S.ib0[ri] = cur.b0;
          S.ib1[ri] = cur.b1;
          S.icount[ri] = hc;
        }
      }
    }
    if (gmax > 0) {
      // Walk groups (as k_mark_groups, from LDS): the topmost node on leaf c's root path holding at most gmax
      // particles whose box passes group_box_ok; its first leaf records it. Such a node lies inside the block root
      // (the root's parent holds more than BK >= gmax leaves).
      __syncthreads();
      if (t < cnt) {
        const int64_t c = f + t;
        uint32_t best = IHS_NONE;
        bool first = true, best_first = false;
        if (hc_leaf_ok(c, K, n, gmax)) {
          if (group_box_ok(S.lb0[t], S.lb1[t], diag_hs)) {
            best = (uint32_t)(L - 1 + c);
            best_first = true;
          }
          uint32_t at = (uint32_t)(L - 1 + c), x = p0;
          while (at != r && x != IHS_NONE) {
            const int xi = (int)((int64_t)x - f);
            if (S.icount[xi] > (uint32_t)gmax) break;
            first = first && (uint32_t)S.child[xi].x == at;
            if (group_box_ok(S.ib0[xi], S.ib1[xi], diag_hs)) {
              best = x;
              best_first = first;
            }
            at = x;
            x = S.parent[xi];
          }
        }
        gstart[c] = best == IHS_NONE ? GROUP_SPLIT : (best_first ? best : IHS_NONE);
      }
    }
    __syncthreads();
  }
}

// Arrival at node q in the top phase's table: true for the second arrival. The first arrival at q inserts q with
// one compare-and-swap per probe (linear probing from a multiplicative hash); the second finds it. Slots are never
// freed, so a node whose probe run is full for its first arrival is full for its second too, and both count in
// ready[q] instead.
__device__ __forceinline__ bool top_arrive(uint32_t* keys, uint32_t q, uint32_t* ready) {
  uint32_t slot = (q * 2654435761u) >> (32 - 13);
  static_assert(TOP_SLOTS == 1 << 13, "the hash yields 13 bits");
  for (int k = 0; k < TOP_PROBES; ++k) {
    uint32_t expected = IHS_NONE;
    if (__hip_atomic_compare_exchange_strong(&keys[slot], &expected, q, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return false;  // inserted: first arrival
    if (expected == q) return true;
    slot = (slot + 1u) & (TOP_SLOTS - 1);
  }
  return __hip_atomic_fetch_add(&ready[q], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
}

// Each thread replays recorded arrivals; every node above the blocks is merged once, by its second arrival, found
// in an LDS table of node ids. As in phase 1 a climb holds its merged node and prefetches the next level's topology.
template <bool HONLY>
__global__ __launch_bounds__(1024) void k_upsweep_top(const uint32_t* __restrict__ pend,
                                                      const uint32_t* __restrict__ npend, const int2* __restrict__ child,
                                                      const uint32_t* __restrict__ parent, uint32_t* ready,
                                                      float4* ncom, float4* nbmin, float4* nbmax, float4* nqa,
                                                      float2* nqb, float inv_theta, const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t keys[TOP_SLOTS];
  for (int i = threadIdx.x; i < TOP_SLOTS; i += blockDim.x) keys[i] = IHS_NONE;
  __syncthreads();
  const uint32_t np = *npend;
  for (uint32_t e = threadIdx.x; e < np; e += blockDim.x) {
    uint32_t q = pend[e];
    int2 ch = child[q];
    uint32_t qp = parent[q];
    NodeVals cur;
    uint32_t from = IHS_NONE;
    for (;;) {
      int2 ch2 = make_int2(0, 0);
      uint32_t qp2 = IHS_NONE;
      if (qp != IHS_NONE) {
        ch2 = child[qp];
        qp2 = parent[qp];
      }
      if (!top_arrive(keys, q, ready)) break;
      if (HONLY) {
        if (from == IHS_NONE)
          cur = merge_box(load_box((uint32_t)ch.x, nbmin, nbmax), load_box((uint32_t)ch.y, nbmin, nbmax));
        else
          cur = merge_box(cur, load_box((uint32_t)ch.x == from ? (uint32_t)ch.y : (uint32_t)ch.x, nbmin, nbmax));
        store_hs(q, cur.b0.w, nbmin);
      } else {
        if (from == IHS_NONE)
          cur = merge_vals(load_node((uint32_t)ch.x, ncom, nbmin, nbmax, nqa, nqb),
                           load_node((uint32_t)ch.y, ncom, nbmin, nbmax, nqa, nqb), inv_theta);
        else
          cur = merge_with_sibling(cur, from, ch, ncom, nbmin, nbmax, nqa, nqb, inv_theta);
        store_node(q, cur, ncom, nbmin, nbmax, nqa, nqb);
      }
      if (qp == IHS_NONE) break;
      from = q;
      q = qp;
      ch = ch2;
      qp = qp2;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Walk groups: node x is a group root when it holds <= gmax particles and its parent holds more.
// Along any leaf-to-root path the particle count grows, so the roots partition the leaves.
__device__ __forceinline__ int64_t node_count(const int2* __restrict__ range, int64_t L, int K, int64_t n, uint32_t x,
                                              int64_t& first) {
  int64_t last;
  if (x >= (uint32_t)(L - 1)) {
    first = last = (int64_t)x - (L - 1);
  } else {
    const int2 rg = range[x];
    first = rg.x;
    last = rg.y;
  }
  return min((last + 1) * K, n) - first * K;
}

__device__ __forceinline__ bool group_ok(const int2* __restrict__ range, int64_t L, int K, int64_t n, uint32_t x,
                                         const float4* __restrict__ nbmin, const float4* __restrict__ nbmax,
                                         float diag_hs, int gmax) {
  int64_t first;
  if (node_count(range, L, K, n, x, first) > gmax) return false;
  return group_box_ok(nbmin[x], nbmax[x], diag_hs);
}

// One thread per leaf: the topmost acceptable node on the leaf's root path is its group; the group's first
// leaf records it. A leaf with no acceptable node is split into single-particle groups.
__global__ __launch_bounds__(256) void k_mark_groups(int64_t n, int K, int64_t L, const int2* __restrict__ range,
                                                     const uint32_t* __restrict__ parent,
                                                     const float4* __restrict__ nbmin,
                                                     const float4* __restrict__ nbmax,
                                                     uint32_t* __restrict__ gstart, const uint32_t* __restrict__ gate,
                                                     float diag_hs, int gmax) {
  if (gate && *gate == 0u) return;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L) return;
  uint32_t best = IHS_NONE;
  for (uint32_t x = (uint32_t)(L - 1 + c); x != IHS_NONE; x = parent[x]) {
    int64_t first;
    if (node_count(range, L, K, n, x, first) > gmax) break;  // counts only grow towards the root
    if (group_ok(range, L, K, n, x, nbmin, nbmax, diag_hs, gmax)) best = x;
  }
  if (best == IHS_NONE) {
    gstart[c] = GROUP_SPLIT;
    return;
  }
  int64_t first;
  node_count(range, L, K, n, best, first);
  gstart[c] = first == c ? best : IHS_NONE;
}

constexpr int GRP_BLK = 1024;

__device__ __forceinline__ uint32_t group_weight(const uint32_t* __restrict__ gstart, int64_t L, int K, int64_t n,
                                                 int64_t c) {
  if (c >= L) return 0u;
  const uint32_t gs = gstart[c];
  if (gs == IHS_NONE) return 0u;
  if (gs == GROUP_SPLIT) return (uint32_t)(min((c + 1) * K, n) - c * K);
  return 1u;
}

// Block-wide (1024) exclusive scan of per-leaf group counts.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t* s_w /*16*/, uint32_t* total) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int ww = 0; ww < GRP_BLK / 64; ++ww) {
    const uint32_t t = s_w[ww];
    if (ww < w) base += t;
    tot += t;
  }
  if (total) *total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(GRP_BLK) void k_group_count(const uint32_t* __restrict__ gstart, int64_t L, int K,
                                                         int64_t n, uint32_t* __restrict__ gblk,
                                                         const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t s_w[GRP_BLK / 64];
  const int64_t c = (int64_t)blockIdx.x * GRP_BLK + threadIdx.x;
  uint32_t tot;
  block_excl_scan_1024(group_weight(gstart, L, K, n, c), s_w, &tot);
  if (threadIdx.x == 0) gblk[blockIdx.x] = tot;
}

// Exclusive scan of the per-block group counts (one workgroup; each thread owns a contiguous run of blocks).
__global__ __launch_bounds__(GRP_BLK) void k_group_offsets(uint32_t* __restrict__ gblk, int nblk,
                                                           uint32_t* __restrict__ ngroups, uint32_t* __restrict__ wq,
                                                           const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t s_w[GRP_BLK / 64];
  if (threadIdx.x < 16) wq[threadIdx.x] = 0u;
  const int per = (nblk + GRP_BLK - 1) / GRP_BLK;
  const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, nblk);
  uint32_t sum = 0;
  for (int b = b0; b < b1; ++b) sum += gblk[b];
  uint32_t total;
  uint32_t run = block_excl_scan_1024(sum, s_w, &total);
  for (int b = b0; b < b1; ++b) {
    const uint32_t c = gblk[b];
    gblk[b] = run;
    run += c;
  }
  if (threadIdx.x == 0) *ngroups = total;
}

__global__ __launch_bounds__(GRP_BLK) void k_group_emit(const uint32_t* __restrict__ gstart, int64_t L, int K, int64_t n,
                                                        const int2* __restrict__ range,
                                                        const uint32_t* __restrict__ gblk, int4* __restrict__ groups,
                                                        const uint32_t* __restrict__ gate,
                                                        const uint32_t* __restrict__ segof,
                                                        const uint32_t* __restrict__ root) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t s_w[GRP_BLK / 64];
  const int64_t c = (int64_t)blockIdx.x * GRP_BLK + threadIdx.x;
  const uint32_t wgt = group_weight(gstart, L, K, n, c);
  const uint32_t idx = gblk[blockIdx.x] + block_excl_scan_1024(wgt, s_w, nullptr);
  if (wgt == 0) return;
  const uint32_t gs = gstart[c];
  if (gs == GROUP_SPLIT) {
    const int rt = (int)root[segof ? segof[c * K] : 0u];
    for (uint32_t q = 0; q < wgt; ++q) groups[idx + q] = make_int4((int)(c * K + q), 1, (int)(L - 1 + c), rt);
    return;
  }
  int64_t first;
  const int64_t cnt = node_count(range, L, K, n, gs, first);
  groups[idx] = make_int4((int)(first * K), (int)cnt, (int)gs, (int)root[segof ? segof[first * K] : 0u]);
}

inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_bbox_morton(hipStream_t s, int split, const void* px, const void* py, const void* pz, int64_t n, int K,
                        TreeBufs& tb, const SegView& sv, const EvalInit& fe) {
  const int64_t L = (n + K - 1) / K;
  // bps * S <= BBOX_PARTS blocks (S <= IHS_MAX_SEGMENTS == BBOX_PARTS).
  const int bps = (int)std::max<int64_t>(1, std::min<int64_t>(BBOX_PARTS / sv.S, (n / sv.S + 255) / 256));
  const unsigned nbk = (unsigned)(bps * sv.S);
  uint32_t* codes0 = sv.S > 1 ? tb.codes0 : nullptr;
  if (split)
    k_bbox<double><<<nbk, 256, 0, s>>>((const double*)px, (const double*)py, (const double*)pz, sv.start, bps,
                                       tb.bbox_part, sv.S, fe);
  else
    k_bbox<float><<<nbk, 256, 0, s>>>((const float*)px, (const float*)py, (const float*)pz, sv.start, bps,
                                      tb.bbox_part, sv.S, fe);
  k_bbox_final<<<1, 1024, 0, s>>>(tb.bbox_part, (int)nbk, bps, sv.S, sv.bbox);
  if (split) {
    k_morton<double><<<grid_for(n, 256), 256, 0, s>>>((const double*)px, (const double*)py, (const double*)pz, n,
                                                      sv.bbox, sv.segof, tb.keys[0], tb.vals[0], codes0, tb.bstart, L);
  } else {
    k_morton<float><<<grid_for(n, 256), 256, 0, s>>>((const float*)px, (const float*)py, (const float*)pz, n,
                                                     sv.bbox, sv.segof, tb.keys[0], tb.vals[0], codes0, tb.bstart, L);
  }
}

void launch_segment_order(hipStream_t s, int64_t n, TreeBufs& tb, const SegView& sv) {
  if (sv.S <= 1) return;
  // Stable radix passes over the segment id with the Morton permutation as payload, in the buffers the Morton
  // passes did not end in; then the codes follow the final permutation.
  const int nb = radix_tiles(n);
  int cur = tb.skeys == tb.keys[0] ? 1 : 0;
  k_seg_keys<<<grid_for(n, 256), 256, 0, s>>>(tb.perm, sv.segof, n, tb.keys[cur], tb.vals[cur]);
  for (int shift = 0; (1 << shift) < sv.S; shift += 8) {
    k_radix_hist<<<nb, RADIX_THREADS, 0, s>>>(tb.keys[cur], n, shift, tb.hist, nb);
    k_radix_scan<<<256, 256, 0, s>>>(tb.hist, nb, tb.dtot);
    k_radix_scatter<<<nb, RADIX_THREADS, 0, s>>>(tb.keys[cur], tb.vals[cur], tb.keys[cur ^ 1], tb.vals[cur ^ 1], n,
                                                 shift, tb.hist, tb.dtot, nb);
    cur ^= 1;
  }
  tb.perm = tb.vals[cur];
  tb.skeys = tb.keys[cur ^ 1];
  k_gather_codes<<<grid_for(n, 256), 256, 0, s>>>(tb.perm, tb.codes0, n, tb.skeys);
}

void launch_radix_sort(hipStream_t s, int64_t n, int bits, TreeBufs& tb) {
  const int nb = radix_tiles(n);
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    k_radix_hist<<<nb, RADIX_THREADS, 0, s>>>(tb.keys[cur], n, shift, tb.hist, nb);
    k_radix_scan<<<256, 256, 0, s>>>(tb.hist, nb, tb.dtot);
    k_radix_scatter<<<nb, RADIX_THREADS, 0, s>>>(tb.keys[cur], tb.vals[cur], tb.keys[cur ^ 1], tb.vals[cur ^ 1], n,
                                                 shift, tb.hist, tb.dtot, nb);
    cur ^= 1;
  }
  tb.skeys = tb.keys[cur];
  tb.perm = tb.vals[cur];
}

int launch_radix_pairs(hipStream_t s, int64_t n, int bits, uint32_t* keys[2], uint32_t* vals[2], int cur,
                       uint32_t* hist, uint32_t* dtot) {
  const int nb = radix_tiles(n);
  for (int shift = 0; shift < bits; shift += 8) {
    k_radix_hist<<<nb, RADIX_THREADS, 0, s>>>(keys[cur], n, shift, hist, nb);
    k_radix_scan<<<256, 256, 0, s>>>(hist, nb, dtot);
    k_radix_scatter<<<nb, RADIX_THREADS, 0, s>>>(keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift, hist,
                                                 dtot, nb);
    cur ^= 1;
  }
  return cur;
}

void launch_fix_ties(hipStream_t s, int64_t n, const uint32_t* ids, TreeBufs& tb, uint32_t* flags) {
  k_fix_ties<<<grid_for(n, 256), 256, 0, s>>>(tb.skeys, tb.perm, ids, n, flags, tb.tie_runs);
  // The ping-pong buffers that hold neither the sorted codes nor the permutation are idle once the sort is done:
  // scratch for runs longer than LDS holds (with scenarios the two results can sit in different slots).
  uint32_t* tkey = tb.skeys == tb.keys[0] ? tb.keys[1] : tb.keys[0];
  uint32_t* tval = tb.perm == tb.vals[0] ? tb.vals[1] : tb.vals[0];
  k_fix_long_ties<<<kLongBlocks, kLongThreads, 0, s>>>(tb.skeys, tb.perm, ids, n, flags, tb.tie_runs, tkey, tval);
}

void launch_karras(hipStream_t s, int64_t n, int K, TreeBufs& tb, const SegView& sv) {
  const int64_t L = (n + K - 1) / K;
  tb.ub_leaves = K == 1 ? UB_LEAVES_1 : UB_LEAVES_K;
  k_karras<<<grid_for(L, 256), 256, 0, s>>>(tb.skeys, L, tb.child, tb.range, tb.parent, tb.ready, tb.npend, sv.segof, K,
                                            tb.bstart, tb.ub_leaves);
}

void launch_groups(hipStream_t s, int64_t n, int K, TreeBufs& tb, const SegView& sv, const uint32_t* gate,
                   float diag_hs, bool gravity, bool marked) {
  const int64_t L = (n + K - 1) / K;
  const int nblk = (int)grid_for(L, GRP_BLK);
  const int gmax = gravity ? tb.group_max : GROUP_MAX;
  int4* out = gravity ? tb.ggroups : tb.groups;
  uint32_t* nout = gravity ? tb.nggroups : tb.ngroups;
  if (!marked)
    k_mark_groups<<<grid_for(L, 256), 256, 0, s>>>(n, K, L, tb.range, tb.parent, tb.nbmin, tb.nbmax, tb.gstart, gate,
                                                   diag_hs, gmax);
  k_group_count<<<nblk, GRP_BLK, 0, s>>>(tb.gstart, L, K, n, tb.gblk, gate);
  k_group_offsets<<<1, GRP_BLK, 0, s>>>(tb.gblk, nblk, nout, tb.wq, gate);
  k_group_emit<<<nblk, GRP_BLK, 0, s>>>(tb.gstart, L, K, n, tb.range, tb.gblk, out, gate, sv.segof, sv.root);
}

template <bool HONLY>
void upsweep_launch(hipStream_t s, int64_t n, int K, int64_t L, float inv_theta, const float4* xm, const float* hs,
                    TreeBufs& tb, const uint32_t* gate, float group_diag_hs) {
  const int gmax = group_diag_hs > 0.f ? GROUP_MAX : 0;
  if (K == 1)
    k_upsweep_blocks<true, HONLY><<<grid_for(L, UB_LEAVES_1), UB_LEAVES_1, 0, s>>>(
        xm, hs, n, K, L, tb.child, tb.range, tb.parent, tb.bstart, tb.ncom, tb.nbmin, tb.nbmax, tb.nqa, tb.nqb,
        inv_theta, tb.pend, tb.npend, gate, tb.gstart, group_diag_hs, gmax);
  else
    k_upsweep_blocks<false, HONLY><<<grid_for(L, UB_LEAVES_K), UB_LEAVES_K, 0, s>>>(
        xm, hs, n, K, L, tb.child, tb.range, tb.parent, tb.bstart, tb.ncom, tb.nbmin, tb.nbmax, tb.nqa, tb.nqb,
        inv_theta, tb.pend, tb.npend, gate, tb.gstart, group_diag_hs, gmax);
  if (L > tb.ub_leaves)
    k_upsweep_top<HONLY><<<1, 1024, 0, s>>>(tb.pend, tb.npend, tb.child, tb.parent, tb.ready, tb.ncom, tb.nbmin,
                                            tb.nbmax, tb.nqa, tb.nqb, inv_theta, gate);
}

void launch_upsweep(hipStream_t s, int64_t n, int K, float theta, const float4* xm, const float* hs, TreeBufs& tb,
                    const uint32_t* gate, float group_diag_hs, bool hs_only) {
  const int64_t L = (n + K - 1) / K;
  const float inv_theta = theta > 0.f ? 1.0f / theta : 3.0e38f;
  if (hs_only) {
    upsweep_launch<true>(s, n, K, L, inv_theta, xm, hs, tb, gate, group_diag_hs);
  } else {
    upsweep_launch<false>(s, n, K, L, inv_theta, xm, hs, tb, gate, group_diag_hs);
    launch_mac_record(s, n, K, theta, tb);
  }
}
