// Host-side launchers of the gfx950 kernels (each .hip file defines its kernels and launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>

#include "ihs_sph.h"

// Error/diagnostic counters kept on device (indices into Workspace::flags).
enum FlagIndex : int {
  FLAG_NBR_OVERFLOW = 0,   // particles whose neighbour list exceeded its capacity
  FLAG_WALK_OVERFLOW = 1,  // tree-walk stack overflows (interactions dropped)
  FLAG_H_CAPPED = 2,       // h-solves stopped at the searched radius (isolated particles)
  FLAG_H_UNCONVERGED = 3,  // h-solves that hit the iteration limit
  FLAG_TIE_RUNS = 4,       // this evaluation: runs of equal Morton codes too long for the short-run sort
  FLAG_NONFINITE = 5,      // particles with non-finite results
  FLAG_H_RETRY = 6,        // this evaluation: seeded solves stopped by the tight search radius (gate of the
                           // wider-radius pass)
  FLAG_RETRIES = 7,        // evaluations that took the wider-radius pass
  FLAG_COUNT = 8
};
// 64-bit counters (Workspace::counters).
enum CounterIndex : int {
  CNT_H_ITERS = 0,   // total Newton updates
  CNT_H_EVALS = 1,   // total density evaluations in the h-solve
  CNT_H_SOLVED = 2,  // particles solved
  CNT_PAIRS = 3,     // neighbour list entries
  CNT_GRAV_CELL = 4, // gravity: target x accepted-node interactions evaluated
  CNT_GRAV_PP = 5,   // gravity: target x leaf-particle interactions evaluated
  CNT_NBR_CAND = 6,  // neighbour search: target x candidate tests
  CNT_WALK_NODES = 7, // nodes popped by all walks (gravity + neighbour search)
  CNT_COUNT = 8
};
// Counters are striped over CNT_SLOTS cache lines (slot = block index mod CNT_SLOTS) so that per-wave
// atomics do not serialise on one address; readers sum the slots.
constexpr int CNT_SLOTS = 64;
constexpr int CNT_STRIDE = 16;  // 128 bytes per slot
static_assert(CNT_COUNT <= CNT_STRIDE, "counter slot too small");
__device__ __forceinline__ void cnt_add(unsigned long long* counters, int id, unsigned long long v) {
  atomicAdd(&counters[(blockIdx.x & (CNT_SLOTS - 1)) * CNT_STRIDE + id], v);
}

// Buffers of one Morton/sort/Karras/upsweep pipeline (sized for a capacity).
// Per-scenario clock: a batch advances several independent impact scenarios ("segments"), each with its own
// Courant dt. Segment s occupies storage positions [start[s], start[s+1]) in every buffer at all times (the sort
// orders by (segment, Morton code)), so the segment of a storage position never changes.
struct SegClock {
  double dt;                 // dt of the segment's last step
  double t;                  // simulated time
  unsigned long long steps;  // steps taken
  uint32_t signal;           // order-preserving key of min h / max(c + |v|, 1) of the current evaluation
  uint32_t pad;
};

// Per-evaluation resets folded into the first kernel of a force evaluation (k_bbox's block 0).
struct EvalInit {
  uint32_t* flags = nullptr;     // FLAG_H_RETRY and FLAG_TIE_RUNS are cleared (null: nothing is reset)
  SegClock* clk = nullptr;       // every scenario's Courant signal is re-armed
  uint32_t* hs_npend = nullptr;  // the h-solve's pending-list lengths are cleared (n_hs_npend of them)
  int n_hs_npend = 0;
  uint32_t* walk_q = nullptr;    // the gravity walk's 8 per-XCD queue counters are cleared (an evaluation that keeps
                                 // the walk groups, whose build clears them otherwise)
};

// The reset, by the threads of block 0 of the evaluation's first kernel (k_bbox, or k_stage when the sort is kept).
__device__ __forceinline__ void eval_init(const EvalInit& fe, int nseg) {
  if (!fe.flags) return;
  if (threadIdx.x == 0) {
    fe.flags[FLAG_H_RETRY] = 0u;
    fe.flags[FLAG_TIE_RUNS] = 0u;
  }
  for (int q = threadIdx.x; q < nseg; q += blockDim.x) fe.clk[q].signal = 0xffffffffu;
  if (fe.hs_npend && (int)threadIdx.x < fe.n_hs_npend) fe.hs_npend[threadIdx.x] = 0u;
  if (fe.walk_q && threadIdx.x < 8) fe.walk_q[threadIdx.x] = 0u;
}

struct SegView {
  int S = 1;
  const uint32_t* segof = nullptr;  // n: segment of storage position t (null when S == 1)
  const uint32_t* start = nullptr;  // S + 1 storage offsets
  const float* soft2 = nullptr;     // S: softening^2
  const float* hcap = nullptr;      // S: search-radius cap
  SegClock* clk = nullptr;          // S
  uint32_t* bbox = nullptr;         // 6 S order-preserving bbox keys
  const uint32_t* root = nullptr;   // S: the segment tree's root node (walk start)
  __device__ __forceinline__ uint32_t of(int64_t t) const { return segof ? segof[t] : 0u; }
};

struct TreeBufs {
  int64_t cap = 0;   // particle capacity
  int64_t lcap = 0;  // leaf capacity
  int nb_cap = 0;    // radix tile capacity
  int group_max = 64;  // gravity walk-group size limit (K <= group_max <= 2 * GROUP_MAX: up to 2 targets per lane)
  uint32_t* bbox = nullptr;  // 6 order-preserving keys
  uint32_t* bbox_part = nullptr;  // 6 BBOX_PARTS: per-block partial boxes of the bbox pass
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* hist = nullptr;  // 256 * nb
  uint32_t* dtot = nullptr;  // 256
  int2* child = nullptr;     // L-1 (left, right)
  int2* range = nullptr;     // L-1 (first, last leaf)
  uint32_t* parent = nullptr;  // 2L-1
  uint32_t* ready = nullptr;   // L-1
  uint32_t* pend = nullptr;    // 2L: upsweep arrivals at nodes above the upsweep's block roots
  uint32_t* npend = nullptr;   // 1: arrivals recorded in pend
  uint32_t* bstart = nullptr;  // L: the upsweep block root whose first leaf is leaf c (none otherwise): block roots
                               // are the subtrees of at most ub_leaves leaves whose parent holds more
  int ub_leaves = 1024;        // leaf capacity of an upsweep block root (set by launch_karras from K)
  float4* ncom = nullptr;   // 2L-1: com xyz, mass
  float4* nbmin = nullptr;  // 2L-1: bmin xyz, max search radius
  float4* nbmax = nullptr;  // 2L-1: bmax xyz, critical radius |bmax-bmin|/theta + |centre-com|
  float4* nqa = nullptr;    // 2L-1: mass-normalised traceless quadrupole about the com,
                            // sum (m / M) (3 y y^T - |y|^2 I): xx, yy, zz, xy
  float2* nqb = nullptr;    // 2L-1: xz, yz
  int quadrupole = 1;       // the walk adds the cells' quadrupole terms
  float theta_mono = 0.f;   // > 0: cells that pass the opening test at this angle act as monopoles
  float quad_tol = 0.f;     // > 0: cells whose quadrupole term is below quad_tol x their monopole's act as monopoles
  float2* nmac = nullptr;   // 2L-1 (theta_mono or quad_tol set): critical radius, squared distance beyond which an
                            // accepted cell is a monopole
  // walk groups: contiguous Morton ranges of at most group_max particles with compact boxes, in leaf order
  uint32_t* gstart = nullptr;   // L: group root node if leaf c is the first leaf of a group; GROUP_SPLIT if
                                // leaf c is split into single-particle groups; none otherwise
  uint32_t* gblk = nullptr;     // per-1024-leaf block: group count, then exclusive offset
  int4* groups = nullptr;       // (first particle, count, node, 0), up to n entries: neighbour-search groups
  uint32_t* ngroups = nullptr;  // 1
  int4* ggroups = nullptr;      // gravity-walk groups (the same list when group_max == GROUP_MAX)
  uint32_t* nggroups = nullptr; // 1
  uint32_t* wq = nullptr;       // 16 work-queue counters (8 per walk kernel, one per XCD)
  // results of the last sort: sorted keys and the permutation (pointers into keys/vals)
  uint32_t* skeys = nullptr;
  uint32_t* perm = nullptr;
  uint32_t* codes0 = nullptr;  // n: Morton codes by storage position (the segment pass re-gathers them)
  uint32_t* tie_runs = nullptr;  // n / 65 + 1: first sorted positions of the runs of equal codes longer than 64
};

constexpr int BBOX_PARTS = 1024;          // blocks of the bbox pass (their partial boxes fold in one workgroup)
static_assert(IHS_MAX_SEGMENTS <= BBOX_PARTS, "every segment needs a bbox block");
constexpr int NMASK_WORDS = 8;           // support masks cover neighbour lists of up to 256 entries
constexpr int GROUP_MAX = 64;            // one wavefront: the largest walk group
constexpr float GROUP_DIAG_HS = 6.0f;     // a gravity group's box diagonal is at most this many search radii (at
                                          // the full search margin; scaled so the bound stays fixed in units of h)
constexpr float NBR_DIAG_HS = 6.0f;       // the same bound for neighbour-search groups (separate lists if unequal)
constexpr uint32_t GROUP_SPLIT = 0xfffffffeu;

constexpr int RADIX_THREADS = 256;
constexpr int RADIX_ITEMS = 8;
constexpr int RADIX_TILE = RADIX_THREADS * RADIX_ITEMS;
inline int radix_tiles(int64_t n) { return (int)((n + RADIX_TILE - 1) / RADIX_TILE); }

// --- k_lbvh.hip ---
// bbox + Morton codes from positions (PT = float or double, SoA), vals = identity.
// Per-segment bbox (sv.bbox) and codes normalised in each segment's box.
void launch_bbox_morton(hipStream_t s, int split, const void* px, const void* py, const void* pz, int64_t n, int K,
                        TreeBufs& tb, const SegView& sv, const EvalInit& fe = EvalInit{});
// S > 1: stable pass by segment after the Morton passes, so the order is (segment, code); re-gathers skeys.
void launch_segment_order(hipStream_t s, int64_t n, TreeBufs& tb, const SegView& sv);
// Stable LSD radix sort of (keys[0], vals[0]) over `bits` low bits; sets tb.skeys / tb.perm.
void launch_radix_sort(hipStream_t s, int64_t n, int bits, TreeBufs& tb);
// The same stable sort over caller buffers: (keys[cur], vals[cur]) sorted on `bits` low key bits, ping-ponging
// through keys/vals[cur ^ 1]; returns the index of the buffers holding the result. hist holds 256 * radix_tiles(n).
int launch_radix_pairs(hipStream_t s, int64_t n, int bits, uint32_t* keys[2], uint32_t* vals[2], int cur,
                       uint32_t* hist, uint32_t* dtot);
// Reorders runs of equal codes in tb.perm by ids[perm] (so the order is (code, id)).
void launch_fix_ties(hipStream_t s, int64_t n, const uint32_t* ids, TreeBufs& tb, uint32_t* flags);
// Leaf cluster codes + Karras topology + counters reset.
// With segments (sv.segof) the tree is a forest: one Karras tree per segment, rooted at sv.root[s].
void launch_karras(hipStream_t s, int64_t n, int K, TreeBufs& tb, const SegView& sv);
// Walk groups in leaf order -> tb.groups / tb.ngroups, and resets the walk work queues. A group is the
// topmost subtree on a leaf's root path holding <= GROUP_MAX (gravity: tb.group_max) particles whose box
// diagonal is at most
// diag_hs times its largest search radius (no radius: no box limit); leaves without one are split into
// single-particle groups. Needs the upsweep's boxes.
// Groups record their segment's root (int4.w), where their walks start. marked: the upsweep has already marked the
// neighbour-search groups (GROUP_MAX, this diag_hs) in tb.gstart.
void launch_groups(hipStream_t s, int64_t n, int K, TreeBufs& tb, const SegView& sv, const uint32_t* gate = nullptr,
                   float diag_hs = GROUP_DIAG_HS, bool gravity = false, bool marked = false);
// Leaf moments and ordered upsweep; xm = sorted (x,y,z,m), hs = sorted search radius (or null). gate (optional): every
// kernel returns at once while *gate is zero; a gated second upsweep over the same topology needs its arrival
// counters re-armed first (launch_retry_hs does). group_diag_hs > 0: the upsweep
// also marks the neighbour-search groups (GROUP_MAX particles, that box bound) in tb.gstart, for launch_groups
// with marked = true. hs_only: the positions and topology are those of the last full upsweep and only hs changed;
// the nodes' largest search radii (nbmin.w) and the group marks are formed again, their moments stand.
void launch_upsweep(hipStream_t s, int64_t n, int K, float theta, const float4* xm, const float* hs, TreeBufs& tb,
                    const uint32_t* gate = nullptr, float group_diag_hs = 0.f, bool hs_only = false);

// --- k_gravity.hip ---
// Wave-coherent Barnes-Hut group walk over sorted particles; accg = (a/G xyz... see kernel), sorted order.
// Persistent: waves pull groups from per-XCD queues (tb.wq[0..8)).
// Groups come from groups/ngroups (tb.ggroups); pot = false skips the potential (accg.w = 0).
// self: some segment has zero softening (the walk then skips self pairs by index).
// Neighbour lists written by the gravity walk (a walk of the same groups; one target per lane): the walk opens
// every node whose box passes the neighbour search's box test, and its near leaves, taken one particle at a time
// with their indices, are tested r^2 < max(hs_i^2, hs_j^2) as the neighbour build does.
struct NbrIo {
  const float* hs = nullptr;
  const float4* nbmin = nullptr;  // node box min and largest search radius
  uint4* nbr4 = nullptr;
  uint32_t* ncount = nullptr;
  int cap = 0;
  int64_t stride = 0;
};
// After a full upsweep (the moments changed): each node's opening record tb.nmac (critical radius, monopole
// distance from theta_mono and quad_tol; nothing when neither is set).
void launch_mac_record(hipStream_t s, int64_t n, int K, float theta, const TreeBufs& tb);
void launch_gravity_walk(hipStream_t s, int64_t n, int K, int kahan, const float4* xm, const TreeBufs& tb,
                         const int4* groups, const uint32_t* ngroups, const SegView& sv, float4* accg, uint32_t* flags,
                         unsigned long long* counters, bool pot, bool self,
                         const NbrIo* nbr = nullptr, float mac2 = 1.0f);
// FP64 direct-sum accelerations (G included) at the particles with the given ids, from the state's positions
// in storage order (accuracy probe of the walk), each over its own scenario with its softening. inv needs n
// entries, tpos nt.
void launch_direct_f64_targets(hipStream_t s, int split, const void* px, const void* py, const void* pz, const float* m,
                               const uint32_t* id, int64_t n, const uint32_t* ids, int64_t nt, const double* seg_soft2,
                               const uint32_t* segof, const uint32_t* seg_start, uint32_t* inv, uint32_t* tpos,
                               double* acc3);
// Tiled direct O(N^2) sum (FP32), sorted or unsorted, acc4 = (acc xyz, potential).
void launch_gravity_direct(hipStream_t s, int64_t n, float soft2, const float4* xm, float4* acc4);

// --- k_sph.hip ---
// Kernels taking `gate` do nothing when *gate == 0 (gate == null: always run); the wider-radius pass of a
// seeded evaluation is launched unconditionally and gated on FLAG_H_RETRY, so a captured step needs no host check.
void launch_nbr_build(hipStream_t s, int64_t n, int K, const float4* xm, const float* hs, const TreeBufs& tb,
                      uint32_t* nbr, uint32_t* ncount, int cap, int64_t stride, uint32_t* flags,
                      unsigned long long* counters, const uint32_t* gate = nullptr);
// The first pass appends unconverged particles to PEND_SHARDS lists (block b to list b mod PEND_SHARDS, each in its
// own region of pend), so that no single counter takes an atomic from every wave.
constexpr int PEND_SHARDS = 64;
struct HsolveScratch {
  uint32_t* pend = nullptr;    // n + PEND_SHARDS * 256: particles left unconverged by the first pass
  float2* bracket = nullptr;   // n: their (lo, hi) after the first update, by particle
  uint32_t* npend = nullptr;   // PEND_SHARDS list lengths
};
// The SPH passes below take pair displacements from xm alone (xlo == null) or, in split precision, from the hi/lo
// positions xm + xlo (pair_d in common.h).
// retry != null: solves stopped by a search radius below hcap increment *retry instead of FLAG_H_CAPPED.
// The pending-list lengths hw.npend start at zero: the evaluation's reset (eval_init) clears them, and before the
// gated second solve launch_retry_hs does.
void launch_hsolve(hipStream_t s, int64_t n, const float4* xm, const uint32_t* xlo, const float* hs, float* h,
                   uint8_t* iters,
                   const uint32_t* nbr, const uint32_t* ncount, int64_t stride, float eta3, float tol, int max_iter,
                   const HsolveScratch& hw, uint32_t* flags, unsigned long long* counters, const SegView& sv,
                   const uint32_t* gate = nullptr, uint32_t* retry = nullptr);
// Wider-radius pass set-up (gated): hs = min(h * margin, max(hcap, h)); counts the pass in FLAG_RETRIES; re-arms
// the upsweep's arrival counters (ready over the L - 1 internal nodes, npend) and clears the h-solve's n_hs_npend
// pending-list lengths, for the second upsweep and solve over the same topology.
void launch_retry_hs(hipStream_t s, int64_t n, const float* h, float* hs, float margin, const SegView& sv,
                     uint32_t* flags, int64_t L, uint32_t* ready, uint32_t* npend, uint32_t* hs_npend,
                     int n_hs_npend);
// Also reduces each segment's Courant signal into sv.clk[s].signal. vh enters as (v, h_prev) and leaves as (v, h).
// mask (optional, lists of at most NMASK_WORDS * 32 entries): receives each list's support mask, bit k of word
// k / 32 at (k / 32) * stride + t set when entry k lies inside the kernel support r < h_ij.
void launch_density(hipStream_t s, int split, int feat, int64_t n, const float4* xm, const uint32_t* xlo, const float* h,
                    const void* u,
                    const uint8_t* mat, const void* eos, const uint32_t* nbr, const uint32_t* ncount, int64_t stride,
                    float4* rpc, float4* vh, float* rho, float4* featv, uint32_t* mask, uint32_t* flags,
                    const SegView& sv);
// divv_feat != null: also write the velocity divergence into divv_feat[t].y.
void launch_forces(hipStream_t s, int kahan, int64_t n, const float4* xm, const uint32_t* xlo, const float4* vh,
                   const float4* rpc,
                   const uint32_t* nbr, const uint32_t* ncount, const uint32_t* mask, int64_t stride, const float* av,
                   float4* asph, float4* divv_feat = nullptr);
// Velocity-only evaluation at unchanged positions: EOS from the stored density and the current u, rpc / vh for
// the forces, and the segments' Courant signals.
void launch_eos(hipStream_t s, int split, int64_t n, const float* rho, const float* h, const float* vx, const float* vy,
                const float* vz, const void* u, const uint8_t* mat, const void* eos, float4* rpc, float4* vh,
                uint32_t* flags, const SegView& sv);

// --- k_step.hip ---
// Master particle state (SoA, current storage order). Positions and u are double in split precision,
// float otherwise (PT).
struct MasterBufs {
  void* px = nullptr;
  void* py = nullptr;
  void* pz = nullptr;
  float* vx = nullptr;
  float* vy = nullptr;
  float* vz = nullptr;
  void* u = nullptr;
  float* h = nullptr;
  float* m = nullptr;
  uint8_t* mat = nullptr;
  uint32_t* id = nullptr;
  uint32_t* prov = nullptr;
  float* rho = nullptr;
  float4* feat = nullptr;   // MLP inputs of the last density pass: (n_ngb, div v, |grad rho|, -)
  uint8_t* hiter = nullptr; // Newton updates in the last h-solve
};

// Device-resident step control (kernels read it, so a captured step needs no host values).
struct StepCtl {
  double dt_fixed;  // > 0: fixed dt; otherwise cfl * the segment's min signal
  double cfl;
  float av[2];      // artificial viscosity alpha, beta
  float damp;       // relaxation damping
  float omega;      // relaxation frame rotation rate about +z
};

// 48-byte interchange record (gpu_sph.rs:25-37).
struct SphParticle48 {
  float pos[3];
  float h;
  float vel[3];
  float u;
  float mass;
  uint32_t mat;
  float rho;
  uint32_t prov;
};

void launch_fe_init(hipStream_t s, uint32_t* flags, const SegView& sv);
// Permutes the state into the new order and writes the working copies xm (x, m), vh (v, h) and hs; with xlo (split
// precision) also each position's remainder below its FP32 rounding (lo_code in common.h).
void launch_gather(hipStream_t s, int split, int64_t n, const uint32_t* perm, const MasterBufs& src,
                   const MasterBufs& dst, float4* xm, float4* vh, float* hs, float h_margin, const SegView& sv,
                   uint32_t* xlo = nullptr);
// The same working copies in the state's own order (k_stage: positions and order unchanged since the last sort),
// with the evaluation's reset fe and the re-arming of the upsweep's arrival counters (ready, npend) over L leaves.
void launch_stage(hipStream_t s, int split, int64_t n, const MasterBufs& st, float4* xm, float4* vh, float* hs,
                  float h_margin, const SegView& sv, uint32_t* xlo, const EvalInit& fe, int64_t L, uint32_t* ready,
                  uint32_t* npend);
void launch_kick_drift(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph,
                       const float4* accg, const StepCtl* ctl, const SegView& sv);
void launch_kick(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph, const float4* accg,
                 const SegView& sv);
// Largest provenance over storage positions [t0, t1) into *out (atomicMax; the caller clears it).
void launch_prov_max(hipStream_t s, int64_t t0, int64_t t1, const MasterBufs& st, uint32_t* out);
// In-place assembly of storage positions [t0, t1) (gpu_sph.rs:318-339): pl = 12 doubles per provenance (centre,
// offset, velocity, spin); *minh receives the order-preserving key of the smallest h (the caller sets it to ~0).
void launch_assemble(hipStream_t s, int split, int64_t t0, int64_t t1, const MasterBufs& st, const double* pl,
                     uint32_t* minh);
void launch_relax(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* asph, const float4* accg,
                  const StepCtl* ctl, const SegView& sv);
// FP64 KE, IE, PE, p, L (9 values) into out; part needs 9 * 1024 doubles.
void launch_energy(hipStream_t s, int split, int64_t n, const MasterBufs& st, const float4* accg, double* part,
                   double* out);
void launch_unpack(hipStream_t s, int split, int64_t n, const SphParticle48* in, const MasterBufs& st);
void launch_pack(hipStream_t s, int split, int64_t n, const MasterBufs& st, SphParticle48* out);
// Scatter a sorted-order array to id order (elem = 4 or 16 bytes).
void launch_scatter_by_id(hipStream_t s, int64_t n, const uint32_t* id, const void* src, void* dst, int elem);
// Double-precision positions (split) or float positions widened -> (n,3) double in id order.
void launch_pos_f64(hipStream_t s, int split, int64_t n, const MasterBufs& st, double* out3);
// Standalone gravity helpers (k_gravity.hip).
void launch_bodies_in(hipStream_t s, const double* pos3, const double* mass, int64_t n, double* px, double* py,
                      double* pz, float* m);
void launch_gather_xm(hipStream_t s, const uint32_t* perm, int64_t n, const double* px, const double* py,
                      const double* pz, const float* m, float4* xm);
void launch_acc_out(hipStream_t s, const uint32_t* perm, int64_t n, const float4* acc4, double* acc3, double* pot);
