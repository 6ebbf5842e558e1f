// C ABI of libihs_sph: handle lifetime, upload/readback, the force-evaluation pipeline, the KDK step
// (captured in a hipGraph), relaxation, standalone gravity and inspection. Mirrors GpuSph
// (ie/crates/engine/src/gpu_sph.rs:742-922) and GravityField (ie/crates/engine/src/gpu_gravity.rs:532-626).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ihs_sph.h"
#include "launch.h"
#include "measure.h"
#include "mlp.h"

namespace {

constexpr float kPi = 3.14159265358979f;

// Scratch of the allocation-free probe ihs_direct_f64_noalloc: per call, the target ids, their storage positions
// and the FP64 accelerations, for at most kProbeMaxTargets targets (2 MiB per GPU, held until the process ends).
// In host code `scratch` is a host-side stand-in, never a device address: each device's copy is looked up once with
// hipGetSymbolAddress (device_scratch) and only that address is passed to kernels and copies.
constexpr int64_t kProbeMaxTargets = 65536;
constexpr size_t kProbeIdsOff = 0;
constexpr size_t kProbeTposOff = 4 * kProbeMaxTargets;
constexpr size_t kProbeAccOff = 8 * kProbeMaxTargets;
constexpr size_t kProbeBytes = kProbeAccOff + 24 * kProbeMaxTargets;
static __device__ __align__(256) unsigned char scratch[kProbeBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {nullptr};

// The device address of `scratch` on the current device `device` (looked up once per device).
hipError_t device_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

struct Scratch {
  // sorted-order compute arrays
  float4* xm = nullptr;
  uint32_t* xlo = nullptr;  // split precision: each position's remainder below xm (lo_code)
  float* hs = nullptr;
  float4* rpc = nullptr;
  float4* vh = nullptr;
  float4* asph = nullptr;
  float4* accg = nullptr;
  float4* apot = nullptr;  // the potential walk's output (an energy read after steps that skipped the potential)
  uint32_t* nbr = nullptr;
  uint32_t* ncount = nullptr;
  uint32_t* nmask = nullptr;  // support masks of the lists (null when lists exceed NMASK_WORDS * 32 entries)
  uint32_t* flags = nullptr;
  unsigned long long* counters = nullptr;
  unsigned long long* counters_aux = nullptr;  // interaction counters of the potential walk (not reported)
  double* epart = nullptr;
  double* eout = nullptr;
  StepCtl* ctl = nullptr;
  void* eos = nullptr;
  // segments (scenarios of a batch; one for a single upload)
  uint32_t* segof = nullptr;    // capacity
  uint32_t* seg_start = nullptr;  // IHS_MAX_SEGMENTS + 1
  float* seg_soft2 = nullptr;
  double* seg_soft2d = nullptr;  // the same in FP64, for the direct-sum probe
  float* seg_hcap = nullptr;
  SegClock* seg_clk = nullptr;
  uint32_t* seg_bbox = nullptr;   // 6 IHS_MAX_SEGMENTS
  uint32_t* seg_root = nullptr;
  // the standalone gravity call's single segment
  uint32_t* g_start = nullptr;  // 2
  float* g_soft2 = nullptr;
  float* g_hcap = nullptr;
  SegClock* g_clk = nullptr;
  uint32_t* g_root = nullptr;
  ihs_particle* staging = nullptr;  // device AoS for upload/readback
  // standalone gravity
  double* gpos = nullptr;
  double* gpx = nullptr;
  double* gpy = nullptr;
  double* gpz = nullptr;
  float* gm = nullptr;
  double* gmass = nullptr;
  double* gacc = nullptr;
  double* gpot = nullptr;
};

}  // namespace

struct ihs_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  ihs_config cfg{};
  int64_t cap = 0, n = 0;
  int split = 0;
  int cur = 0;
  MasterBufs st[2];
  TreeBufs tb;
  Scratch sc;
  MlpDev mlp;
  int nseg = 1;
  int n_mat = 0;                        // EOS records uploaded (materials index below this)
  std::vector<int64_t> seg_start_host;  // nseg + 1
  bool zero_soft = false;               // some scenario has no softening: the walk skips self pairs by index
  int64_t force_evals = 0;
  int64_t list_builds = 0;  // evaluations that built neighbour lists (the mean neighbour count's denominator)
  bool have_fe = false;
  bool pot_fresh = false;  // accg.w holds the potential of the last force evaluation
  // The last force evaluation solved h at the current positions (nothing has drifted since), so the next
  // evaluation's search radius only needs room for round-off.
  bool h_settled = false;
  // The last full force evaluation saw the current positions (no drift since): its tree, neighbour lists,
  // density and gravity still hold.
  bool pos_fresh = false;
  // tb holds the sort and topology of st[cur]'s positions, in st[cur]'s order (set by every force evaluation,
  // cleared by whatever moves, reorders or replaces the particles or builds another tree in tb).
  bool tree_fresh = false;
  HsolveScratch hw;
  // One captured KDK step per starting parity of the ping-pong state buffers.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  int graph_next[2] = {0, 1};  // the buffer a captured step leaves the state in
  int graph_builds[2] = {2, 2};  // list-building evaluations in a captured step
  // stage profiling
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_stage;
  size_t prof_used = 0;
  double prof_ms[IHS_STAGE_COUNT] = {0};
  int64_t prof_calls[IHS_STAGE_COUNT] = {0};
  std::vector<void*> allocs;
  MeasureBufs mb;  // disk / remnant measurement and clump buffers, allocated on first use
  // Two-phase readback (gpu_store.rs:126-152): ihs_begin_readback packs the state as the stream reaches it into
  // snap_dev and copies it to the pinned snap_host, recording snap_ev; ihs_take_readback hands over that snapshot
  // once the event has completed, whatever was queued after it. Allocated on the first begin.
  void* probe_scratch = nullptr;  // this device's copy of `scratch` (set at ihs_create)
  SphParticle48* snap_dev = nullptr;
  ihs_particle* snap_host = nullptr;
  hipEvent_t snap_ev = nullptr;
  int64_t snap_n = 0;
  bool snap_pending = false;
  std::string err;
};

namespace {

int fail(ihs_handle* h, int code, const char* fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIP_TRY(h, expr)                                                                    \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail((h), IHS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(ihs_handle* h, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    h->allocs.push_back(q);
    *p = (T*)q;
  }
  return e;
}

hipError_t alloc_raw(ihs_handle* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
  if (e == hipSuccess) h->allocs.push_back(*p);
  return e;
}

int alloc_all(ihs_handle* h) {
  const int64_t C = h->cap;
  const int K = h->cfg.leaf_size;
  const int64_t Lc = (C + K - 1) / K + 1;
  for (int b = 0; b < 2; ++b) {
    MasterBufs& m = h->st[b];
    HIP_TRY(h, alloc_raw(h, &m.px, 8 * C));
    HIP_TRY(h, alloc_raw(h, &m.py, 8 * C));
    HIP_TRY(h, alloc_raw(h, &m.pz, 8 * C));
    HIP_TRY(h, alloc_raw(h, &m.u, 8 * C));
    HIP_TRY(h, dalloc(h, &m.vx, C));
    HIP_TRY(h, dalloc(h, &m.vy, C));
    HIP_TRY(h, dalloc(h, &m.vz, C));
    HIP_TRY(h, dalloc(h, &m.h, C));
    HIP_TRY(h, dalloc(h, &m.m, C));
    HIP_TRY(h, dalloc(h, &m.mat, C));
    HIP_TRY(h, dalloc(h, &m.id, C));
    HIP_TRY(h, dalloc(h, &m.prov, C));
    HIP_TRY(h, dalloc(h, &m.rho, C));
    HIP_TRY(h, dalloc(h, &m.feat, C));
    HIP_TRY(h, dalloc(h, &m.hiter, C));
  }
  TreeBufs& t = h->tb;
  t.cap = C;
  t.lcap = Lc;
  t.nb_cap = radix_tiles(C);
  t.group_max = h->cfg.group_max;
  HIP_TRY(h, dalloc(h, &t.bbox, 8));
  HIP_TRY(h, dalloc(h, &t.bbox_part, 6 * BBOX_PARTS));
  for (int b = 0; b < 2; ++b) {
    HIP_TRY(h, dalloc(h, &t.keys[b], C));
    HIP_TRY(h, dalloc(h, &t.vals[b], C));
  }
  HIP_TRY(h, dalloc(h, &t.hist, (size_t)256 * t.nb_cap));
  HIP_TRY(h, dalloc(h, &t.dtot, 256));
  HIP_TRY(h, dalloc(h, &t.child, Lc));
  HIP_TRY(h, dalloc(h, &t.range, Lc));
  HIP_TRY(h, dalloc(h, &t.parent, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.ready, Lc));
  HIP_TRY(h, dalloc(h, &t.pend, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.npend, 1));
  HIP_TRY(h, dalloc(h, &t.bstart, Lc));
  HIP_TRY(h, dalloc(h, &t.ncom, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.nbmin, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.nbmax, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.nqa, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.nqb, 2 * Lc));
  t.quadrupole = h->cfg.quadrupole;
  t.theta_mono = h->cfg.theta_mono;
  t.quad_tol = h->cfg.quad_tol;
  if (t.quadrupole && (t.theta_mono > 0.f || t.quad_tol > 0.f)) HIP_TRY(h, dalloc(h, &t.nmac, 2 * Lc));
  HIP_TRY(h, dalloc(h, &t.gstart, Lc));
  HIP_TRY(h, dalloc(h, &t.gblk, Lc / 1024 + 2));
  HIP_TRY(h, dalloc(h, &t.groups, C + 1));
  HIP_TRY(h, dalloc(h, &t.ngroups, 1));
  if (t.group_max != GROUP_MAX || NBR_DIAG_HS != GROUP_DIAG_HS) {
    HIP_TRY(h, dalloc(h, &t.ggroups, C + 1));
    HIP_TRY(h, dalloc(h, &t.nggroups, 1));
  } else {
    t.ggroups = t.groups;
    t.nggroups = t.ngroups;
  }
  HIP_TRY(h, dalloc(h, &t.wq, 16));
  Scratch& s = h->sc;
  HIP_TRY(h, dalloc(h, &s.xm, C));
  if (h->split) HIP_TRY(h, dalloc(h, &s.xlo, C));
  HIP_TRY(h, dalloc(h, &s.hs, C));
  HIP_TRY(h, dalloc(h, &s.rpc, C));
  HIP_TRY(h, dalloc(h, &s.vh, C));
  HIP_TRY(h, dalloc(h, &s.asph, C));
  HIP_TRY(h, dalloc(h, &s.accg, C));
  HIP_TRY(h, dalloc(h, &s.apot, C));
  HIP_TRY(h, dalloc(h, &s.nbr, (size_t)C * h->cfg.nbr_capacity));
  HIP_TRY(h, dalloc(h, &s.ncount, C));
  if (h->cfg.nbr_capacity <= NMASK_WORDS * 32) HIP_TRY(h, dalloc(h, &s.nmask, (size_t)NMASK_WORDS * C));
  HIP_TRY(h, dalloc(h, &s.flags, FLAG_COUNT));
  HIP_TRY(h, dalloc(h, &s.counters, CNT_SLOTS * CNT_STRIDE));
  HIP_TRY(h, dalloc(h, &s.counters_aux, CNT_SLOTS * CNT_STRIDE));
  HIP_TRY(h, dalloc(h, &s.epart, 9 * 1024));
  HIP_TRY(h, dalloc(h, &s.eout, 16));
  HIP_TRY(h, dalloc(h, &s.ctl, 1));
  HIP_TRY(h, dalloc(h, &s.segof, C));
  HIP_TRY(h, dalloc(h, &s.seg_start, IHS_MAX_SEGMENTS + 1));
  HIP_TRY(h, dalloc(h, &s.seg_soft2, IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.seg_soft2d, IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.seg_hcap, IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.seg_clk, IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.seg_bbox, 6 * IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.seg_root, IHS_MAX_SEGMENTS));
  HIP_TRY(h, dalloc(h, &s.g_start, 2));
  HIP_TRY(h, dalloc(h, &s.g_soft2, 1));
  HIP_TRY(h, dalloc(h, &s.g_hcap, 1));
  HIP_TRY(h, dalloc(h, &s.g_clk, 1));
  HIP_TRY(h, dalloc(h, &s.g_root, 1));
  HIP_TRY(h, dalloc(h, &t.codes0, C));
  HIP_TRY(h, dalloc(h, &t.tie_runs, C / 65 + 1));
  HIP_TRY(h, alloc_raw(h, &s.eos, 16 * sizeof(ihs_eos)));
  HIP_TRY(h, dalloc(h, &s.staging, C));
  HIP_TRY(h, dalloc(h, &h->hw.pend, C + PEND_SHARDS * 256));
  HIP_TRY(h, dalloc(h, &h->hw.bracket, C));
  HIP_TRY(h, dalloc(h, &h->hw.npend, PEND_SHARDS));
  HIP_TRY(h, hipMemset(s.flags, 0, FLAG_COUNT * sizeof(uint32_t)));
  HIP_TRY(h, hipMemset(s.counters, 0, CNT_SLOTS * CNT_STRIDE * sizeof(unsigned long long)));
  HIP_TRY(h, hipMemset(s.nbr, 0, (size_t)C * h->cfg.nbr_capacity * sizeof(uint32_t)));
  HIP_TRY(h, hipMemset(s.ncount, 0, C * sizeof(uint32_t)));
  StepCtl ctl{};
  ctl.dt_fixed = 0.0;
  ctl.cfl = h->cfg.cfl;
  ctl.av[0] = 1.0f;
  ctl.av[1] = 2.0f;
  ctl.damp = 1.0f;
  HIP_TRY(h, hipMemcpy(s.ctl, &ctl, sizeof ctl, hipMemcpyHostToDevice));
  if (mlp_alloc(h->mlp, C) != 0) return fail(h, IHS_ERR_HIP, "MLP buffer allocation failed");
  return IHS_OK;
}

// Records an event ending `stage` (stage < 0: a start marker) when profiling is on.
void prof_mark(ihs_handle* h, int stage) {
  if (!h->prof_on) return;
  if (h->prof_used == h->prof_ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    h->prof_ev.push_back(e);
    h->prof_stage.push_back(-1);
  }
  hipEventRecord(h->prof_ev[h->prof_used], h->stream);
  h->prof_stage[h->prof_used] = stage;
  h->prof_used++;
}

void prof_flush(ihs_handle* h) {
  if (!h->prof_on || h->prof_used == 0) return;
  hipStreamSynchronize(h->stream);
  for (size_t i = 1; i < h->prof_used; ++i) {
    const int st = h->prof_stage[i];
    if (st < 0) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->prof_ev[i - 1], h->prof_ev[i]) == hipSuccess) {
      h->prof_ms[st] += ms;
      h->prof_calls[st] += 1;
    }
  }
  h->prof_used = 0;
}

bool valid_leaf(int k) { return k >= 1 && k <= 64 && (k & (k - 1)) == 0; }

SegView seg_view(ihs_handle* h) {
  SegView v;
  v.S = h->nseg;
  v.segof = h->nseg > 1 ? h->sc.segof : nullptr;
  v.start = h->sc.seg_start;
  v.soft2 = h->sc.seg_soft2;
  v.hcap = h->sc.seg_hcap;
  v.clk = h->sc.seg_clk;
  v.bbox = h->sc.seg_bbox;
  v.root = h->sc.seg_root;
  return v;
}

// Search-radius margin of an evaluation whose h was solved at unchanged positions: room for round-off only.
constexpr float kMarginSettled = 1.02f;

// One force evaluation: sort, permute, tree, neighbour lists, h-solve, density/EOS, forces, gravity.
// after_drift: the positions moved since the previous evaluation (the second evaluation of a KDK step). Then a
// trained network seeds the h-solve (LEARNED) and sets a tight search radius around its seed, with a wider-radius
// pass that runs only if some solve reached that radius; a recording run stores (features, converged h) rows.
// At positions the previous evaluation sorted, none moved since (the first evaluation of a KDK step: the kick
// changed only v and u; or a repeated evaluation), the sort would return the identity and the Karras pass the same
// topology: both are kept and the state stays in its buffers (k_stage), with results identical to sorting again.
void force_eval(ihs_handle* h, bool after_drift, bool need_pot = true) {
  hipStream_t s = h->stream;
  const int64_t n = h->n;
  const int K = h->cfg.leaf_size;
  const bool sorted = h->tree_fresh && !after_drift;
  const bool newton_mode = h->cfg.h_mode != IHS_H_FIXED;
  // At positions unchanged since the last full evaluation and with no solve to run, the neighbour lists and walk
  // groups that evaluation left are the ones a rebuild would find (the same positions and h): with theta_settled set
  // they are kept, and gravity takes a walk that builds no lists, at theta_settled.
  const bool keep_lists = sorted && h->cfg.theta_settled > 0.f && !(newton_mode && !h->h_settled);
  MasterBufs& S = h->st[h->cur];
  MasterBufs& D = sorted ? S : h->st[h->cur ^ 1];
  Scratch& sc = h->sc;
  const bool kahan = h->split != 0;
  const bool newton = h->cfg.h_mode != IHS_H_FIXED;
  const bool seeded = h->cfg.h_mode == IHS_H_LEARNED && after_drift && h->mlp.trained;
  const bool record = h->mlp.record && after_drift && h->cfg.h_mode == IHS_H_NEWTON;
  const float eta = std::cbrt(3.0f * h->cfg.n_ngb / (4.0f * kPi));
  const SegView sv = seg_view(h);
  prof_mark(h, -1);
  EvalInit fe;
  fe.flags = sc.flags;
  fe.clk = sv.clk;
  fe.hs_npend = h->hw.npend;
  fe.n_hs_npend = PEND_SHARDS;
  if (keep_lists) fe.walk_q = h->tb.wq;
  float margin = 1.0f;
  if (newton) margin = (!after_drift && h->h_settled) ? kMarginSettled : h->cfg.h_margin;
  if (sorted) {
    launch_stage(s, h->split, n, S, sc.xm, sc.vh, sc.hs, margin, sv, sc.xlo, fe, (n + K - 1) / K, h->tb.ready,
                 h->tb.npend);
  } else {
    launch_bbox_morton(s, h->split, S.px, S.py, S.pz, n, K, h->tb, sv, fe);
    launch_radix_sort(s, n, 30, h->tb);
    launch_segment_order(s, n, h->tb, sv);
    launch_fix_ties(s, n, S.id, h->tb, sc.flags);
    prof_mark(h, IHS_STAGE_SORT);
    launch_gather(s, h->split, n, h->tb.perm, S, D, sc.xm, sc.vh, sc.hs, margin, sv, sc.xlo);
    h->cur ^= 1;
  }
  prof_mark(h, IHS_STAGE_GATHER);
  if (seeded) {
    mlp_seed_h(s, h->mlp, n, D, sc.hs, sc.eos, sv, eta, h->cfg.h_margin_seeded);
    prof_mark(h, IHS_STAGE_HSOLVE);
  }
  // Walk groups are bounded in units of h, whatever margin set this evaluation's search radius.
  const float eff_margin = seeded ? h->cfg.h_margin_seeded : margin;
  const float diag_scale = newton ? h->cfg.h_margin / eff_margin : 1.0f;
  if (!sorted) launch_karras(s, n, K, h->tb, sv);
  if (!keep_lists) {
    // Kept sort: the node moments in memory are those of these positions (the last full upsweep's); only hs changed.
    launch_upsweep(s, n, K, h->cfg.theta, sc.xm, sc.hs, h->tb, nullptr, NBR_DIAG_HS * diag_scale, sorted);
    launch_groups(s, n, K, h->tb, sv, nullptr, NBR_DIAG_HS * diag_scale, false, true);
    if (h->tb.ggroups != h->tb.groups) launch_groups(s, n, K, h->tb, sv, nullptr, GROUP_DIAG_HS * diag_scale, true);
  }
  prof_mark(h, IHS_STAGE_TREE);
  // One traversal for gravity and the neighbour lists when the walk and neighbour groups coincide: the walk opens
  // the nodes the neighbour search needs and writes the lists (gravity does not depend on h, so it runs first).
  const bool fused = !keep_lists && h->cfg.fused_walk && h->tb.ggroups == h->tb.groups && h->tb.group_max <= GROUP_MAX;
  if (keep_lists) {
    const float r = h->cfg.theta / h->cfg.theta_settled;
    launch_gravity_walk(s, n, K, kahan, sc.xm, h->tb, h->tb.ggroups, h->tb.nggroups, sv, sc.accg, sc.flags,
                        sc.counters, need_pot, h->zero_soft, nullptr, r * r);
    prof_mark(h, IHS_STAGE_GRAVITY);
  } else if (fused) {
    NbrIo io;
    io.hs = sc.hs;
    io.nbmin = h->tb.nbmin;
    io.nbr4 = (uint4*)sc.nbr;
    io.ncount = sc.ncount;
    io.cap = h->cfg.nbr_capacity;
    io.stride = n;
    launch_gravity_walk(s, n, K, kahan, sc.xm, h->tb, h->tb.ggroups, h->tb.nggroups, sv, sc.accg, sc.flags,
                        sc.counters, need_pot, h->zero_soft, &io);
    prof_mark(h, IHS_STAGE_GRAVITY);
  } else {
    launch_nbr_build(s, n, K, sc.xm, sc.hs, h->tb, sc.nbr, sc.ncount, h->cfg.nbr_capacity, n, sc.flags, sc.counters);
    prof_mark(h, IHS_STAGE_NBR);
  }
  if (!keep_lists) h->list_builds++;
  // At positions unchanged since the last solve (the first evaluation of a KDK step: the kick changed only v and
  // u) every h already meets the tolerance, and a solve would only confirm it: it is skipped. A particle the last
  // solve left capped or unconverged is solved again at the next evaluation after a drift.
  const bool solve = newton && !(h->h_settled && !after_drift);  // never with keep_lists
  if (solve) {
    const float eta3 = eta * eta * eta;
    if (record) mlp_record_features(s, h->mlp, n, D, sc.eos, sv, eta);
    launch_hsolve(s, n, sc.xm, sc.xlo, sc.hs, D.h, D.hiter, sc.nbr, sc.ncount, n, eta3, h->cfg.h_tol,
                  h->cfg.h_max_iter, h->hw, sc.flags, sc.counters, sv, nullptr,
                  seeded ? sc.flags + FLAG_H_RETRY : nullptr);  // the evaluation's reset cleared the pending lists
    if (record) mlp_record_targets(s, h->mlp, n, D);
    prof_mark(h, IHS_STAGE_HSOLVE);
    if (seeded) {
      // Wider-radius pass, gated on the device: every kernel returns at once unless a solve hit the tight radius.
      const uint32_t* gate = sc.flags + FLAG_H_RETRY;
      launch_retry_hs(s, n, D.h, sc.hs, h->cfg.h_margin, sv, sc.flags, (n + K - 1) / K, h->tb.ready, h->tb.npend,
                      h->hw.npend, PEND_SHARDS);
      launch_upsweep(s, n, K, h->cfg.theta, sc.xm, sc.hs, h->tb, gate, NBR_DIAG_HS, true);  // only hs grew
      launch_groups(s, n, K, h->tb, sv, gate, NBR_DIAG_HS, false, true);
      launch_nbr_build(s, n, K, sc.xm, sc.hs, h->tb, sc.nbr, sc.ncount, h->cfg.nbr_capacity, n, sc.flags,
                       sc.counters, gate);
      launch_hsolve(s, n, sc.xm, sc.xlo, sc.hs, D.h, D.hiter, sc.nbr, sc.ncount, n, eta3, h->cfg.h_tol,
                    h->cfg.h_max_iter, h->hw, sc.flags, sc.counters, sv, gate, nullptr);
      prof_mark(h, IHS_STAGE_HSOLVE);
    }
    h->h_settled = true;
  }
  // The network's inputs come from the evaluation before a drift. With reuse_positions that evaluation is
  // velocity-only (it refreshes div v), so the position-dependent inputs are taken here, at the same positions.
  const int feat = (h->cfg.h_mode == IHS_H_LEARNED || h->mlp.record) && (!after_drift || h->cfg.reuse_positions);
  launch_density(s, h->split, feat, n, sc.xm, sc.xlo, D.h, D.u, D.mat, sc.eos, sc.nbr, sc.ncount, n, sc.rpc, sc.vh, D.rho,
                 D.feat, sc.nmask, sc.flags, sv);
  prof_mark(h, IHS_STAGE_DENSITY);
  launch_forces(s, kahan, n, sc.xm, sc.xlo, sc.vh, sc.rpc, sc.nbr, sc.ncount, sc.nmask, n,
                (const float*)((const char*)sc.ctl + offsetof(StepCtl, av)), sc.asph, feat ? D.feat : nullptr);
  prof_mark(h, IHS_STAGE_FORCES);
  if (!fused && !keep_lists) {
    launch_gravity_walk(s, n, K, kahan, sc.xm, h->tb, h->tb.ggroups, h->tb.nggroups, sv, sc.accg, sc.flags,
                        sc.counters, need_pot, h->zero_soft);
    prof_mark(h, IHS_STAGE_GRAVITY);
  }
  h->force_evals++;
  h->have_fe = true;
  h->pot_fresh = need_pot;
  h->pos_fresh = true;
  h->tree_fresh = true;
}

// Velocity-only force evaluation (configuration reuse_positions): the first evaluation of a KDK step sees the
// positions of the previous step's second evaluation, so its sort, tree, neighbour lists, h, density and gravity
// are those of that evaluation; only the EOS (u was kicked) and the velocity-dependent SPH forces are recomputed.
void force_eval_velocity(ihs_handle* h) {
  hipStream_t s = h->stream;
  const int64_t n = h->n;
  MasterBufs& S = h->st[h->cur];
  Scratch& sc = h->sc;
  const SegView sv = seg_view(h);
  const bool learned = h->cfg.h_mode == IHS_H_LEARNED;
  prof_mark(h, -1);
  launch_fe_init(s, sc.flags, sv);
  launch_eos(s, h->split, n, S.rho, S.h, S.vx, S.vy, S.vz, S.u, S.mat, sc.eos, sc.rpc, sc.vh, sc.flags, sv);
  prof_mark(h, IHS_STAGE_DENSITY);
  launch_forces(s, h->split != 0, n, sc.xm, sc.xlo, sc.vh, sc.rpc, sc.nbr, sc.ncount, sc.nmask, n,
                (const float*)((const char*)sc.ctl + offsetof(StepCtl, av)), sc.asph, learned ? S.feat : nullptr);
  prof_mark(h, IHS_STAGE_FORCES);
  h->force_evals++;
}

void kdk_step(ihs_handle* h) {
  if (h->cfg.reuse_positions && h->pos_fresh) force_eval_velocity(h);
  else force_eval(h, false, false);
  launch_kick_drift(h->stream, h->split, h->n, h->st[h->cur], h->sc.asph, h->sc.accg, h->sc.ctl, seg_view(h));
  h->h_settled = false;
  h->pos_fresh = false;
  h->tree_fresh = false;
  prof_mark(h, IHS_STAGE_KICK);
  // The step needs accelerations only; the potential is a diagnostic, summed when an energy is read (potential()).
  force_eval(h, true, false);
  launch_kick(h->stream, h->split, h->n, h->st[h->cur], h->sc.asph, h->sc.accg, seg_view(h));
  prof_mark(h, IHS_STAGE_KICK);
  prof_flush(h);
}

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// Copies host floats to a temporary device buffer; returns nullptr and sets the error on failure.
float* upload_tmp(ihs_handle* h, const float* src, size_t floats) {
  float* d = nullptr;
  if (hipMalloc(&d, std::max<size_t>(floats, 1) * sizeof(float)) != hipSuccess) {
    fail(h, IHS_ERR_HIP, "temporary allocation of %zu floats failed", floats);
    return nullptr;
  }
  if (hipMemcpy(d, src, floats * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(d);
    fail(h, IHS_ERR_HIP, "upload of %zu floats failed", floats);
    return nullptr;
  }
  return d;
}

void drop_graph(ihs_handle* h) {
  for (int p = 0; p < 2; ++p) {
    if (h->graph_exec[p]) hipGraphExecDestroy(h->graph_exec[p]);
    if (h->graph[p]) hipGraphDestroy(h->graph[p]);
    h->graph_exec[p] = nullptr;
    h->graph[p] = nullptr;
  }
}

}  // namespace

extern "C" {

int ihs_abi_version(void) { return IHS_ABI_VERSION; }

void ihs_default_config(ihs_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->capacity = 1 << 20;
  c->leaf_size = 16;
  c->precision = IHS_PREC_FP32;
  c->h_mode = IHS_H_NEWTON;
  c->theta = 0.5f;
  c->cfl = 0.1f;
  c->n_ngb = 48.0f;
  c->h_tol = 1.0e-4f;
  c->h_max_iter = 40;
  c->h_margin = 1.25f;
  c->h_cap_factor = 4.0f;
  c->nbr_capacity = 192;
  c->use_graph = 1;
  c->h_margin_seeded = 1.03f;
  c->group_max = 64;
  c->quadrupole = 1;
  c->reuse_positions = 0;
  c->fused_walk = 1;
  c->theta_mono = 0.0f;
  c->quad_tol = 0.0f;
  c->theta_settled = 0.0f;
}

int ihs_create(int device, const ihs_config* cfg, ihs_handle** out) {
  if (!out) return IHS_ERR_ARG;
  *out = nullptr;
  ihs_config c;
  if (cfg) c = *cfg;
  else ihs_default_config(&c);
  if (c.capacity < 1 || !valid_leaf(c.leaf_size) || c.nbr_capacity < 1 || c.h_mode < 0 || c.h_mode > 2 ||
      c.precision < 0 || c.precision > 1 || !(c.theta >= 0.f) || !(c.h_margin >= 1.0f) ||
      !(c.h_margin_seeded >= 1.0f) || c.group_max < c.leaf_size || c.group_max > 2 * GROUP_MAX ||
      c.quadrupole < 0 || c.quadrupole > 1 || c.reuse_positions < 0 || c.reuse_positions > 1 || c.fused_walk < 0 ||
      c.fused_walk > 1 || !(c.theta_mono >= 0.f) || (c.theta_mono > 0.f && !(c.theta_mono <= c.theta)) ||
      !(c.quad_tol >= 0.f) || !(c.theta_settled >= 0.f) || c.theta_settled > c.theta)
    return IHS_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IHS_ERR_HIP;
  if (device < 0 || device >= ndev) return IHS_ERR_ARG;
  c.nbr_capacity = (c.nbr_capacity + 3) & ~3;  // lists are stored in 16-byte chunks of four
  ihs_handle* h = new ihs_handle();
  h->device = device;
  h->cfg = c;
  h->cap = c.capacity;
  h->split = c.precision == IHS_PREC_SPLIT;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return IHS_ERR_HIP;
  }
  if (alloc_all(h) != IHS_OK || device_scratch(device, &h->probe_scratch) != hipSuccess) {
    ihs_destroy(h);
    return IHS_ERR_HIP;
  }
  *out = h;
  return IHS_OK;
}

int ihs_destroy(ihs_handle* h) {
  if (!h) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graph(h);
  mlp_free(h->mlp);
  measure_release(h->mb);
  if (h->snap_host) hipHostFree(h->snap_host);
  if (h->snap_ev) hipEventDestroy(h->snap_ev);
  for (hipEvent_t e : h->prof_ev) hipEventDestroy(e);
  for (void* p : h->allocs) hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return IHS_OK;
}

const char* ihs_last_error(const ihs_handle* h) { return h ? h->err.c_str() : "null handle"; }

int ihs_upload_batch(ihs_handle* h, const ihs_particle* p, const int64_t* counts, int32_t nseg, const ihs_eos* eos,
                     int32_t n_mat, const float* softening) {
  if (!h) return IHS_ERR_ARG;
  if (!p || !counts || !softening || nseg < 1 || nseg > IHS_MAX_SEGMENTS)
    return fail(h, IHS_ERR_ARG, "a batch needs 1..%d scenarios", IHS_MAX_SEGMENTS);
  if (!eos || n_mat < 1 || n_mat > 16) return fail(h, IHS_ERR_ARG, "need 1..16 EOS records, got %d", n_mat);
  const int K = h->cfg.leaf_size;
  std::vector<int64_t> start(nseg + 1, 0);
  for (int sg = 0; sg < nseg; ++sg) {
    if (counts[sg] < 1) return fail(h, IHS_ERR_ARG, "scenario %d has no particles", sg);
    if (nseg > 1 && counts[sg] % K != 0)
      return fail(h, IHS_ERR_ARG, "scenario %d: %lld particles is not a multiple of the leaf size %d (a leaf may not "
                  "straddle two scenarios)", sg, (long long)counts[sg], K);
    if (!(softening[sg] >= 0.f)) return fail(h, IHS_ERR_ARG, "scenario %d has negative softening", sg);
    start[sg + 1] = start[sg] + counts[sg];
  }
  const int64_t n = start[nseg];
  if (n > h->cap) return fail(h, IHS_ERR_CAPACITY, "upload of %lld particles exceeds capacity %lld", (long long)n,
                              (long long)h->cap);
  std::vector<float> hcap(nseg, 0.f), soft2(nseg);
  std::vector<double> soft2d(nseg);
  std::vector<uint32_t> segof(n), root(nseg), start32(nseg + 1);
  const int64_t L = (n + K - 1) / K;
  for (int sg = 0; sg < nseg; ++sg) {
    float hmax = 0.f;
    for (int64_t i = start[sg]; i < start[sg + 1]; ++i) {
      if (p[i].mat >= (uint32_t)n_mat)
        return fail(h, IHS_ERR_ARG, "particle %lld has mat %u >= n_mat %d", (long long)i, p[i].mat, n_mat);
      if (!(p[i].h > 0.f) || !std::isfinite(p[i].h))
        return fail(h, IHS_ERR_ARG, "particle %lld has h <= 0", (long long)i);
      hmax = std::max(hmax, p[i].h);
      segof[i] = (uint32_t)sg;
    }
    hcap[sg] = h->cfg.h_cap_factor * hmax;
    soft2[sg] = softening[sg] * softening[sg];
    soft2d[sg] = (double)softening[sg] * (double)softening[sg];
    const int64_t leaf0 = start[sg] / K, nleaf = (start[sg + 1] - start[sg] + K - 1) / K;
    root[sg] = nleaf >= 2 ? (uint32_t)leaf0 : (uint32_t)(L - 1 + leaf0);  // a one-leaf tree is its leaf
    start32[sg] = (uint32_t)start[sg];
  }
  start32[nseg] = (uint32_t)n;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->n = n;
  h->nseg = nseg;
  h->zero_soft = false;
  for (int sg = 0; sg < nseg; ++sg) h->zero_soft |= !(soft2[sg] > 0.f);
  h->seg_start_host = start;
  h->n_mat = n_mat;
  h->cur = 0;
  h->have_fe = false;
  h->h_settled = false;
  h->pos_fresh = false;
  h->tree_fresh = false;
  Scratch& sc = h->sc;
  HIP_TRY(h, hipMemcpy(sc.eos, eos, n_mat * sizeof(ihs_eos), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.segof, segof.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.seg_start, start32.data(), (nseg + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.seg_soft2, soft2.data(), nseg * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.seg_soft2d, soft2d.data(), nseg * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.seg_hcap, hcap.data(), nseg * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(sc.seg_root, root.data(), nseg * sizeof(uint32_t), hipMemcpyHostToDevice));
  {
    std::vector<SegClock> clk(nseg);
    for (auto& c : clk) c = SegClock{0.0, 0.0, 0ull, 0xffffffffu, 0u};
    HIP_TRY(h, hipMemcpy(sc.seg_clk, clk.data(), nseg * sizeof(SegClock), hipMemcpyHostToDevice));
  }
  HIP_TRY(h, hipMemcpy(sc.staging, p, n * sizeof(ihs_particle), hipMemcpyHostToDevice));
  launch_unpack(h->stream, h->split, n, (const SphParticle48*)sc.staging, h->st[0]);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipGetLastError());
  if (h->cfg.h_mode != IHS_H_FIXED) {
    // The uploaded h is a guess: converge it before stepping. Each pass can grow h by the search
    // margin, so repeat while the number of solves stopped at the searched radius keeps falling.
    uint32_t prev = 0xffffffffu;
    for (int pass = 0; pass < 16; ++pass) {
      uint32_t before = 0, after = 0;
      HIP_TRY(h, hipMemcpy(&before, sc.flags + FLAG_H_CAPPED, 4, hipMemcpyDeviceToHost));
      h->h_settled = false;  // every pass may grow h by the full margin
      force_eval(h, false);
      HIP_TRY(h, hipStreamSynchronize(h->stream));
      HIP_TRY(h, hipMemcpy(&after, sc.flags + FLAG_H_CAPPED, 4, hipMemcpyDeviceToHost));
      const uint32_t capped = after - before;
      if (capped == 0 || capped >= prev) break;
      prev = capped;
    }
    HIP_TRY(h, hipGetLastError());
  }
  // Counters describe stepping only.
  HIP_TRY(h, hipMemset(sc.counters, 0, CNT_SLOTS * CNT_STRIDE * sizeof(unsigned long long)));
  {
    uint32_t fl[FLAG_COUNT] = {0};
    HIP_TRY(h, hipMemcpy(sc.flags, fl, sizeof fl, hipMemcpyHostToDevice));
  }
  h->force_evals = 0;
  h->list_builds = 0;
  return IHS_OK;
}

int ihs_upload(ihs_handle* h, const ihs_particle* p, int64_t n, const ihs_eos* eos, int32_t n_mat, float softening) {
  if (!h) return IHS_ERR_ARG;
  if (n < 1 || !p) return fail(h, IHS_ERR_ARG, "upload needs at least one particle");
  return ihs_upload_batch(h, p, &n, 1, eos, n_mat, &softening);
}

int ihs_segment_clock(ihs_handle* h, int32_t seg, double* t, double* dt, int64_t* steps) {
  if (!h) return IHS_ERR_ARG;
  if (seg < 0 || seg >= h->nseg) return fail(h, IHS_ERR_ARG, "scenario %d out of range (%d uploaded)", seg, h->nseg);
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  SegClock c;
  HIP_TRY(h, hipMemcpy(&c, h->sc.seg_clk + seg, sizeof c, hipMemcpyDeviceToHost));
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (steps) *steps = (int64_t)c.steps;
  return IHS_OK;
}

static int update_ctl(ihs_handle* h, void (*fn)(StepCtl&, const void*), const void* arg) {
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  StepCtl ctl;
  HIP_TRY(h, hipMemcpy(&ctl, h->sc.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  fn(ctl, arg);
  HIP_TRY(h, hipMemcpy(h->sc.ctl, &ctl, sizeof ctl, hipMemcpyHostToDevice));
  return IHS_OK;
}

int ihs_set_dt(ihs_handle* h, double dt, float damp) {
  if (!h) return IHS_ERR_ARG;
  struct A {
    double dt;
    float damp;
  } a{dt, damp};
  return update_ctl(h, [](StepCtl& c, const void* p) {
    const A* q = (const A*)p;
    c.dt_fixed = q->dt > 0.0 ? q->dt : 0.0;
    c.damp = q->damp;
  }, &a);
}

int ihs_set_av(ihs_handle* h, float alpha, float beta) {
  if (!h) return IHS_ERR_ARG;
  float a[2] = {alpha, beta};
  return update_ctl(h, [](StepCtl& c, const void* p) {
    c.av[0] = ((const float*)p)[0];
    c.av[1] = ((const float*)p)[1];
  }, a);
}

int ihs_set_omega(ihs_handle* h, float omega) {
  if (!h) return IHS_ERR_ARG;
  return update_ctl(h, [](StepCtl& c, const void* p) { c.omega = *(const float*)p; }, &omega);
}

int ihs_force_eval(ihs_handle* h) {
  if (!h) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  hipSetDevice(h->device);
  force_eval(h, false);
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

int ihs_step_kdk(ihs_handle* h, int32_t substeps) {
  if (!h) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  if (substeps < 0) return fail(h, IHS_ERR_ARG, "negative substeps");
  hipSetDevice(h->device);
  const bool graph = h->cfg.use_graph && !h->mlp.record && !h->prof_on;
  const bool fresh_needed = h->cfg.reuse_positions && !h->pos_fresh;
  if (graph && substeps > 0 &&
      ((h->cfg.h_mode != IHS_H_FIXED && !h->h_settled) || fresh_needed || !h->tree_fresh)) {
    // Captured steps assume h was solved (and, reusing positions, the forces evaluated) at the current
    // positions, and the tree sorted for them, as after any step; a step after an upload without a solve, after
    // relaxation or after another tree was built in the handle runs directly.
    kdk_step(h);
    --substeps;
  }
  if (!graph) {
    for (int i = 0; i < substeps; ++i) kdk_step(h);
    HIP_TRY(h, hipGetLastError());
    return IHS_OK;
  }
  // Each full evaluation flips the double-buffered state, so a step flips it twice, or once when its first
  // evaluation reuses positions: one graph per starting buffer, each leaving the state where its capture did.
  for (int i = 0; i < substeps; ++i) {
    const int par = h->cur;
    if (!h->graph_exec[par]) {
      const int64_t fe0 = h->force_evals, lb0 = h->list_builds;
      HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      kdk_step(h);
      hipError_t e = hipStreamEndCapture(h->stream, &h->graph[par]);
      h->force_evals = fe0;
      h->graph_builds[par] = (int)(h->list_builds - lb0);
      h->list_builds = lb0;
      h->graph_next[par] = h->cur;
      h->cur = par;
      if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
      HIP_TRY(h, hipGraphInstantiate(&h->graph_exec[par], h->graph[par], nullptr, nullptr, 0));
    }
    HIP_TRY(h, hipGraphLaunch(h->graph_exec[par], h->stream));
    h->cur = h->graph_next[par];
    h->force_evals += 2;
    h->list_builds += h->graph_builds[par];
    h->have_fe = true;
    h->pot_fresh = false;
  }
  return IHS_OK;
}

int ihs_relax(ihs_handle* h, int32_t steps) {
  if (!h) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  hipSetDevice(h->device);
  for (int i = 0; i < steps; ++i) {
    force_eval(h, false);
    launch_relax(h->stream, h->split, h->n, h->st[h->cur], h->sc.asph, h->sc.accg, h->sc.ctl, seg_view(h));
    h->h_settled = false;
    h->pos_fresh = false;
    h->tree_fresh = false;
  }
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

int64_t ihs_readback(ihs_handle* h, ihs_particle* out, int64_t cap, int32_t blocking) {
  if (!h || !out) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  if (!blocking) {
    const hipError_t q = hipStreamQuery(h->stream);
    if (q == hipErrorNotReady) return IHS_EAGAIN;
  }
  const int64_t n = std::min(cap, h->n);
  launch_pack(h->stream, h->split, h->n, h->st[h->cur], (SphParticle48*)h->sc.staging);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(out, h->sc.staging, n * sizeof(ihs_particle), hipMemcpyDeviceToHost));
  return n;
}

int ihs_begin_readback(ihs_handle* h) {
  if (!h) return IHS_ERR_ARG;
  if (h->n < 1 || h->snap_pending) return IHS_OK;  // nothing to copy, or a snapshot in flight (gpu_store.rs:127-129)
  hipSetDevice(h->device);
  if (!h->snap_dev) {
    HIP_TRY(h, dalloc(h, &h->snap_dev, h->cap));
    HIP_TRY(h, hipHostMalloc((void**)&h->snap_host, (size_t)h->cap * sizeof(ihs_particle), hipHostMallocDefault));
    HIP_TRY(h, hipEventCreateWithFlags(&h->snap_ev, hipEventDisableTiming));
  }
  // h->cur already names the buffer the state will be in once the queued steps have run.
  launch_pack(h->stream, h->split, h->n, h->st[h->cur], h->snap_dev);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(h->snap_host, h->snap_dev, (size_t)h->n * sizeof(ihs_particle), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipEventRecord(h->snap_ev, h->stream));
  h->snap_n = h->n;
  h->snap_pending = true;
  return IHS_OK;
}

int64_t ihs_take_readback(ihs_handle* h, ihs_particle* out, int64_t cap) {
  if (!h || !out || cap < 0) return IHS_ERR_ARG;
  if (!h->snap_pending) return IHS_EAGAIN;  // idle: nothing begun (gpu_store.rs:147-152 returns None)
  hipSetDevice(h->device);
  const hipError_t q = hipEventQuery(h->snap_ev);
  if (q == hipErrorNotReady) return IHS_EAGAIN;
  if (q != hipSuccess) return fail(h, IHS_ERR_HIP, "readback: %s", hipGetErrorString(q));
  const int64_t n = std::min(cap, h->snap_n);
  std::memcpy(out, h->snap_host, (size_t)n * sizeof(ihs_particle));
  h->snap_pending = false;
  return n;
}

int64_t ihs_count(const ihs_handle* h) { return h ? h->n : IHS_ERR_ARG; }

int ihs_sync(ihs_handle* h) {
  if (!h) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

// The potential of the last force evaluation's positions (in .w): accg when that evaluation summed it; otherwise one
// gravity walk into apot over that evaluation's tree and walk groups, which still describe the positions (a step ends
// with a kick, which moves only v and u). That walk does not open the nodes the neighbour search needs, so its
// potential is the plain walk's at the same theta; its interaction counts are not reported.
const float4* potential(ihs_handle* h) {
  if (h->pot_fresh) return h->sc.accg;
  hipMemsetAsync(h->tb.wq, 0, 8 * sizeof(uint32_t), h->stream);
  launch_gravity_walk(h->stream, h->n, h->cfg.leaf_size, h->split != 0, h->sc.xm, h->tb, h->tb.ggroups,
                      h->tb.nggroups, seg_view(h), h->sc.apot, h->sc.flags, h->sc.counters_aux, true, h->zero_soft);
  return h->sc.apot;
}

int ihs_stats_get(ihs_handle* h, ihs_stats* out, int32_t reset) {
  if (!h || !out) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  std::memset(out, 0, sizeof *out);
  if (h->have_fe) launch_energy(h->stream, h->split, h->n, h->st[h->cur], potential(h), h->sc.epart, h->sc.eout);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  double e[9] = {0};
  if (h->have_fe) HIP_TRY(h, hipMemcpy(e, h->sc.eout, sizeof e, hipMemcpyDeviceToHost));
  StepCtl ctl;
  uint32_t fl[FLAG_COUNT];
  unsigned long long cn[CNT_COUNT] = {0};
  std::vector<unsigned long long> slots(CNT_SLOTS * CNT_STRIDE);
  HIP_TRY(h, hipMemcpy(&ctl, h->sc.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  HIP_TRY(h, hipMemcpy(fl, h->sc.flags, sizeof fl, hipMemcpyDeviceToHost));
  HIP_TRY(h, hipMemcpy(slots.data(), h->sc.counters, slots.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int sl = 0; sl < CNT_SLOTS; ++sl)
    for (int c = 0; c < CNT_COUNT; ++c) cn[c] += slots[sl * CNT_STRIDE + c];
  out->ke = e[0];
  out->ie = e[1];
  out->pe = e[2];
  for (int q = 0; q < 3; ++q) {
    out->p[q] = e[3 + q];
    out->L[q] = e[6 + q];
  }
  SegClock clk0;
  HIP_TRY(h, hipMemcpy(&clk0, h->sc.seg_clk, sizeof clk0, hipMemcpyDeviceToHost));
  out->t = clk0.t;
  out->dt = clk0.dt;
  out->steps = (int64_t)clk0.steps;
  out->force_evals = h->force_evals;
  const double solved = (double)cn[CNT_H_SOLVED];
  out->mean_h_iters = solved > 0 ? (double)cn[CNT_H_ITERS] / solved : 0.0;
  out->mean_h_evals = solved > 0 ? (double)cn[CNT_H_EVALS] / solved : 0.0;
  out->mean_neighbours =
      h->list_builds > 0 && h->n > 0 ? (double)cn[CNT_PAIRS] / ((double)h->list_builds * (double)h->n) : 0.0;
  out->nbr_overflow = fl[FLAG_NBR_OVERFLOW];
  out->walk_overflow = fl[FLAG_WALK_OVERFLOW];
  out->h_capped = fl[FLAG_H_CAPPED];
  out->h_unconverged = fl[FLAG_H_UNCONVERGED];
  out->nonfinite = fl[FLAG_NONFINITE];
  out->tie_runs = fl[FLAG_TIE_RUNS];
  out->grav_cell_interactions = (int64_t)cn[CNT_GRAV_CELL];
  out->grav_pp_interactions = (int64_t)cn[CNT_GRAV_PP];
  out->nbr_candidates = (int64_t)cn[CNT_NBR_CAND];
  out->walk_nodes = (int64_t)cn[CNT_WALK_NODES];
  if (reset) {
    HIP_TRY(h, hipMemset(h->sc.counters, 0, slots.size() * sizeof(unsigned long long)));
    fl[FLAG_NBR_OVERFLOW] = fl[FLAG_WALK_OVERFLOW] = fl[FLAG_H_CAPPED] = fl[FLAG_H_UNCONVERGED] = fl[FLAG_NONFINITE] = 0;
    HIP_TRY(h, hipMemcpy(h->sc.flags, fl, sizeof fl, hipMemcpyHostToDevice));
    h->force_evals = 0;
    h->list_builds = 0;
  }
  return IHS_OK;
}

int ihs_segment_energy(ihs_handle* h, int32_t seg, double* out9) {
  if (!h || !out9) return IHS_ERR_ARG;
  if (seg < 0 || seg >= h->nseg) return fail(h, IHS_ERR_ARG, "scenario %d out of range (%d uploaded)", seg, h->nseg);
  if (!h->have_fe) return fail(h, IHS_ERR_STATE, "no force evaluation of the current state (the potential is unset)");
  hipSetDevice(h->device);
  const int64_t t0 = h->seg_start_host[seg], t1 = h->seg_start_host[seg + 1];
  // The scenario's storage range [t0, t1): a view of the master state from t0.
  const MasterBufs& S = h->st[h->cur];
  const size_t pw = h->split ? 8 : 4;
  MasterBufs V = S;
  V.px = (char*)S.px + pw * t0;
  V.py = (char*)S.py + pw * t0;
  V.pz = (char*)S.pz + pw * t0;
  V.u = (char*)S.u + pw * t0;
  V.vx = S.vx + t0;
  V.vy = S.vy + t0;
  V.vz = S.vz + t0;
  V.m = S.m + t0;
  launch_energy(h->stream, h->split, t1 - t0, V, potential(h) + t0, h->sc.epart, h->sc.eout);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(out9, h->sc.eout, 9 * sizeof(double), hipMemcpyDeviceToHost));
  return IHS_OK;
}

int ihs_profile(ihs_handle* h, int32_t enable) {
  if (!h) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->prof_on = enable != 0;
  h->prof_used = 0;
  return IHS_OK;
}

int ihs_profile_read(ihs_handle* h, double* ms, int64_t* calls, int32_t reset) {
  if (!h) return IHS_ERR_ARG;
  prof_flush(h);
  for (int i = 0; i < IHS_STAGE_COUNT; ++i) {
    if (ms) ms[i] = h->prof_ms[i];
    if (calls) calls[i] = h->prof_calls[i];
    if (reset) {
      h->prof_ms[i] = 0.0;
      h->prof_calls[i] = 0;
    }
  }
  return IHS_OK;
}

int ihs_gravity(ihs_handle* h, const double* pos3, const double* mass, int64_t n, double softening, double theta,
                int32_t mode, double* acc3, double* pot) {
  if (!h) return IHS_ERR_ARG;
  if (!pos3 || !mass || !acc3 || n < 1) return fail(h, IHS_ERR_ARG, "gravity needs n >= 1 bodies and buffers");
  if (n > h->cap) return fail(h, IHS_ERR_CAPACITY, "gravity on %lld bodies exceeds capacity", (long long)n);
  if (mode != IHS_GRAVITY_DIRECT && mode != IHS_GRAVITY_TREE) return fail(h, IHS_ERR_ARG, "bad gravity mode");
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  // The bodies take over the working copies, the tree and the accelerations: nothing a force evaluation left for
  // the particles stands until the next one.
  h->have_fe = false;
  h->pot_fresh = false;
  h->pos_fresh = false;
  h->tree_fresh = false;
  Scratch& sc = h->sc;
  if (!sc.gpos) {
    HIP_TRY(h, dalloc(h, &sc.gpos, 3 * h->cap));
    HIP_TRY(h, dalloc(h, &sc.gpx, h->cap));
    HIP_TRY(h, dalloc(h, &sc.gpy, h->cap));
    HIP_TRY(h, dalloc(h, &sc.gpz, h->cap));
    HIP_TRY(h, dalloc(h, &sc.gm, h->cap));
    HIP_TRY(h, dalloc(h, &sc.gmass, h->cap));
    HIP_TRY(h, dalloc(h, &sc.gacc, 3 * h->cap));
    HIP_TRY(h, dalloc(h, &sc.gpot, h->cap));
  }
  hipStream_t s = h->stream;
  HIP_TRY(h, hipMemcpyAsync(sc.gpos, pos3, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(sc.gmass, mass, n * sizeof(double), hipMemcpyHostToDevice, s));
  launch_bodies_in(s, sc.gpos, sc.gmass, n, sc.gpx, sc.gpy, sc.gpz, sc.gm);
  const float soft2 = (float)(softening * softening);
  if (mode == IHS_GRAVITY_DIRECT) {
    launch_gather_xm(s, nullptr, n, sc.gpx, sc.gpy, sc.gpz, sc.gm, sc.xm);
    launch_gravity_direct(s, n, soft2, sc.xm, sc.accg);
    launch_acc_out(s, nullptr, n, sc.accg, sc.gacc, sc.gpot);
  } else {
    const int K = h->cfg.leaf_size;
    // One segment: these bodies, their softening, a walk from node 0.
    const uint32_t st2[2] = {0u, (uint32_t)n}, root0 = 0u;
    HIP_TRY(h, hipMemcpy(sc.g_start, st2, sizeof st2, hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(sc.g_soft2, &soft2, sizeof soft2, hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(sc.g_root, &root0, sizeof root0, hipMemcpyHostToDevice));
    SegView gv;
    gv.S = 1;
    gv.start = sc.g_start;
    gv.soft2 = sc.g_soft2;
    gv.hcap = sc.g_hcap;
    gv.clk = sc.g_clk;
    gv.bbox = h->tb.bbox;
    gv.root = sc.g_root;
    launch_fe_init(s, sc.flags, gv);
    launch_bbox_morton(s, 1, sc.gpx, sc.gpy, sc.gpz, n, K, h->tb, gv);
    launch_radix_sort(s, n, 30, h->tb);  // identity payload: ties already ordered by index
    launch_gather_xm(s, h->tb.perm, n, sc.gpx, sc.gpy, sc.gpz, sc.gm, sc.xm);
    launch_karras(s, n, K, h->tb, gv);
    launch_upsweep(s, n, K, (float)theta, sc.xm, nullptr, h->tb);
    if (h->tb.ggroups != h->tb.groups) launch_groups(s, n, K, h->tb, gv, nullptr, GROUP_DIAG_HS, true);
    else launch_groups(s, n, K, h->tb, gv);
    launch_gravity_walk(s, n, K, 0, sc.xm, h->tb, h->tb.ggroups, h->tb.nggroups, gv, sc.accg, sc.flags, sc.counters,
                        true, !(soft2 > 0.f));
    launch_acc_out(s, h->tb.perm, n, sc.accg, sc.gacc, sc.gpot);
    h->have_fe = false;  // the tree buffers now describe these bodies, not the SPH state
  }
  HIP_TRY(h, hipMemcpyAsync(acc3, sc.gacc, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (pot) HIP_TRY(h, hipMemcpyAsync(pot, sc.gpot, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

int64_t ihs_debug_tree(ihs_handle* h, uint32_t* sorted_codes, uint32_t* order, uint32_t* ccodes, uint32_t* left,
                       uint32_t* right, uint32_t* parent) {
  if (!h) return IHS_ERR_ARG;
  if (!h->have_fe) return fail(h, IHS_ERR_STATE, "no force evaluation has built a tree for the current state");
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->n, K = h->cfg.leaf_size, L = (n + K - 1) / K;
  if (sorted_codes) HIP_TRY(h, hipMemcpy(sorted_codes, h->tb.skeys, n * 4, hipMemcpyDeviceToHost));
  if (order) HIP_TRY(h, hipMemcpy(order, h->st[h->cur].id, n * 4, hipMemcpyDeviceToHost));
  // Leaf c's cluster code is the first sorted code of its bucket.
  if (ccodes)
    HIP_TRY(h, hipMemcpy2D(ccodes, 4, h->tb.skeys, (size_t)4 * h->cfg.leaf_size, 4, (size_t)L, hipMemcpyDeviceToHost));
  std::vector<int2> ch(L > 1 ? L - 1 : 1);
  if (L > 1) HIP_TRY(h, hipMemcpy(ch.data(), h->tb.child, (L - 1) * sizeof(int2), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < 2 * L - 1; ++i) {
    if (left) left[i] = i < L - 1 ? (uint32_t)ch[i].x : 0u;
    if (right) right[i] = i < L - 1 ? (uint32_t)ch[i].y : 0u;
  }
  if (parent) HIP_TRY(h, hipMemcpy(parent, h->tb.parent, (2 * L - 1) * 4, hipMemcpyDeviceToHost));
  return L;
}

int ihs_debug_field(ihs_handle* h, int32_t field, void* out) {
  if (!h || !out) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  hipSetDevice(h->device);
  const int64_t n = h->n;
  hipStream_t s = h->stream;
  MasterBufs& S = h->st[h->cur];
  void* tmp = h->sc.staging;  // >= 48 bytes per particle
  size_t bytes = 0;
  switch (field) {
    case IHS_FIELD_H: launch_scatter_by_id(s, n, S.id, S.h, tmp, 4); bytes = 4 * n; break;
    case IHS_FIELD_RHO: launch_scatter_by_id(s, n, S.id, S.rho, tmp, 4); bytes = 4 * n; break;
    case IHS_FIELD_RPC: launch_scatter_by_id(s, n, S.id, h->sc.rpc, tmp, 16); bytes = 16 * n; break;
    case IHS_FIELD_ASPH: launch_scatter_by_id(s, n, S.id, h->sc.asph, tmp, 16); bytes = 16 * n; break;
    case IHS_FIELD_AGRAV: launch_scatter_by_id(s, n, S.id, h->sc.accg, tmp, 16); bytes = 16 * n; break;
    case IHS_FIELD_FEAT: launch_scatter_by_id(s, n, S.id, S.feat, tmp, 16); bytes = 16 * n; break;
    case IHS_FIELD_NCOUNT: launch_scatter_by_id(s, n, S.id, h->sc.ncount, tmp, 4); bytes = 4 * n; break;
    case IHS_FIELD_HS: launch_scatter_by_id(s, n, S.id, h->sc.hs, tmp, 4); bytes = 4 * n; break;
    case IHS_FIELD_HITER: {
      std::vector<uint8_t> it(n);
      std::vector<uint32_t> id(n);
      HIP_TRY(h, hipStreamSynchronize(s));
      HIP_TRY(h, hipMemcpy(it.data(), S.hiter, n, hipMemcpyDeviceToHost));
      HIP_TRY(h, hipMemcpy(id.data(), S.id, 4 * n, hipMemcpyDeviceToHost));
      for (int64_t t = 0; t < n; ++t) ((uint32_t*)out)[id[t]] = it[t];
      return IHS_OK;
    }
    case IHS_FIELD_POS64: launch_pos_f64(s, h->split, n, S, (double*)tmp); bytes = 24 * n; break;
    default: return fail(h, IHS_ERR_ARG, "unknown field %d", field);
  }
  HIP_TRY(h, hipStreamSynchronize(s));
  HIP_TRY(h, hipMemcpy(out, tmp, bytes, hipMemcpyDeviceToHost));
  return IHS_OK;
}

int ihs_direct_f64(ihs_handle* h, const uint32_t* ids, int64_t nt, double* acc3) {
  if (!h || !ids || !acc3 || nt < 1) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  for (int64_t k = 0; k < nt; ++k)
    if (ids[k] >= (uint64_t)h->n) return fail(h, IHS_ERR_ARG, "target id %u >= n", ids[k]);
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  uint32_t *d_ids = nullptr, *d_tpos = nullptr;
  double* d_acc = nullptr;
  hipError_t e = hipMalloc(&d_ids, nt * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&d_tpos, nt * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&d_acc, nt * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(d_ids, ids, nt * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const MasterBufs& S = h->st[h->cur];
    launch_direct_f64_targets(h->stream, h->split, S.px, S.py, S.pz, S.m, S.id, h->n, d_ids, nt, h->sc.seg_soft2d,
                              h->nseg > 1 ? h->sc.segof : nullptr, h->sc.seg_start, h->hw.pend, d_tpos, d_acc);
    e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(acc3, d_acc, nt * 3 * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (d_ids) hipFree(d_ids);
  if (d_tpos) hipFree(d_tpos);
  if (d_acc) hipFree(d_acc);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "direct sum: %s", hipGetErrorString(e));
  return IHS_OK;
}

int ihs_direct_f64_noalloc(ihs_handle* h, const uint32_t* ids, int64_t nt, double* acc3) {
  if (!h || !ids || !acc3 || nt < 1) return IHS_ERR_ARG;
  if (nt > kProbeMaxTargets)
    return fail(h, IHS_ERR_CAPACITY, "the allocation-free probe takes at most %lld targets per call, got %lld",
                (long long)kProbeMaxTargets, (long long)nt);
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "no particles uploaded");
  for (int64_t k = 0; k < nt; ++k)
    if (ids[k] >= (uint64_t)h->n) return fail(h, IHS_ERR_ARG, "target id %u >= n", ids[k]);
  if (!h->probe_scratch) return fail(h, IHS_ERR_STATE, "probe scratch unavailable on device %d", h->device);
  hipSetDevice(h->device);
  // Every word the kernels read is written first in this call (ids copied, positions mapped, accelerations
  // stored), so nothing is cleared; the storage-index table is the handle's own workspace (hw.pend).
  char* base = (char*)h->probe_scratch;
  uint32_t* d_ids = (uint32_t*)(base + kProbeIdsOff);
  uint32_t* d_tpos = (uint32_t*)(base + kProbeTposOff);
  double* d_acc = (double*)(base + kProbeAccOff);
  HIP_TRY(h, hipMemcpyAsync(d_ids, ids, nt * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream));
  const MasterBufs& S = h->st[h->cur];
  launch_direct_f64_targets(h->stream, h->split, S.px, S.py, S.pz, S.m, S.id, h->n, d_ids, nt, h->sc.seg_soft2d,
                            h->nseg > 1 ? h->sc.segof : nullptr, h->sc.seg_start, h->hw.pend, d_tpos, d_acc);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(acc3, d_acc, nt * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return IHS_OK;
}

// ---- learned smoothing-length seed ----

int ihs_mlp_set_weights(ihs_handle* h, const float* w, int64_t nparam) {
  if (!h || !w) return IHS_ERR_ARG;
  if (nparam != MLP_NPARAM) return fail(h, IHS_ERR_ARG, "expected %d parameters, got %lld", MLP_NPARAM, (long long)nparam);
  for (int64_t i = 0; i < nparam; ++i)
    if (!std::isfinite(w[i])) return fail(h, IHS_ERR_ARG, "parameter %lld is not finite", (long long)i);
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  drop_graph(h);  // captured steps bake in whether the seed runs
  MlpDev& m = h->mlp;
  HIP_TRY(h, hipMemcpy(m.w, w, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemset(m.adam_m, 0, MLP_NPARAM * sizeof(float)));
  HIP_TRY(h, hipMemset(m.adam_v, 0, MLP_NPARAM * sizeof(float)));
  m.adam_t = 0;
  mlp_sync_bf16(h->stream, m);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipGetLastError());
  m.trained = true;
  return IHS_OK;
}

int ihs_mlp_get_weights(ihs_handle* h, float* w, int64_t nparam) {
  if (!h || !w) return IHS_ERR_ARG;
  if (nparam < MLP_NPARAM) return fail(h, IHS_ERR_ARG, "need room for %d parameters", MLP_NPARAM);
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(w, h->mlp.w, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return IHS_OK;
}

int ihs_mlp_record(ihs_handle* h, int64_t max_samples) {
  if (!h || max_samples < 0) return IHS_ERR_ARG;
  if (max_samples > 0 && h->cfg.h_mode != IHS_H_NEWTON)
    return fail(h, IHS_ERR_STATE, "recording needs the plain Newton h-solve (h_mode NEWTON)");
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  MlpDev& m = h->mlp;
  if (max_samples == 0) {
    m.record = false;
    return IHS_OK;
  }
  if (max_samples > m.sample_cap) {
    if (m.samples) HIP_TRY(h, hipFree(m.samples));
    m.samples = nullptr;
    m.sample_cap = 0;
    HIP_TRY(h, hipMalloc(&m.samples, (size_t)max_samples * MLP_SAMPLE * sizeof(float)));
    m.sample_cap = max_samples;
  }
  m.nsamples = 0;
  m.record = true;
  return IHS_OK;
}

int64_t ihs_mlp_samples(ihs_handle* h, float* out, int64_t cap) {
  if (!h || cap < 0) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int64_t k = std::min(cap, h->mlp.nsamples);
  if (out && k > 0)
    HIP_TRY(h, hipMemcpy(out, h->mlp.samples, (size_t)k * MLP_SAMPLE * sizeof(float), hipMemcpyDeviceToHost));
  return out ? k : h->mlp.nsamples;
}

int ihs_mlp_train(ihs_handle* h, const float* rows_host, int64_t n_rows, int32_t steps, int32_t batch, float lr,
                  uint64_t seed, float* loss) {
  if (!h || steps < 0 || batch < 1 || !(lr > 0.f)) return IHS_ERR_ARG;
  MlpDev& m = h->mlp;
  if (!m.trained) return fail(h, IHS_ERR_STATE, "set initial weights before training");
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int64_t rows = rows_host ? n_rows : m.nsamples;
  if (rows < 1) return fail(h, IHS_ERR_STATE, "no training rows");
  float* tmp = nullptr;
  const float* table = m.samples;
  if (rows_host) {
    tmp = upload_tmp(h, rows_host, (size_t)rows * MLP_SAMPLE);
    if (!tmp) return IHS_ERR_HIP;
    table = tmp;
  }
  float* dloss = nullptr;
  hipError_t e = hipMalloc(&dloss, std::max<int>(steps, 1) * sizeof(float));
  if (e != hipSuccess) {
    if (tmp) hipFree(tmp);
    return fail(h, IHS_ERR_HIP, "loss buffer: %s", hipGetErrorString(e));
  }
  drop_graph(h);
  uint64_t rng = seed;
  const int64_t nb = std::min<int64_t>(batch, rows);
  for (int32_t it = 0; it < steps; ++it) {
    const uint64_t off = splitmix64(rng) % (uint64_t)rows;
    uint64_t stride = rows > 1 ? 1 + splitmix64(rng) % (uint64_t)(rows - 1) : 1;
    while (gcd64(stride, (uint64_t)rows) != 1) stride = stride + 1 < (uint64_t)rows ? stride + 1 : 1;
    mlp_gradient(h->stream, m, table, MLP_SAMPLE, table + MLP_IN, MLP_SAMPLE, nb, rows, (int64_t)off,
                 (int64_t)stride);
    const float frac = steps > 1 ? (float)it / (float)(steps - 1) : 0.f;
    const float lr_it = lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(kPi * frac)));
    mlp_adam(h->stream, m, lr_it);
    hipMemcpyAsync(dloss + it, m.grad + MLP_NPARAM, sizeof(float), hipMemcpyDeviceToDevice, h->stream);
  }
  e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && loss && steps > 0) e = hipMemcpy(loss, dloss, steps * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(dloss);
  if (tmp) hipFree(tmp);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "training: %s", hipGetErrorString(e));
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

int ihs_mlp_gradient(ihs_handle* h, const float* x, const float* y, int64_t n, float* grad) {
  if (!h || !x || !y || !grad || n < 1) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  float* dx = upload_tmp(h, x, (size_t)n * MLP_IN);
  if (!dx) return IHS_ERR_HIP;
  float* dy = upload_tmp(h, y, (size_t)n);
  if (!dy) {
    hipFree(dx);
    return IHS_ERR_HIP;
  }
  mlp_gradient(h->stream, h->mlp, dx, MLP_IN, dy, 1, n, n, 0, 1);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(grad, h->mlp.grad, (MLP_NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(dx);
  hipFree(dy);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "gradient: %s", hipGetErrorString(e));
  return IHS_OK;
}

int ihs_mlp_predict(ihs_handle* h, const float* x, int64_t n, float* y) {
  if (!h || !x || !y || n < 1) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  float* dx = upload_tmp(h, x, (size_t)n * MLP_IN);
  if (!dx) return IHS_ERR_HIP;
  float* dy = nullptr;
  hipError_t e = hipMalloc(&dy, (size_t)n * sizeof(float));
  if (e == hipSuccess) {
    mlp_forward(h->stream, h->mlp, dx, n, dy);
    e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dy);
  }
  hipFree(dx);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "predict: %s", hipGetErrorString(e));
  return IHS_OK;
}

// ---- Disk / remnant measurement (k_measure.hip) ----

// Packs the current state into the staging records in upload order (what ihs_readback copies out) and points at
// one scenario's particles.
static int measure_snapshot(ihs_handle* h, int32_t segment, const SphParticle48** P, int64_t* n) {
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "measurement needs uploaded particles");
  if (segment < 0 || segment >= h->nseg)
    return fail(h, IHS_ERR_ARG, "segment %d outside [0, %d)", segment, h->nseg);
  HIP_TRY(h, measure_reserve(h->mb, h->cap));
  launch_pack(h->stream, h->split, h->n, h->st[h->cur], (SphParticle48*)h->sc.staging);
  *P = (const SphParticle48*)h->sc.staging + h->seg_start_host[segment];
  *n = h->seg_start_host[segment + 1] - h->seg_start_host[segment];
  return IHS_OK;
}

int ihs_measure_impact(ihs_handle* h, int32_t segment, int64_t n_earth, ihs_measure* out) {
  if (!h || !out) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  const SphParticle48* P = nullptr;
  int64_t n = 0;
  if (int rc = measure_snapshot(h, segment, &P, &n)) return rc;
  if (n_earth < 0 || n_earth > n) return fail(h, IHS_ERR_ARG, "n_earth %lld outside [0, %lld]", (long long)n_earth,
                                              (long long)n);
  MeasureFrame fr;
  HIP_TRY(h, measure_frame(h->stream, h->mb, P, n, MM_IMPACT_RUN, n_earth, &fr));
  ihs_measure m;
  std::memset(&m, 0, sizeof m);
  m.e_disk = fr.cat[CAT_DISK_E];
  m.t_disk = fr.cat[CAT_DISK_T];
  m.disk_kg = m.e_disk + m.t_disk;
  m.earth_frac = m.disk_kg > 0.0 ? 100.0 * m.e_disk / m.disk_kg : 0.0;
  m.e_rem = fr.cat[CAT_REM_E];
  m.t_rem = fr.cat[CAT_REM_T];
  m.e_esc = fr.cat[CAT_ESC_E];
  m.t_esc = fr.cat[CAT_ESC_T];
  m.r_remnant = fr.r_remnant;
  m.m_total = fr.m_total;
  m.m_remnant = fr.m_remnant;
  for (int q = 0; q < 3; ++q) {
    m.com[q] = fr.com[q];
    m.v_com[q] = fr.v_com[q];
  }
  m.n_disk = fr.n_disk;
  if (fr.n_disk >= 2) {  // main.rs:803-805
    ClumpQuery q;
    q.impact_run = 1;
    q.link = 2.0 * (fr.cat[CAT_DISK_H] / (double)fr.n_disk);  // main.rs:807-808
    q.g = 6.674e-11;
    for (int a = 0; a < 3; ++a) q.central[a] = fr.com[a];
    q.central_mass = fr.m_remnant;
    q.central_radius = fr.r_remnant;
    ClumpSummary sum;
    ihs_clump best;
    HIP_TRY(h, clump_set(h->stream, h->mb, fr.n_disk, q, nullptr, nullptr, &sum, &best));
    m.link = q.link;
    m.n_clumps = sum.n2;
    m.n_bound = sum.n_bound;
    if (sum.best >= 0) {
      m.clump_kg = best.mass;
      m.clump_earth_frac = 100.0 * best.earth_mass / best.mass;
      m.clump_n = best.n_members;
    }
  }
  *out = m;
  return IHS_OK;
}

// The engine's disk and its find_clumps query (gpu_sph.rs:399-640): softening 1e4 m, the remnant as the
// central body, linking length twice the mean h of all particles.
static int engine_disk(ihs_handle* h, int32_t segment, int mode, MeasureFrame* fr, ClumpQuery* q, int64_t* n_out) {
  const SphParticle48* P = nullptr;
  int64_t n = 0;
  if (int rc = measure_snapshot(h, segment, &P, &n)) return rc;
  HIP_TRY(h, measure_frame(h->stream, h->mb, P, n, mode, 0, fr));
  q->impact_run = 0;
  q->link = 2.0 * (fr->sum_h / (double)n);
  q->g = 6.674e-11;
  q->soft2 = 1.0e4 * 1.0e4;
  for (int a = 0; a < 3; ++a) q->central[a] = fr->com[a];
  q->central_mass = fr->m_remnant;
  q->central_radius = fr->r_remnant;
  *n_out = n;
  return IHS_OK;
}

int ihs_disk_stats(ihs_handle* h, int32_t segment, ihs_disk_summary* out) {
  if (!h || !out) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  MeasureFrame fr;
  ClumpQuery q;
  int64_t n = 0;
  if (int rc = engine_disk(h, segment, MM_ENGINE_BOUND, &fr, &q, &n)) return rc;
  constexpr double kMoon = 7.342e22;  // gpu_sph.rs:401
  ihs_disk_summary d;
  std::memset(&d, 0, sizeof d);
  d.e_disk = fr.cat[CAT_DISK_E];
  d.t_disk = fr.cat[CAT_DISK_T];
  d.esc = fr.cat[CAT_ESC_ALL];
  d.r_remnant = fr.r_remnant;
  d.m_remnant = fr.m_remnant;
  d.n_disk = fr.n_disk;
  const double disk = d.e_disk + d.t_disk;
  if (fr.n_disk >= 2) {  // gpu_sph.rs:456-460
    ClumpSummary sum;
    ihs_clump best;
    HIP_TRY(h, clump_set(h->stream, h->mb, fr.n_disk, q, nullptr, nullptr, &sum, &best));
    d.link = q.link;
    d.n_clumps = sum.n2;
    if (sum.best >= 0) d.biggest = best.mass;
  }
  d.disk = disk / kMoon;
  d.earth_pct = disk > 0.0 ? 100.0 * d.e_disk / disk : 0.0;
  d.remnant_km = fr.r_remnant / 1e3;
  d.escaped = d.esc / kMoon;
  d.moon = d.biggest / kMoon;
  *out = d;
  return IHS_OK;
}

int ihs_disk_clumps(ihs_handle* h, int32_t segment, int32_t require_bound, double* frame, ihs_clump* clumps,
                    int64_t cap, int64_t* n_clumps) {
  if (!h || !frame || !n_clumps || (cap > 0 && !clumps)) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  MeasureFrame fr;
  ClumpQuery q;
  int64_t n = 0;
  if (int rc = engine_disk(h, segment, require_bound ? MM_ENGINE_BOUND : MM_ENGINE_ANY, &fr, &q, &n)) return rc;
  frame[0] = fr.m_total;
  for (int a = 0; a < 3; ++a) {
    frame[1 + a] = fr.com[a];
    frame[4 + a] = fr.v_com[a];
  }
  frame[7] = fr.r_remnant;
  frame[8] = fr.m_remnant;
  frame[9] = 6.674e-11 * fr.m_remnant;
  frame[10] = fr.cat[CAT_DISK_E] + fr.cat[CAT_DISK_T];
  frame[11] = q.link;
  frame[12] = (double)fr.n_disk;
  *n_clumps = 0;
  if (n < 2 || fr.n_disk < 2) return IHS_OK;  // gpu_sph.rs:481-483, 500-502
  std::vector<ihs_clump> all;
  HIP_TRY(h, clump_set(h->stream, h->mb, fr.n_disk, q, &all, nullptr, nullptr, nullptr));
  int64_t kept = 0;
  for (const ihs_clump& c : all) {
    if (!(c.mass > 0.0)) continue;  // accretion.rs:127-129
    if (kept < cap) clumps[kept] = c;
    ++kept;
  }
  *n_clumps = kept;
  return IHS_OK;
}

int ihs_find_clumps(ihs_handle* h, int64_t n, const double* pos3, const double* vel3, const double* mass,
                    const double* rho, double linking_length, double g, double softening, const double* central3,
                    double central_mass, double central_radius, ihs_clump* clumps, int64_t cap, int64_t* n_clumps,
                    int64_t* label) {
  if (!h || !n_clumps || n < 0 || (n > 0 && (!pos3 || !vel3 || !mass || !rho)) || !central3 || (cap > 0 && !clumps))
    return IHS_ERR_ARG;
  if (!(linking_length >= 0.0) || !std::isfinite(linking_length))
    return fail(h, IHS_ERR_ARG, "linking length must be finite and >= 0");
  if (n > (int64_t)0x7fffffff) return fail(h, IHS_ERR_CAPACITY, "find_clumps takes at most 2^31 - 1 bodies");
  *n_clumps = 0;
  if (n == 0) return IHS_OK;
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, measure_reserve(h->mb, std::max<int64_t>(n, h->mb.cap)));
  HIP_TRY(h, clump_upload(h->stream, h->mb, n, pos3, vel3, mass, rho));
  ClumpQuery q;
  q.link = linking_length;
  q.g = g;
  q.soft2 = softening * softening;
  for (int a = 0; a < 3; ++a) q.central[a] = central3[a];
  q.central_mass = central_mass;
  q.central_radius = central_radius;
  std::vector<ihs_clump> all;
  std::vector<int32_t> lab;
  HIP_TRY(h, clump_set(h->stream, h->mb, n, q, &all, label ? &lab : nullptr, nullptr, nullptr));
  std::vector<int64_t> remap(all.size(), -1);
  int64_t kept = 0;
  for (size_t c = 0; c < all.size(); ++c) {
    if (!(all[c].mass > 0.0)) continue;  // accretion.rs:127-129
    if (kept < cap) clumps[kept] = all[c];
    remap[c] = kept++;
  }
  *n_clumps = kept;
  if (label)
    for (int64_t i = 0; i < n; ++i) label[i] = remap[(size_t)lab[i]];
  return IHS_OK;
}

// ---- IC assembly on the device (gpu_sph.rs:279-339) ----

int ihs_body_bulk(ihs_handle* h, int32_t segment, uint32_t prov, double* out5) {
  if (!h || !out5) return IHS_ERR_ARG;
  hipSetDevice(h->device);
  const SphParticle48* P = nullptr;
  int64_t n = 0;
  if (int rc = measure_snapshot(h, segment, &P, &n)) return rc;
  HIP_TRY(h, body_bulk(h->stream, h->mb, P, n, prov, out5));
  return IHS_OK;
}

int ihs_assemble(ihs_handle* h, int32_t segment, const double* placements, int32_t nbody, float* softening,
                 float* dt) {
  if (!h || !placements || !softening || !dt) return IHS_ERR_ARG;
  if (nbody < 1 || nbody > IHS_MAX_BODIES) return fail(h, IHS_ERR_ARG, "need 1..%d placements, got %d", IHS_MAX_BODIES,
                                                          nbody);
  for (int k = 0; k < 9 * nbody; ++k)
    if (!std::isfinite(placements[k])) return fail(h, IHS_ERR_ARG, "placement %d is not finite", k / 9);
  hipSetDevice(h->device);
  const SphParticle48* P = nullptr;
  int64_t n = 0;
  if (int rc = measure_snapshot(h, segment, &P, &n)) return rc;
  const int64_t t0 = h->seg_start_host[segment], t1 = h->seg_start_host[segment + 1];
  MasterBufs& S = h->st[h->cur];
  uint32_t* dflags = nullptr;
  double* dpl = nullptr;
  HIP_TRY(h, hipMalloc(&dflags, 2 * sizeof(uint32_t)));
  hipError_t e = hipMalloc(&dpl, (size_t)12 * nbody * sizeof(double));
  auto done = [&](hipError_t err) {
    hipFree(dflags);
    if (dpl) hipFree(dpl);
    return err;
  };
  // Every particle must have a placement: the reference drops the others, which would change the count.
  uint32_t init[2] = {0u, 0xffffffffu}, got[2] = {0u, 0u};
  if (e == hipSuccess) e = hipMemcpyAsync(dflags, init, sizeof init, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) launch_prov_max(h->stream, t0, t1, S, dflags);
  if (e == hipSuccess) e = hipMemcpyAsync(got, dflags, sizeof got, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "assemble: %s", hipGetErrorString(done(e)));
  if (got[0] >= (uint32_t)nbody) {
    done(hipSuccess);
    return fail(h, IHS_ERR_ARG, "particles of provenance %u have no placement (%d given); the in-place assembly keeps "
                "every particle", got[0], nbody);
  }
  // Each body's robust centre (gpu_sph.rs:326).
  std::vector<double> pl((size_t)12 * nbody);
  double max_speed = 0.0;
  for (int k = 0; k < nbody; ++k) {
    double bulk[5];
    e = body_bulk(h->stream, h->mb, P, n, (uint32_t)k, bulk);
    if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "body_bulk: %s", hipGetErrorString(done(e)));
    const double* q = placements + 9 * k;
    for (int a = 0; a < 3; ++a) {
      pl[12 * k + a] = bulk[a];
      pl[12 * k + 3 + a] = q[a];
      pl[12 * k + 6 + a] = q[3 + a];
      pl[12 * k + 9 + a] = q[6 + a];
    }
    max_speed = std::max(max_speed, std::sqrt(q[3] * q[3] + q[4] * q[4] + q[5] * q[5]));  // gpu_sph.rs:337
  }
  e = hipMemcpyAsync(dpl, pl.data(), pl.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) launch_assemble(h->stream, h->split, t0, t1, S, dpl, dflags + 1);
  if (e == hipSuccess) e = hipMemcpyAsync(got, dflags, sizeof got, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  done(e);
  if (e != hipSuccess) return fail(h, IHS_ERR_HIP, "assemble: %s", hipGetErrorString(e));
  // softening = min h / 4; the shock-safe dt against the fastest approach (gpu_sph.rs:333-338).
  uint32_t k = got[1];
  const float min_h = (k >> 31) ? __builtin_bit_cast(float, k ^ 0x80000000u) : __builtin_bit_cast(float, ~k);
  *softening = 0.25f * min_h;
  *dt = (float)(0.01 * (double)min_h / (20000.0 + max_speed));
  // The scenario's softening, clock and cached evaluation state follow the new configuration.
  std::vector<float> soft2(h->nseg);
  HIP_TRY(h, hipMemcpy(soft2.data(), h->sc.seg_soft2, h->nseg * sizeof(float), hipMemcpyDeviceToHost));
  soft2[segment] = *softening * *softening;
  const double s2d = (double)*softening * (double)*softening;
  HIP_TRY(h, hipMemcpy(h->sc.seg_soft2 + segment, &soft2[segment], sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->sc.seg_soft2d + segment, &s2d, sizeof(double), hipMemcpyHostToDevice));
  h->zero_soft = false;
  for (float v : soft2) h->zero_soft |= !(v > 0.f);
  const SegClock clk{0.0, 0.0, 0ull, 0xffffffffu, 0u};
  HIP_TRY(h, hipMemcpy(h->sc.seg_clk + segment, &clk, sizeof clk, hipMemcpyHostToDevice));
  h->have_fe = false;
  h->h_settled = false;
  h->pos_fresh = false;
  h->tree_fresh = false;
  return IHS_OK;
}

// ---- Checkpoint / resume: a versioned binary image of the device state ----

namespace {
constexpr char kCkptMagic[8] = {'I', 'H', 'S', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;

struct CkptHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int64_t n;
  int64_t total_bytes;
  int32_t nseg;
  int32_t split;
  ihs_config cfg;
  StepCtl ctl;
  int64_t force_evals;
  uint8_t h_settled, zero_soft, mlp_trained;
  uint8_t n_mat;  // EOS records in use (0 in files that predate the field: then the table's 16 bound)
  uint8_t pad[4];
};

// Indices of the sections ckpt_sections lists, for the checks on the host before a load.
enum CkptSection {
  CK_SEG_START = 0, CK_SEGOF, CK_SOFT2, CK_SOFT2D, CK_HCAP, CK_CLK, CK_ROOT, CK_EOS, CK_MLP,
  CK_PX, CK_PY, CK_PZ, CK_U, CK_VX, CK_VY, CK_VZ, CK_H, CK_M, CK_MAT, CK_ID, CK_PROV, CK_RHO, CK_FEAT, CK_HITER,
  CK_COUNT
};

// The device sections of a checkpoint, in file order: the scenario tables, the EOS table, the network weights and
// the master state in its storage order.
std::vector<std::pair<void*, size_t>> ckpt_sections(ihs_handle* h, const MasterBufs& S, int64_t n, int nseg) {
  Scratch& sc = h->sc;
  const size_t pw = h->split ? 8 : 4;
  return {
      {sc.seg_start, sizeof(uint32_t) * (nseg + 1)}, {sc.segof, sizeof(uint32_t) * n},
      {sc.seg_soft2, sizeof(float) * nseg},          {sc.seg_soft2d, sizeof(double) * nseg},
      {sc.seg_hcap, sizeof(float) * nseg},           {sc.seg_clk, sizeof(SegClock) * nseg},
      {sc.seg_root, sizeof(uint32_t) * nseg},        {sc.eos, 16 * sizeof(ihs_eos)},
      {h->mlp.w, sizeof(float) * MLP_NPARAM},
      {S.px, pw * n}, {S.py, pw * n}, {S.pz, pw * n}, {S.u, pw * n},
      {S.vx, 4 * n}, {S.vy, 4 * n}, {S.vz, 4 * n}, {S.h, 4 * n}, {S.m, 4 * n}, {S.mat, (size_t)n},
      {S.id, 4 * n}, {S.prov, 4 * n}, {S.rho, 4 * n}, {S.feat, 16 * n}, {S.hiter, (size_t)n},
  };
}

int64_t ckpt_total(ihs_handle* h) {
  int64_t t = sizeof(CkptHeader);
  for (auto& sct : ckpt_sections(h, h->st[h->cur], h->n, h->nseg)) t += (int64_t)sct.second;
  return t;
}
}  // namespace

int64_t ihs_checkpoint_bytes(ihs_handle* h) {
  if (!h) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "nothing to checkpoint: no particles uploaded");
  return ckpt_total(h);
}

int64_t ihs_checkpoint_save(ihs_handle* h, void* buf, int64_t cap) {
  if (!h || !buf) return IHS_ERR_ARG;
  if (h->n < 1) return fail(h, IHS_ERR_STATE, "nothing to checkpoint: no particles uploaded");
  const int64_t total = ckpt_total(h);
  if (cap < total) return fail(h, IHS_ERR_CAPACITY, "checkpoint needs %lld bytes, buffer holds %lld", (long long)total,
                               (long long)cap);
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  CkptHeader hd;
  std::memset(&hd, 0, sizeof hd);
  std::memcpy(hd.magic, kCkptMagic, sizeof hd.magic);
  hd.version = kCkptVersion;
  hd.header_bytes = sizeof(CkptHeader);
  hd.n = h->n;
  hd.total_bytes = total;
  hd.nseg = h->nseg;
  hd.split = h->split;
  hd.cfg = h->cfg;
  HIP_TRY(h, hipMemcpy(&hd.ctl, h->sc.ctl, sizeof(StepCtl), hipMemcpyDeviceToHost));
  hd.force_evals = h->force_evals;
  hd.h_settled = h->h_settled;
  hd.zero_soft = h->zero_soft;
  hd.mlp_trained = h->mlp.trained;
  hd.n_mat = (uint8_t)h->n_mat;
  char* out = (char*)buf;
  std::memcpy(out, &hd, sizeof hd);
  int64_t off = sizeof hd;
  for (auto& sct : ckpt_sections(h, h->st[h->cur], h->n, h->nseg)) {
    HIP_TRY(h, hipMemcpy(out + off, sct.first, sct.second, hipMemcpyDeviceToHost));
    off += (int64_t)sct.second;
  }
  return off;
}

int ihs_checkpoint_load(ihs_handle* h, const void* buf, int64_t bytes) {
  if (!h || !buf) return IHS_ERR_ARG;
  CkptHeader hd;
  if (bytes < (int64_t)sizeof hd) return fail(h, IHS_ERR_ARG, "checkpoint truncated (%lld bytes)", (long long)bytes);
  std::memcpy(&hd, buf, sizeof hd);
  if (std::memcmp(hd.magic, kCkptMagic, sizeof hd.magic) != 0) return fail(h, IHS_ERR_ARG, "not a checkpoint");
  if (hd.version != kCkptVersion || hd.header_bytes != sizeof hd)
    return fail(h, IHS_ERR_ARG, "checkpoint version %u (header %u bytes) is not version %u", hd.version,
                hd.header_bytes, kCkptVersion);
  if (hd.n < 1 || hd.n > h->cap) return fail(h, IHS_ERR_CAPACITY, "checkpoint of %lld particles, capacity %lld",
                                             (long long)hd.n, (long long)h->cap);
  if (hd.nseg < 1 || hd.nseg > IHS_MAX_SEGMENTS) return fail(h, IHS_ERR_ARG, "checkpoint has %d scenarios", hd.nseg);
  ihs_config mine = h->cfg, theirs = hd.cfg;
  mine.capacity = theirs.capacity = 0;
  if (std::memcmp(&mine, &theirs, sizeof mine) != 0)
    return fail(h, IHS_ERR_ARG, "checkpoint was written with another configuration (create the handle from its "
                "header's ihs_config)");
  hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->n = hd.n;
  h->nseg = hd.nseg;
  h->cur = 0;
  auto secs = ckpt_sections(h, h->st[0], hd.n, hd.nseg);
  int64_t total = sizeof hd;
  for (auto& sct : secs) total += (int64_t)sct.second;
  if (total != hd.total_bytes || total != bytes) {
    h->n = 0;
    return fail(h, IHS_ERR_ARG, "checkpoint size %lld does not match its header (%lld)", (long long)bytes,
                (long long)total);
  }
  const char* in = (const char*)buf;
  if (secs.size() != CK_COUNT) {
    h->n = 0;
    return fail(h, IHS_ERR_STATE, "checkpoint section table out of step with its reader");
  }
  std::vector<const char*> at(CK_COUNT);
  {
    int64_t off = sizeof hd;
    for (int k = 0; k < CK_COUNT; ++k) {
      at[k] = in + off;
      off += (int64_t)secs[k].second;
    }
  }
  // Everything a kernel indexes with is checked here, before any of it reaches the device: the scenario table,
  // the particle ids (the scatter to upload order writes through them), the material indices (the EOS table
  // has 16 records) and h; the per-particle scenario index and the tree roots are rebuilt from the table.
  const int64_t n = hd.n;
  const int nseg = hd.nseg;
  const int K = h->cfg.leaf_size;
  std::vector<uint32_t> start32(nseg + 1);
  std::memcpy(start32.data(), at[CK_SEG_START], sizeof(uint32_t) * (nseg + 1));
  auto refuse = [&](const char* what) {
    h->n = 0;
    return fail(h, IHS_ERR_ARG, "corrupt checkpoint: %s", what);
  };
  if (start32[0] != 0u || (int64_t)start32[nseg] != n) return refuse("scenario table does not span the particles");
  for (int sg = 0; sg < nseg; ++sg) {
    if (start32[sg + 1] <= start32[sg]) return refuse("scenario table not increasing");
    if (nseg > 1 && (start32[sg + 1] - start32[sg]) % (uint32_t)K != 0) return refuse("scenario not leaf-aligned");
  }
  {
    std::vector<uint8_t> seen(n, 0);
    const uint32_t* id = (const uint32_t*)at[CK_ID];
    for (int64_t t = 0; t < n; ++t) {
      uint32_t v;
      std::memcpy(&v, id + t, sizeof v);
      if (v >= (uint64_t)n || seen[v]) return refuse("particle ids are not a permutation of 0..n-1");
      seen[v] = 1;
    }
  }
  const int mat_bound = hd.n_mat >= 1 && hd.n_mat <= 16 ? hd.n_mat : 16;
  for (int64_t t = 0; t < n; ++t)
    if ((uint8_t)at[CK_MAT][t] >= mat_bound) return refuse("material index outside the EOS table");
  for (int64_t t = 0; t < n; ++t) {
    float v;
    std::memcpy(&v, at[CK_H] + 4 * t, sizeof v);
    if (!(v > 0.f) || !std::isfinite(v)) return refuse("smoothing length not positive and finite");
  }
  for (int sg = 0; sg < nseg; ++sg) {
    float f;
    double d;
    std::memcpy(&f, at[CK_SOFT2] + 4 * sg, sizeof f);
    std::memcpy(&d, at[CK_SOFT2D] + 8 * sg, sizeof d);
    if (!(f >= 0.f) || !std::isfinite(f) || !(d >= 0.0) || !std::isfinite(d)) return refuse("softening");
  }
  if (hd.mlp_trained)
    for (int k = 0; k < MLP_NPARAM; ++k) {
      float w;
      std::memcpy(&w, at[CK_MLP] + 4 * k, sizeof w);
      if (!std::isfinite(w)) return refuse("network weight not finite");
    }
  std::vector<uint32_t> segof(n), root(nseg);
  const int64_t L = (n + K - 1) / K;
  for (int sg = 0; sg < nseg; ++sg) {
    for (int64_t t = start32[sg]; t < start32[sg + 1]; ++t) segof[t] = (uint32_t)sg;
    const int64_t leaf0 = start32[sg] / K, nleaf = (start32[sg + 1] - start32[sg] + K - 1) / K;
    root[sg] = nleaf >= 2 ? (uint32_t)leaf0 : (uint32_t)(L - 1 + leaf0);
  }
  for (int k = 0; k < CK_COUNT; ++k) {
    const void* src = k == CK_SEGOF ? (const void*)segof.data() : k == CK_ROOT ? (const void*)root.data() : at[k];
    HIP_TRY(h, hipMemcpy(secs[k].first, src, secs[k].second, hipMemcpyHostToDevice));
  }
  HIP_TRY(h, hipMemcpy(h->sc.ctl, &hd.ctl, sizeof(StepCtl), hipMemcpyHostToDevice));
  h->seg_start_host.assign(start32.begin(), start32.end());
  h->n_mat = mat_bound;
  h->zero_soft = hd.zero_soft != 0;
  h->h_settled = hd.h_settled != 0;
  h->pos_fresh = false;
  h->tree_fresh = false;
  h->have_fe = false;
  h->force_evals = hd.force_evals;
  h->list_builds = 0;
  MlpDev& m = h->mlp;
  m.trained = hd.mlp_trained != 0;
  if (m.trained) {
    HIP_TRY(h, hipMemset(m.adam_m, 0, MLP_NPARAM * sizeof(float)));
    HIP_TRY(h, hipMemset(m.adam_v, 0, MLP_NPARAM * sizeof(float)));
    m.adam_t = 0;
    mlp_sync_bf16(h->stream, m);
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipGetLastError());
  return IHS_OK;
}

}  // extern "C"
