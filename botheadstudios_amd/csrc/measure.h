// Disk / remnant measurement and friends-of-friends clumps: the interface between ihs_api.hip and k_measure.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ihs_sph.h"
#include "launch.h"

// Which reference function's orbit classification a frame follows.
enum MeasureMode : int {
  MM_IMPACT_RUN = 0,    // impact-run measure (main.rs:732-778): perigee None when unbound; target = index < n_earth
  MM_ENGINE_BOUND = 1,  // gpu_sph.rs disk_stats_json: escaped = !is_bound; disk = bound and perigee > r_remnant
  MM_ENGINE_ANY = 2,    // gpu_sph.rs moonlet_bodies & co: disk = orbit::perigee (hyperbolic too) > r_remnant
};

// Category slots of a frame's mass sums.
enum { CAT_DISK_E = 0, CAT_DISK_T, CAT_REM_E, CAT_REM_T, CAT_ESC_E, CAT_ESC_T, CAT_ESC_ALL, CAT_DISK_H, CAT_COUNT };

// Phase 1 results (device struct, copied to the host at the end of the phase).
struct MeasureFrame {
  double m_total;
  double com[3];
  double v_com[3];
  double sum_h;  // sum of h over all particles
  double r_remnant;
  double m_remnant;
  double cat[CAT_COUNT];
  uint32_t n_disk;
  uint32_t n_clumps;  // phase 2: clump count
  uint32_t n_items;   // phase 2: pair-potential work items
  uint32_t n_sel;     // body_bulk: bodies selected by the last filter
};

// Phase 2 parameters.
struct ClumpQuery {
  int impact_run = 0;  // moon_candidate arithmetic: FP32 position differences, r >= 1 m, fallback density 2700
  double link = 0.0;
  double g = 6.674e-11;
  double soft2 = 0.0;
  double central[3] = {0.0, 0.0, 0.0};
  double central_mass = 0.0;
  double central_radius = 0.0;
};

// Phase 2 summary over clumps of at least two members.
struct ClumpSummary {
  uint32_t n2;              // clumps with >= 2 members
  uint32_t n_bound;         // ... of which self-bound
  int32_t best;             // most massive bound clump outside Roche (lowest first member on ties), -1 if none
  uint32_t pad;
  unsigned long long best_mass;  // bit pattern of that clump's mass (positive doubles order as integers)
};

// Device pointers (passed to kernels by value).
struct MeasureDev {
  // sort
  uint32_t* rk[2] = {nullptr, nullptr};
  uint32_t* rv[2] = {nullptr, nullptr};
  uint32_t* khi = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* dtot = nullptr;
  uint64_t* key64 = nullptr;
  uint64_t* skey = nullptr;
  // reductions and scans
  void* part = nullptr;  // per-block double-double partial sums
  void* tsum = nullptr;  // per-tile double-double sums
  uint32_t* flag = nullptr;
  uint32_t* scan = nullptr;
  uint32_t* bsum = nullptr;
  MeasureFrame* frame = nullptr;
  ClumpSummary* summary = nullptr;
  uint64_t* bbox = nullptr;  // 6 order-preserving keys of the clump set's box
  // body set (input order) and its clump-ordered copy
  double *ax = nullptr, *ay = nullptr, *az = nullptr, *avx = nullptr, *avy = nullptr, *avz = nullptr;
  double *am = nullptr, *arho = nullptr;
  float *axf = nullptr, *ayf = nullptr, *azf = nullptr;
  uint32_t *aearth = nullptr, *asrc = nullptr;
  double *bx = nullptr, *by = nullptr, *bz = nullptr, *bvx = nullptr, *bvy = nullptr, *bvz = nullptr;
  double *bm = nullptr, *brho = nullptr;
  float *bxf = nullptr, *byf = nullptr, *bzf = nullptr;
  uint32_t *bearth = nullptr, *bsrc = nullptr;
  uint32_t* parent = nullptr;
  uint32_t* label = nullptr;
  uint32_t* cstart = nullptr;  // C + 1
  uint32_t* ioff = nullptr;    // C + 1
  ihs_clump* crec = nullptr;
  void* pe_part = nullptr;     // per work item
  SphParticle48* rec[2] = {nullptr, nullptr};  // body_bulk: the body's records, the clipped subset
};

struct MeasureBufs : MeasureDev {
  int64_t cap = 0;
  std::vector<void*> allocs;
  int64_t pe_cap = 0;  // pair-potential items pe_part holds (grown on demand)
};

hipError_t measure_reserve(MeasureBufs& B, int64_t cap);
void measure_release(MeasureBufs& B);

// Phase 1 over the snapshot P[0, n) (upload order): centre of mass, remnant, classification; the disk set is left
// in the body-set arrays. out receives the frame.
hipError_t measure_frame(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, int mode, int64_t n_earth,
                         MeasureFrame* out);

// Phase 2 over the k bodies in the body-set arrays. all (may be null) receives every clump in order of lowest
// member; labels (may be null) each body's clump index; sum the summary.
hipError_t clump_set(hipStream_t s, MeasureBufs& B, int64_t k, const ClumpQuery& q, std::vector<ihs_clump>* all,
                     std::vector<int32_t>* labels, ClumpSummary* sum, ihs_clump* best);

// body_bulk (gpu_sph.rs:279-316) over the records of provenance `prov` in P[0, n): out = centre xyz, mass, radius.
hipError_t body_bulk(hipStream_t s, MeasureBufs& B, const SphParticle48* P, int64_t n, uint32_t prov, double out[5]);

// Uploads n host bodies into the body-set arrays (find_clumps input).
hipError_t clump_upload(hipStream_t s, MeasureBufs& B, int64_t n, const double* pos3, const double* vel3,
                        const double* mass, const double* rho);
