/*
 * CPU oracle for the SPH giant-impact step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's own CPU twins (integrity-engine, Rust f64) and of the
 * LBVH arithmetic in its WGSL shaders, used as the checker for the HIP library. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Citations are relative to the reference repository root (`integrity-engine/` abbreviated `ie/`).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_pressure / or_sound_speed_sq   — eos.rs:235-377 invariants for every catalogue material
 *                                       (P(rho0,0)=0, K=A within 2%, monotone, hot limit a*rho*u,
 *                                       continuity at E_iv/E_cv, c ~ sqrt(A/rho0)) + exact parameters.
 *   or_sph_w / or_sph_dw              — closed forms: integral of W = 1, W(0,h)=8/(pi h^3),
 *                                       dW/dr = finite difference of W.
 *   or_bbox_f32 / or_morton_f32       — expand_bits known values, coincident -> equal code,
 *                                       brute-force bit interleave, exact float min/max.
 *   or_sort_pairs                     — definition (lexicographic (code, index)) vs numpy lexsort.
 *   or_karras                         — verify_tree structural invariants (gpu-bh-verify:420-459)
 *                                       + brute-force LCP split definition on small inputs.
 *   or_upsweep_f32                    — root mass / COM vs f64 sums (gpu-bh-verify:527-550),
 *                                       root AABB == bbox exactly.
 *   or_bh_walk_f32 / or_octree_f64    — theta=0.5 RMS < 1e-2, max < 1e-1 and theta->0 recovery
 *                                       vs or_direct_f64 on the N=1500 splitmix cloud
 *                                       (bhtree.rs:205-249, gpu-bh-verify:552-569).
 *   or_direct_f64                     — two-body closed form, momentum conservation, uniform
 *                                       sphere potential -3GM^2/5R.
 *   or_grid_pairs                     — pair set == brute force (neighbors.rs:103-145).
 *   or_density / or_forces / or_kdk   — brute vs grid paths agree; momentum conservation of the
 *                                       symmetric pair forces (uniform and mixed h); KDK reversibility
 *                                       at AV=0; sph-verify N=300 known-answer config runs; lattice-sum
 *                                       identities for a linear pressure field and for uniform pressure
 *                                       over 1/rho^2 linear; the two-particle P/rho^2 coefficient;
 *                                       Monaghan's viscosity on one approaching pair (mean c, 0.01 h^2,
 *                                       beta mu^2, heating); h_ij = (h_i+h_j)/2 in density and forces
 *                                       (tests/test_oracle_sph_pins.py).
 *   or_relax / or_kdk (physics)       — the reference's hydrostatic balance of a relaxed basalt sphere
 *                                       (hydrostatic.rs:442-488) and head-on collision energy
 *                                       conservation with shock heating (hydrostatic.rs:1082-1148).
 *   or_hsolve                         — NEW (no reference: the reference keeps h fixed,
 *                                       hydrostatic.rs:58-61). Parity unpinned by the reference;
 *                                       pinned by residual <= tol, bracketing/monotonicity,
 *                                       neighbour count ~48 on a lattice, quadratic convergence.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tillotson parameter block, 10 doubles: rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta. */
double or_pressure(const double* e, double rho, double u);
double or_sound_speed_sq(const double* e, double rho, double u);
double or_sound_speed_wgsl(const double* e, double rho, double u);
float or_pressure_f32(const float* e, float rho, float u);
double or_sph_w(double r, double h);
double or_sph_dw(double r, double h);

/* LBVH leg, FP32 / integer, bit-exact with the reference's CPU Morton path. xyz is n*3 floats. */
uint32_t or_expand_bits(uint32_t v);
void or_bbox_f32(const float* xyz, int64_t n, float* lo3, float* hi3);
void or_morton_f32(const float* xyz, int64_t n, const float* lo3, const float* hi3, uint32_t* codes);
void or_sort_pairs(const uint32_t* codes, int64_t n, uint32_t* sorted_codes, uint32_t* order);
void or_cluster_codes(const uint32_t* sorted_codes, int64_t n, int64_t k, uint32_t* ccodes);
void or_karras(const uint32_t* ccodes, int64_t L, uint32_t* left, uint32_t* right, uint32_t* parent);
/* fix_dups = 1: the duplicate-run split consistent with the index-augmented ranges (see .cpp). */
void or_karras_ex(const uint32_t* ccodes, int64_t L, int32_t fix_dups, uint32_t* left, uint32_t* right,
                  uint32_t* parent);
/* Moments over the Karras arena (2L-1 nodes): com4 = (com xyz, mass), bmin3, bmax3 (f32).
 * sbody = bodies permuted to sorted order, n*4 floats (x,y,z,m). */
void or_upsweep_f32(const float* sbody, int64_t n, int64_t k, const uint32_t* left, const uint32_t* right,
                    int64_t L, float* com4, float* bmin3, float* bmax3);
/* cs_gravity_bh per-particle walk; acc is n*3 floats in ORIGINAL index order. */
void or_bh_walk_f32(const float* sbody, const uint32_t* order, int64_t n, int64_t k, const uint32_t* left,
                    const uint32_t* right, const float* com4, const float* bmin3, const float* bmax3,
                    int64_t L, float theta, float soft2, float* acc);
/* The same walk on sorted targets, counting per target: accepted nodes, leaf-particle interactions and opened nodes
 * (counts: nt*3 int64); acc3 (may be NULL) receives each target's acceleration (nt*3). The algorithmic work of a
 * Barnes-Hut step is priced on these counts. */
void or_bh_walk_count(const float* sbody, int64_t n, int64_t k, const uint32_t* left, const uint32_t* right,
                      const float* com4, const float* bmin3, const float* bmax3, int64_t L, float theta,
                      float soft2, const int64_t* targets, int64_t nt, int64_t* counts, float* acc3);

/* FP64 gravity. Direct softened sum on targets (all if targets == NULL, nt = n). acc nt*3, pot nt
 * (pot may be NULL; pot_i = -sum_j G m_j / sqrt(r^2+eps^2), j != i). */
void or_direct_f64(const double* xyz, const double* m, int64_t n, double soft, const int64_t* targets,
                   int64_t nt, double* acc, double* pot, int nthreads);
/* bhtree.rs octree Barnes-Hut (leaf_cap = 1 reproduces the reference; >1 = bucketed leaves). */
void or_octree_f64(const double* xyz, const double* m, int64_t n, double theta, double soft, int leaf_cap,
                   double* acc);

/* neighbors.rs: unique pairs (i<j) within one cell of each other; returns the number written
 * (pairs as 2*count int64), or -needed if cap is too small. Only pairs with |r| < radius are kept
 * when radius > 0 (the callers' radius filter). */
int64_t or_grid_pairs(const double* xyz, int64_t n, double cell, double radius, int64_t* pairs, int64_t cap);

/* HydroBody state: xyz/vel n*3, u/h/m n, mat n (int32), eos n_mat*10. */
typedef struct {
  int64_t n;
  double* xyz;
  double* vel;
  double* u;
  double* h;
  double* m;
  const int32_t* mat;
  const double* eos;
  int32_t n_mat;
  double softening;
  double av_alpha;
  double av_beta;
  double theta;     /* octree opening angle for gravity_mode 0 */
  int32_t gravity_mode; /* 0 = octree BH (hydrostatic.rs), 1 = direct f64, 2 = none */
  int32_t cs_fd;    /* sound-speed finite difference: 0 = eos.rs (1e-4), 1 = WGSL (1e-3) */
  int32_t brute;    /* 1 = O(N^2) pair sweep (sph-verify cpu_forces), 0 = neighbour grid */
  int32_t leaf_cap; /* octree leaf bucket (1 = reference) */
  int32_t nthreads;
  int32_t h_solve;  /* 1: Newton-solve h (or_hsolve) before every density pass of or_kdk (new; the reference keeps h
                       fixed) */
  int32_t h_pad;
  double h_eta;     /* eta of the h-solve (48 neighbours: 2.2545) */
  double h_tol;     /* its |f|/rho tolerance */
} or_state;

void or_density(const or_state* s, double* rho);
void or_forces(const or_state* s, const double* rho, double* acc, double* dudt);
double or_courant_dt(const or_state* s, const double* rho, double cfl);
/* KDK steps (hydrostatic.rs:345-359 / sph-verify CpuState::step). dt > 0: fixed; dt <= 0: Courant
 * dt with cfl (impact-run loop: density -> dt -> step). Returns the total time advanced. */
double or_kdk(or_state* s, double dt, double cfl, int32_t steps);
/* Damped relaxation steps (hydrostatic.rs:283-291): density, gravity + pressure-only SPH accelerations (no
 * artificial viscosity), v = (v + a dt) damp, x += v dt. */
void or_relax(or_state* s, double dt, double damp, int32_t steps);
/* FP64 energy (KE, IE, PE with direct O(N^2) softened PE), momentum and angular momentum about
 * the origin: out[0..8] = KE, IE, PE, px, py, pz, Lx, Ly, Lz. */
void or_energy(const or_state* s, double* out);

/* NEW (no reference counterpart): Newton smoothing-length solve on the own-h gather density
 * rho~_i(h) = sum_j m_j W(r_ij, h) (self included), f(h) = rho~ - m_i (eta/h)^3 = 0.
 * h is in/out (initial guess in). Returns the number of particles that did not converge.
 * iters (may be NULL) receives per-particle Newton updates; resid the final |f|/rho~. */
int64_t or_hsolve(const double* xyz, const double* m, int64_t n, double eta, double tol, int32_t max_iter,
                  double h_max, double* h, int32_t* iters, double* resid, int nthreads);
/* Residual |f(h)|/rho~(h) and neighbour count (r < h_i, j != i) at given h. */
void or_hresidual(const double* xyz, const double* m, int64_t n, double eta, const double* h, double* resid,
                  int32_t* nngb, int nthreads);

#ifdef __cplusplus
}
#endif
