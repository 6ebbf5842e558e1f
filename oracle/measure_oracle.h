/*
 * CPU oracle for the disk / remnant measurement and accretion clumps — TEST INFRASTRUCTURE ONLY.
 * See measure_oracle.cpp for the reference citations. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never does.
 *
 * Parity pins (tests/test_oracle_measure.py):
 *   or_sum_oi        — equals math.fsum (the correctly rounded sum) on random and cancelling sequences, and is
 *                      independent of the order of its terms.
 *   or_perigee       — closed forms: a Kepler ellipse of known a, e (q = a(1-e)), a circular orbit (q = r), a
 *                      radial fall (None), a hyperbola (q = h^2/mu/(1+e) for the engine variant, None for
 *                      impact-run), a zero radius.
 *   or_measure_ir    — category masses partition m_total; equal masses put r_remnant at the ceil(0.85 n)-th
 *                      smallest radius; the FoF components equal scipy's connected components of the brute-force
 *                      link graph; clump potential equals a numpy pair sum; a cold bound blob outside Roche is the
 *                      Moon candidate, and a hot or inside-Roche one is not.
 *   or_engine_frame  — the same remnant and partition checks; a bound-only disk is a subset of the any-orbit disk.
 *   or_body_bulk     — a clean ball: the plain centre of mass and the ceil(0.98 n)-th radius; strays far outside
 *                      (1% of the mass at 20 radii) leave the centre and radius of the ball, where a plain mean
 *                      moves by a fifth of a radius.
 *   or_assemble_n    — each placed body's bulk centre lands on its offset, velocities are the bulk velocity plus
 *                      spin x r, dropped provenances, softening = min h / 4 and the dt formula.
 *   or_find_clumps   — accretion.rs's own tests (two cold blobs accrete with mass, momentum and centre of mass
 *                      conserved; the Roche gate; a hot group is unbound), grid pairs == brute-force pairs above
 *                      512 bodies, components == scipy.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float pos[3];
  float h;
  float vel[3];
  float u;
  float mass;
  uint32_t mat;
  float rho;
  uint32_t prov;
} or_particle;

typedef struct {
  double earth_frac, disk_kg, e_disk, t_disk, e_rem, t_rem, e_esc, t_esc, r_remnant;
  double clump_kg, clump_earth_frac;
  int64_t clump_n, n_clumps, n_bound;
  double m_total, m_remnant, com[3], v_com[3], link;
  int64_t n_disk;
} or_measure;

typedef struct {
  int64_t first;
  int64_t n_members;
  double mass;
  double com_pos[3];
  double com_vel[3];
  double rho;
  double radius;
  double internal_ke;
  double self_pe;
  double earth_mass;
  int32_t bound;
  int32_t outside_roche;
} or_clump;

double or_sum_oi(const double* terms, int64_t n);
/* which = 0: impact-run perigee (main.rs:232-247); 1: orbit::perigee (orbit.rs:96-116). Returns 1 for Some. */
int or_perigee(int32_t which, const double* rel_p, const double* rel_v, double mu, double* pg);
/* cls (may be NULL): 0 remnant, 1 disk, 2 escaped, per particle. */
void or_measure_ir(const or_particle* P, int64_t n, int64_t n_earth, or_measure* out, int32_t* cls);
/* frame: m_total, com[3], v_com[3], r_remnant, m_remnant, mu, disk_mass, link (12 doubles);
 * esc_e_t (may be NULL): e_disk, t_disk, escaped. cls as above (escaped = unbound). */
void or_engine_frame(const or_particle* P, int64_t n, int32_t require_bound, double* frame, double* esc_e_t,
                     int32_t* cls);
/* Returns the number of clumps of positive mass; writes up to cap of them ordered by lowest member; label (may be
 * NULL) receives each body's clump index or -1. */
int64_t or_find_clumps(int64_t n, const double* pos3, const double* vel3, const double* mass, const double* rho,
                       double linking_length, double g, double softening, const double* central3,
                       double central_mass, double central_radius, or_clump* out, int64_t cap, int64_t* label);

/* body_bulk (gpu_sph.rs:279-316): out5 = centre xyz, mass, radius. */
void or_body_bulk(const or_particle* P, int64_t n, uint32_t prov, double* out5);
/* assemble_from_relaxed_n (gpu_sph.rs:318-339): placements = nbody x 9 (offset, vel, spin); returns the kept count. */
int64_t or_assemble_n(const or_particle* P, int64_t n, const double* placements, int32_t nbody, or_particle* out,
                      float* softening, float* dt);

#ifdef __cplusplus
}
#endif
