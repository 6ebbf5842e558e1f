/*
 * libihs_sph — MI355X (gfx950) SPH + self-gravity giant-impact step behind a C ABI.
 *
 * The entry points mirror the reference's step interface, `GpuSph`
 * (integrity-engine/crates/engine/src/gpu_sph.rs:742-922), and its gravity backend, `GravityField` /
 * `GpuGravity` (integrity-engine/crates/engine/src/gpu_gravity.rs:181-230,532-626). Conventions:
 *   - every function returns an int status (IHS_OK = 0, < 0 = error) unless documented otherwise,
 *     never aborts, and leaves a message for ihs_last_error();
 *   - one handle = one HIP device + one stream; thread-compatible, not thread-safe (like GpuSph);
 *   - the caller owns host memory, the library owns device memory;
 *   - particles cross the boundary as the reference's 48-byte AoS record; the device keeps SoA.
 */
#ifndef IHS_SPH_H
#define IHS_SPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IHS_ABI_VERSION 1

#define IHS_OK 0
#define IHS_ERR_ARG (-1)
#define IHS_ERR_HIP (-2)
#define IHS_ERR_CAPACITY (-3)
#define IHS_ERR_STATE (-4)
#define IHS_EAGAIN (-11)

/* SphParticle, gpu_sph.rs:25-37 (std430, 48 bytes). */
typedef struct {
  float pos[3];
  float h;
  float vel[3];
  float u;
  float mass;
  uint32_t mat;
  float rho;
  uint32_t prov;
} ihs_particle;

/* SphEos, gpu_sph.rs:39-55 (48 bytes): Tillotson rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta. */
typedef struct {
  float rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta, _p0, _p1;
} ihs_eos;

enum { IHS_PREC_FP32 = 0, IHS_PREC_SPLIT = 1 };
/* Smoothing length: FIXED keeps the uploaded h (the reference's behaviour, hydrostatic.rs:58-61);
 * NEWTON solves rho~(h) = m (eta/h)^3 for n_ngb neighbours each force evaluation; LEARNED seeds
 * that Newton solve with the built-in network (Newton checks every seed). */
enum { IHS_H_FIXED = 0, IHS_H_NEWTON = 1, IHS_H_LEARNED = 2 };
enum { IHS_GRAVITY_DIRECT = 0, IHS_GRAVITY_TREE = 1 };

typedef struct {
  int32_t capacity;      /* max particles (GpuSph::new capacity) */
  int32_t leaf_size;     /* LBVH leaf bucket K: 1,2,4,...,64 (default 16) */
  int32_t precision;     /* IHS_PREC_* (default FP32) */
  int32_t h_mode;        /* IHS_H_* (default NEWTON) */
  float theta;           /* Barnes-Hut opening angle (0.5) */
  float cfl;             /* Courant factor for adaptive dt (0.1, impact-run:583) */
  float n_ngb;           /* target neighbours within h (48) */
  float h_tol;           /* density-consistency tolerance |f|/rho (1e-4) */
  int32_t h_max_iter;    /* Newton iteration cap (40) */
  float h_margin;        /* neighbour search radius = h_margin * previous h (1.25) */
  float h_cap_factor;    /* search radius never exceeds h_cap_factor * max uploaded h (4) */
  int32_t nbr_capacity;  /* neighbour-list entries per particle (192) */
  int32_t use_graph;     /* capture the KDK step in a hipGraph (1) */
  float h_margin_seeded; /* LEARNED: search radius = h_margin_seeded * seed after a drift (1.03); a wider-radius
                            pass (h_margin) runs only when some solve reaches it */
  int32_t group_max;     /* particles per gravity-walk group, leaf_size..128 (64): one wavefront walks one group,
                            each lane owning one (<= 64) or two (<= 128) particles */
  int32_t quadrupole;    /* 1: accepted cells act with monopole + traceless quadrupole (default); 0: monopole only,
                            as the reference (bh_gravity.wgsl:375-386) */
  int32_t reuse_positions; /* 1: the first evaluation of a KDK step, at the positions the previous step's second
                              evaluation saw, keeps that evaluation's tree, neighbours, h, density and gravity and
                              recomputes only the EOS and the velocity-dependent SPH forces (0: recompute all, as
                              the reference does, sph_step.wgsl:213-237) */
  int32_t fused_walk;    /* 1: the gravity walk also writes the neighbour lists (one traversal of the tree per
                            evaluation instead of two; the same neighbour sets, listed in walk order), when walk and
                            neighbour groups coincide (group_max 64); 0: a separate neighbour build (default 1) */
  float theta_mono;      /* with quadrupole: a cell that also passes the opening test at this smaller angle acts as
                            a monopole (its quadrupole term is below the error of the cells accepted at theta);
                            0: every accepted cell takes its quadrupole (default 0) */
  float quad_tol;        /* with quadrupole: a cell whose quadrupole term is bounded below quad_tol times its
                            monopole term (3.5 |q|_F / r^2 at the group's nearest point) acts as a monopole; 0: off */
  float theta_settled;   /* > 0: an evaluation at positions and h unchanged since the previous evaluation (the first
                            of a KDK step, after a kick) keeps that evaluation's neighbour lists and walk groups,
                            which the unchanged positions and h reproduce, and sums gravity with a walk that builds
                            no lists at this opening angle (<= theta; without the neighbour search's opened nodes it
                            needs a smaller angle for the same force error); 0: every evaluation rebuilds the lists
                            (default 0) */
  int32_t reserved[3];
} ihs_config;

typedef struct ihs_handle ihs_handle;

typedef struct {
  double ke, ie, pe;      /* FP64: kinetic, internal, gravitational (1/2 sum m phi, tree potential; after KDK
                             steps, which sum no potential, one gravity walk over the last tree forms it) */
  double p[3];            /* FP64 momentum */
  double L[3];            /* FP64 angular momentum about the origin */
  double t;               /* simulated time (scenario 0 of a batch) */
  double dt;              /* last step's dt (scenario 0 of a batch) */
  int64_t steps;          /* integration steps taken (scenario 0 of a batch) */
  int64_t force_evals;    /* force evaluations since the last counter reset */
  double mean_h_iters;    /* Newton updates per particle per force evaluation */
  double mean_h_evals;    /* density evaluations per particle per force evaluation */
  double mean_neighbours; /* neighbour-list entries per particle */
  uint32_t nbr_overflow, walk_overflow, h_capped, h_unconverged, nonfinite;
  uint32_t tie_runs;       /* last evaluation's sort: runs of equal Morton codes longer than 64 (sorted by a
                              workgroup each) */
  int64_t grav_cell_interactions; /* target x accepted-node interactions evaluated (since reset) */
  int64_t grav_pp_interactions;   /* target x leaf-particle interactions evaluated (since reset) */
  int64_t nbr_candidates;         /* target x candidate tests of the neighbour search (since reset) */
  int64_t walk_nodes;             /* tree nodes popped by all group walks (since reset) */
} ihs_stats;

/* Pipeline stages timed by ihs_profile (HIP events on the handle's stream). */
enum {
  IHS_STAGE_SORT = 0,    /* bbox, Morton keys, radix sort, tie order */
  IHS_STAGE_GATHER = 1,  /* state permutation into Morton order */
  IHS_STAGE_TREE = 2,    /* Karras build + upsweep */
  IHS_STAGE_NBR = 3,     /* neighbour lists */
  IHS_STAGE_HSOLVE = 4,  /* learned seed + Newton h-solve */
  IHS_STAGE_DENSITY = 5, /* density + EOS + Courant signal */
  IHS_STAGE_FORCES = 6,  /* SPH forces + du/dt */
  IHS_STAGE_GRAVITY = 7, /* Barnes-Hut walk */
  IHS_STAGE_KICK = 8,    /* kick / drift */
  IHS_STAGE_COUNT = 9
};

/* Debug/inspection fields for ihs_debug_field (all returned in particle-id order). */
enum {
  IHS_FIELD_H = 0,       /* float */
  IHS_FIELD_RHO = 1,     /* float */
  IHS_FIELD_RPC = 2,     /* float4: rho, P/rho^2, c, P */
  IHS_FIELD_ASPH = 3,    /* float4: SPH acceleration xyz, du/dt */
  IHS_FIELD_AGRAV = 4,   /* float4: gravitational acceleration xyz, potential (after ihs_force_eval or ihs_relax;
                            KDK steps sum no potential and leave 0) */
  IHS_FIELD_HITER = 5,   /* uint32: Newton updates in the last h-solve */
  IHS_FIELD_NCOUNT = 6,  /* uint32: neighbour-list length */
  IHS_FIELD_FEAT = 7,    /* float4: n_ngb within h, div v, |grad rho|, 0 */
  IHS_FIELD_POS64 = 8,   /* double[3] */
  IHS_FIELD_HS = 9       /* float: searched radius */
};

void ihs_default_config(ihs_config* cfg);
int ihs_abi_version(void);
/* GpuSph::new (gpu_sph.rs:767). */
int ihs_create(int device, const ihs_config* cfg, ihs_handle** out);
int ihs_destroy(ihs_handle* h);
const char* ihs_last_error(const ihs_handle* h);
/* GpuSph::upload (gpu_sph.rs:839-850): n beyond capacity is an error (the reference clamps silently). */
int ihs_upload(ihs_handle* h, const ihs_particle* p, int64_t n, const ihs_eos* eos, int32_t n_mat, float softening);
/* Batched scenarios (impact-parameter sweeps): nseg independent scenarios concatenated in p, counts[s] particles
 * each, with their own softening. Every pass of a step covers all of them in one launch: one tree per scenario,
 * walks confined to it, a Courant dt per scenario (ihs_segment_clock). With nseg > 1 each count must be a multiple
 * of the leaf size. ihs_upload is the nseg = 1 case. No counterpart in the reference, which runs scenarios one at
 * a time (impact-run:645-671). */
#define IHS_MAX_SEGMENTS 1024
int ihs_upload_batch(ihs_handle* h, const ihs_particle* p, const int64_t* counts, int32_t nseg, const ihs_eos* eos,
                     int32_t n_mat, const float* softening);
/* Simulated time, last dt and step count of scenario seg. */
int ihs_segment_clock(ihs_handle* h, int32_t seg, double* t, double* dt, int64_t* steps);
/* GpuSph::set_dt (gpu_sph.rs:853-857). dt <= 0 selects the adaptive Courant dt (impact-run:403-410). */
int ihs_set_dt(ihs_handle* h, double dt, float damp);
/* GpuSph::set_av (gpu_sph.rs:864-868). */
int ihs_set_av(ihs_handle* h, float alpha, float beta);
/* SphParams::omega for relaxation in a rotating frame (gpu_sph.rs:68-76, sph_step.wgsl:252-256). */
int ihs_set_omega(ihs_handle* h, float omega);
/* GpuSph::encode_kdk (gpu_sph.rs:901-908): substeps KDK steps, two force evaluations each. Enqueued on
 * the handle's stream; returns without waiting. */
int ihs_step_kdk(ihs_handle* h, int32_t substeps);
/* GpuSph::encode_relax (gpu_sph.rs:893-898): steps x (force evaluation + damped kick/drift). */
int ihs_relax(ihs_handle* h, int32_t steps);
/* One force evaluation at the current state, no integration (for inspection). */
int ihs_force_eval(ihs_handle* h);
/* The current state in upload order (the particle count written, <= cap). blocking = 1 waits for the queued work;
 * blocking = 0 returns IHS_EAGAIN while the stream is busy. For a frame loop that keeps queuing steps, use the
 * two-phase pair below. */
int64_t ihs_readback(ihs_handle* h, ihs_particle* out, int64_t cap, int32_t blocking);
/* GpuSph::begin_readback (gpu_store.rs:126-145): enqueues, behind the work already queued, a copy of the state as
 * it will then be into a library-owned pinned host buffer, and returns without waiting. A no-op when nothing is
 * uploaded or a snapshot is already in flight, as in the reference. */
int ihs_begin_readback(ihs_handle* h);
/* GpuSph::take_readback (gpu_store.rs:147-152): the snapshot of the last ihs_begin_readback once its copy has
 * completed (the particle count written, <= cap, in upload order), however many steps were queued after it;
 * IHS_EAGAIN while the copy is pending or when no snapshot was begun (the reference's None). Never waits. */
int64_t ihs_take_readback(ihs_handle* h, ihs_particle* out, int64_t cap);
/* GpuSph::count (gpu_sph.rs:870). */
int64_t ihs_count(const ihs_handle* h);
int ihs_sync(ihs_handle* h);
/* FP64 energy / momentum reductions and solver counters (blocks). reset=1 clears the counters. The energies,
 * momentum and angular momentum are totals over every scenario of a batch (ihs_segment_energy gives one
 * scenario's); the clock fields are scenario 0's. */
int ihs_stats_get(ihs_handle* h, ihs_stats* out, int32_t reset);
/* One scenario's FP64 sums (blocks), as gpu_sph.rs:624-649 total_energy and impact-run:528-536 over its
 * particles: out9 = KE, IE, PE (1/2 sum m phi, the last evaluation's tree potential), p xyz, L xyz about the
 * origin. Needs a force evaluation of the current state. */
int ihs_segment_energy(ihs_handle* h, int32_t seg, double* out9);
/* GravityField::accelerations / GpuGravity::accelerations[_tree] (gpu_gravity.rs:181-230,556-578):
 * softened self-gravity on n bodies (pos3 and acc3 are n*3 doubles). mode = IHS_GRAVITY_*;
 * pot (may be NULL) receives the potential. Synchronous. Uses the handle's device and leaf size, and its tree and
 * working buffers: the particles need a force evaluation (or a step) before their energies or tree are read. */
int ihs_gravity(ihs_handle* h, const double* pos3, const double* mass, int64_t n, double softening, double theta,
                int32_t mode, double* acc3, double* pot);
/* Inspection for parity tests: the LBVH of the last force evaluation. sorted_codes/order have n
 * entries (order = particle ids in Morton order), ccodes L, left/right/parent 2L-1 (children of
 * leaves are 0, the root's parent is 0xffffffff). Returns L. */
int64_t ihs_debug_tree(ihs_handle* h, uint32_t* sorted_codes, uint32_t* order, uint32_t* ccodes, uint32_t* left,
                       uint32_t* right, uint32_t* parent);
int ihs_debug_field(ihs_handle* h, int32_t field, void* out);
/* Force-error probe: FP64 direct-sum gravitational accelerations (G, the upload's softening; the particle itself
 * excluded) at the particles with the given upload-order ids, from the current positions (FP64 in split
 * precision). Compare with IHS_FIELD_AGRAV of the last force evaluation (gpu-bh-verify's cpu_direct check,
 * ie/tools/gpu-bh-verify/src/main.rs:140-158, on the device). acc3 receives nt*3 doubles. Synchronous. */
int ihs_direct_f64(ihs_handle* h, const uint32_t* ids, int64_t nt, double* acc3);
/* ihs_direct_f64 without any device allocation: no counterpart in the reference. For callers that cannot
 * allocate (code under HIP graph capture, tight serving loops that check the walk's force error as they run,
 * harnesses that forbid hipMalloc). The per-call staging lives in a static 2 MiB array on each GPU (held until the
 * process ends), so a call takes at most 65,536 targets (IHS_ERR_CAPACITY beyond). Every handle on one GPU shares
 * that array: calls that overlap on one GPU (from several threads) race. Synchronous; the results equal
 * ihs_direct_f64's bit for bit. */
int ihs_direct_f64_noalloc(ihs_handle* h, const uint32_t* ids, int64_t nt, double* acc3);
/* Per-stage timing: enable=1 makes ihs_step_kdk launch directly (no graph) with HIP events around every
 * stage; ihs_profile_read returns accumulated milliseconds and launch counts per IHS_STAGE_*. */
int ihs_profile(ihs_handle* h, int32_t enable);
int ihs_profile_read(ihs_handle* h, double* ms, int64_t* calls, int32_t reset);

/* Learned smoothing-length seed (no reference counterpart; the reference keeps h fixed). A 6->32->32->1 GELU
 * network, IHS_MLP_NPARAM parameters laid out W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3, used as BF16
 * with FP32 accumulation. Features per particle (previous evaluation's state): x0 = 100 div v dt, ln(rho/rho0),
 * (n_ngb - 47)/47, |grad rho| h / rho, ln(h / (eta (m/rho0)^(1/3))), material id. Output y: the correction to
 * the continuity estimate, so that the seed is h_prev exp((x0/3 + y)/100) (clamped to [h/1.2, 1.2 h] and to the
 * searched radius); a recorded target is 100 ln(h_new/h_prev) - x0/3. In IHS_H_LEARNED mode the second force
 * evaluation of each KDK step starts the Newton solve from that seed; Newton checks every seed. */
#define IHS_MLP_NPARAM 1313
#define IHS_MLP_IN 6
#define IHS_MLP_SAMPLE 8 /* floats per recorded sample: 6 features, target, 0 */
/* Sets the FP32 master weights (and their BF16 copy), clears the Adam state and enables the seed. */
int ihs_mlp_set_weights(ihs_handle* h, const float* w, int64_t nparam);
int ihs_mlp_get_weights(ihs_handle* h, float* w, int64_t nparam);
/* max_samples > 0: (re)start recording (features, converged target) pairs at every post-drift Newton solve
 * (KDK steps then launch directly, not from a graph) until max_samples rows are held; 0 stops recording and
 * keeps the rows. Needs h_mode NEWTON. */
int ihs_mlp_record(ihs_handle* h, int64_t max_samples);
/* Copies up to cap recorded rows (IHS_MLP_SAMPLE floats each) to out (may be NULL); returns the row count. */
int64_t ihs_mlp_samples(ihs_handle* h, float* out, int64_t cap);
/* Adam on the mean squared error over the recorded rows (rows_host == NULL) or over n_rows host rows
 * (IHS_MLP_SAMPLE floats each), on the device. Each step draws `batch` rows as (off + k * stride) mod rows with
 * off and a stride coprime with rows drawn from seed; the step size decays from lr to lr/20 on a cosine.
 * loss (may be NULL) receives the mean 0.5 * err^2 of each step's batch. Requires weights to start from. */
int ihs_mlp_train(ihs_handle* h, const float* rows_host, int64_t n_rows, int32_t steps, int32_t batch, float lr,
                  uint64_t seed, float* loss);
/* Exact gradient of mean(0.5 (f(x) - y)^2) over n host samples (x: n*6, y: n) at the current BF16 weights:
 * grad receives IHS_MLP_NPARAM values then the loss. */
int ihs_mlp_gradient(ihs_handle* h, const float* x, const float* y, int64_t n, float* grad);
/* Network output for n host feature rows (x: n*6) into y (n). */
int ihs_mlp_predict(ihs_handle* h, const float* x, int64_t n, float* y);

/* ---- Disk / remnant measurement and accretion clumps, on the device-resident state ----
 * All of these read the particles of one scenario (segment, 0 for a single upload) as the readback would hand
 * them over (FP32 fields, upload order) without copying them to the host. Sums are double-double (about 104
 * bits) in a fixed order, so a result does not depend on the storage order of the particles. */

/* One friends-of-friends clump (accretion.rs:29-41 Clump). `first` is the clump's lowest member index (members
 * are reported by label arrays, not lists); clumps are ordered by `first`. */
typedef struct {
  int64_t first;
  int64_t n_members;
  double mass;
  double com_pos[3];
  double com_vel[3];
  double rho;          /* mass / sum(m_i / rho_i) */
  double radius;       /* (3 mass / (4 pi rho))^(1/3) */
  double internal_ke;  /* sum 1/2 m_i |v_i - v_com|^2 */
  double self_pe;      /* -sum_{i<j} G m_i m_j / r_ij */
  double earth_mass;   /* members from the target body (index < n_earth, or prov 0) */
  int32_t bound;       /* internal_ke + self_pe < 0 */
  int32_t outside_roche;
} ihs_clump;

/* impact-run `Measure` (ie/tools/impact-run/src/main.rs:715-731) plus the intermediate values it is built from. */
typedef struct {
  double earth_frac; /* % of the orbiting disk from the target */
  double disk_kg, e_disk, t_disk, e_rem, t_rem, e_esc, t_esc, r_remnant;
  double clump_kg, clump_earth_frac;
  int64_t clump_n, n_clumps, n_bound;
  double m_total, m_remnant, com[3], v_com[3], link;
  int64_t n_disk;
} ihs_measure;

/* impact-run `measure(body, n_earth)` (main.rs:732-778) with `moon_candidate` (main.rs:802-878): remnant = the
 * smallest radius about the centre of mass enclosing 85% of the mass; a particle escapes when its orbit about the
 * remnant is unbound, is disk when its perigee lies above the remnant radius, remnant otherwise; target particles
 * are those with index < n_earth. The Moon candidate is the most massive friends-of-friends clump of the disk
 * (linking length twice the disk's mean h) that is self-bound and outside the fluid Roche limit; ties in mass go
 * to the clump with the lowest member index (the reference leaves them to hash-map order). Synchronous. */
int ihs_measure_impact(ihs_handle* h, int32_t segment, int64_t n_earth, ihs_measure* out);

/* gpu_sph.rs `disk_stats_json` (gpu_sph.rs:399-469) as numbers: disk (M_moon), earth_pct, remnant_km,
 * escaped (M_moon), moon (M_moon) are the JSON fields; the rest are the same quantities in SI units. */
typedef struct {
  double disk, earth_pct, remnant_km, escaped, moon;
  double e_disk, t_disk, esc, r_remnant, m_remnant, biggest, link;
  int64_t n_disk, n_clumps;
} ihs_disk_summary;
int ihs_disk_stats(ihs_handle* h, int32_t segment, ihs_disk_summary* out);

/* The orbiting disk and its clumps as gpu_sph.rs moonlet_bodies / largest_moonlet_orbit / disk_moonlets select
 * them (gpu_sph.rs:478-640): disk = perigee above the remnant radius (require_bound = 0: hyperbolic passes
 * included, as those functions do; 1: bound orbits only, as disk_stats_json does); clumps by
 * accretion::find_clumps (linking length twice the mean h of all particles, G, softening 1e4 m, the remnant as
 * the central body). frame receives m_total, com[3], v_com[3], r_remnant, m_remnant, mu, disk_mass, link, n_disk
 * (13 doubles). Writes up to cap clumps; *n_clumps receives the total. */
int ihs_disk_clumps(ihs_handle* h, int32_t segment, int32_t require_bound, double* frame, ihs_clump* clumps,
                    int64_t cap, int64_t* n_clumps);

/* ---- IC assembly on the device: the steps around the GPU relaxation (gpu_sph.rs:198-339) ----
 * A caller uploads the bodies far apart (each body's particles tagged with its source index as prov), relaxes
 * them with ihs_relax (artificial viscosity off, ihs_set_av(h, 0, 0), as the reference's relax discipline asks),
 * then places them on the collision in place, with no readback and no second upload. */

/* body_bulk (gpu_sph.rs:279-316): the robust centre of mass, total mass and the radius enclosing 98% of the mass
 * of the particles with provenance prov, after two passes that re-centre on the particles within 1.5 radii.
 * out5 = centre x, y, z, mass, radius (all zero for a provenance with no particles). Synchronous. */
int ihs_body_bulk(ihs_handle* h, int32_t segment, uint32_t prov, double* out5);

/* assemble_from_relaxed_n (gpu_sph.rs:318-339) in place: every particle of body k = prov is re-centred on its
 * body_bulk centre, moved to offset_k and set moving at vel_k + spin_k x (r - centre). placements holds nbody x
 * 9 doubles: offset xyz, velocity xyz, spin (angular velocity) xyz. Positions and velocities are rounded to FP32,
 * as the reference's output records are. Returns the reference's softening (min h / 4, now the scenario's
 * softening) and its shock-safe dt, 0.01 min h / (20 km/s + the fastest placement speed), for ihs_set_dt. The
 * scenario's clock restarts at 0. Particles whose prov has no placement are an error (the reference drops them).
 * Synchronous. */
#define IHS_MAX_BODIES 64
int ihs_assemble(ihs_handle* h, int32_t segment, const double* placements, int32_t nbody, float* softening,
                 float* dt);

/* ---- Checkpoint / resume (no counterpart in the reference, whose snapshots stay in memory, impact-run:394-434) ----
 * A versioned binary image of the device state: an "IHSCKPT1" header holding the ihs_config, step control and
 * flags, then the scenario tables, the EOS table, the network weights and the master particle state (FP64
 * positions and u in split precision) in its storage order. A run resumed from a checkpoint continues bit for
 * bit as the run that wrote it (with reuse_positions off). The header holds, in order, the magic, version,
 * header size, n, total size, scenario count and precision, then at byte 40 the ihs_config to create the resuming
 * handle with (any capacity >= n will do). */
int64_t ihs_checkpoint_bytes(ihs_handle* h);
/* Writes the checkpoint into buf (cap bytes); returns the bytes written. Synchronous. */
int64_t ihs_checkpoint_save(ihs_handle* h, void* buf, int64_t cap);
/* Restores a checkpoint into a handle created with the same configuration. Synchronous. Before anything reaches
 * the device, the sections kernels index with are checked on the host (particle ids a permutation of 0..n-1,
 * material indices inside the EOS table the checkpoint was written with, a scenario table that spans the
 * particles in leaf-aligned increasing runs, h > 0, finite softening and weights), and the per-particle scenario
 * index and tree roots are rebuilt from the scenario table; a corrupt file is refused (IHS_ERR_ARG) and the handle
 * then holds no particles. */
int ihs_checkpoint_load(ihs_handle* h, const void* buf, int64_t bytes);

/* accretion::find_clumps (accretion.rs:73-155) on n host bodies (pos3, vel3: n*3 doubles; mass, rho: n):
 * friends-of-friends within linking_length, then each clump's mass, centre, density, radius, internal kinetic
 * and softened self-potential energy, boundness and the Roche gate against the central body. Clumps of zero
 * mass are dropped. Writes up to cap clumps (ordered by lowest member), *n_clumps receives the total, and
 * label (may be NULL; n entries) receives each body's clump index (-1 for a dropped clump). Synchronous. */
int ihs_find_clumps(ihs_handle* h, int64_t n, const double* pos3, const double* vel3, const double* mass,
                    const double* rho, double linking_length, double g, double softening, const double* central3,
                    double central_mass, double central_radius, ihs_clump* clumps, int64_t cap, int64_t* n_clumps,
                    int64_t* label);

#ifdef __cplusplus
}
#endif
#endif
