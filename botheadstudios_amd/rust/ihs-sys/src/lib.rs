//! Raw bindings to `libihs_sph`, the MI355X (gfx950) SPH + self-gravity giant-impact step behind a C ABI
//! (`include/ihs_sph.h`).
//!
//! One declaration per header entry point, the header's constants, and `#[repr(C)]` mirrors of its structs. The
//! entry points replace the reference's `GpuSph` (ie/crates/engine/src/gpu_sph.rs:742-922) and its gravity
//! backend (ie/crates/engine/src/gpu_gravity.rs:181-230,532-626). Conventions (from the header):
//! - every function returns a status (`IHS_OK` = 0, negative = error) unless it returns a count, never aborts,
//!   and leaves a message for `ihs_last_error`;
//! - one handle is one HIP device and one stream: usable from one thread at a time;
//! - the caller owns host memory, the library owns device memory;
//! - particles cross the boundary as the reference's 48-byte `SphParticle` record (`ihs_particle` here, the same
//!   layout, so `bytemuck::cast_slice` converts between the two without a copy).
//!
//! The `wrapper` feature (on by default) adds [`GpuSphNative`], with `GpuSph`'s method names.
#![allow(non_camel_case_types, non_snake_case)]

use std::marker::{PhantomData, PhantomPinned};
use std::os::raw::{c_char, c_int, c_void};

#[cfg(feature = "wrapper")]
mod native;
#[cfg(feature = "wrapper")]
pub use native::{GpuSphNative, IhsError};

pub const IHS_ABI_VERSION: c_int = 1;

pub const IHS_OK: c_int = 0;
pub const IHS_ERR_ARG: c_int = -1;
pub const IHS_ERR_HIP: c_int = -2;
pub const IHS_ERR_CAPACITY: c_int = -3;
pub const IHS_ERR_STATE: c_int = -4;
pub const IHS_EAGAIN: c_int = -11;

pub const IHS_PREC_FP32: i32 = 0;
pub const IHS_PREC_SPLIT: i32 = 1;

pub const IHS_H_FIXED: i32 = 0;
pub const IHS_H_NEWTON: i32 = 1;
pub const IHS_H_LEARNED: i32 = 2;

pub const IHS_GRAVITY_DIRECT: i32 = 0;
pub const IHS_GRAVITY_TREE: i32 = 1;

pub const IHS_STAGE_SORT: i32 = 0;
pub const IHS_STAGE_GATHER: i32 = 1;
pub const IHS_STAGE_TREE: i32 = 2;
pub const IHS_STAGE_NBR: i32 = 3;
pub const IHS_STAGE_HSOLVE: i32 = 4;
pub const IHS_STAGE_DENSITY: i32 = 5;
pub const IHS_STAGE_FORCES: i32 = 6;
pub const IHS_STAGE_GRAVITY: i32 = 7;
pub const IHS_STAGE_KICK: i32 = 8;
pub const IHS_STAGE_COUNT: i32 = 9;

pub const IHS_FIELD_H: i32 = 0;
pub const IHS_FIELD_RHO: i32 = 1;
pub const IHS_FIELD_RPC: i32 = 2;
pub const IHS_FIELD_ASPH: i32 = 3;
pub const IHS_FIELD_AGRAV: i32 = 4;
pub const IHS_FIELD_HITER: i32 = 5;
pub const IHS_FIELD_NCOUNT: i32 = 6;
pub const IHS_FIELD_FEAT: i32 = 7;
pub const IHS_FIELD_POS64: i32 = 8;
pub const IHS_FIELD_HS: i32 = 9;

pub const IHS_MAX_SEGMENTS: i32 = 1024;
pub const IHS_MLP_NPARAM: i64 = 1313;
pub const IHS_MLP_IN: i64 = 6;
pub const IHS_MLP_SAMPLE: i64 = 8;
pub const IHS_MAX_BODIES: i32 = 64;

/// `SphParticle` (gpu_sph.rs:25-37): 48 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_particle {
    pub pos: [f32; 3],
    pub h: f32,
    pub vel: [f32; 3],
    pub u: f32,
    pub mass: f32,
    pub mat: u32,
    pub rho: f32,
    pub prov: u32,
}

/// `SphEos` (gpu_sph.rs:39-55): Tillotson rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta; 48 bytes.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_eos {
    pub rho0: f32,
    pub a: f32,
    pub b: f32,
    pub cap_a: f32,
    pub cap_b: f32,
    pub e0: f32,
    pub e_iv: f32,
    pub e_cv: f32,
    pub alpha: f32,
    pub beta: f32,
    pub _p0: f32,
    pub _p1: f32,
}

/// Handle configuration; `ihs_default_config` fills the reference's defaults.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_config {
    pub capacity: i32,
    pub leaf_size: i32,
    pub precision: i32,
    pub h_mode: i32,
    pub theta: f32,
    pub cfl: f32,
    pub n_ngb: f32,
    pub h_tol: f32,
    pub h_max_iter: i32,
    pub h_margin: f32,
    pub h_cap_factor: f32,
    pub nbr_capacity: i32,
    pub use_graph: i32,
    pub h_margin_seeded: f32,
    pub group_max: i32,
    pub quadrupole: i32,
    pub reuse_positions: i32,
    pub fused_walk: i32,
    pub theta_mono: f32,
    pub quad_tol: f32,
    pub theta_settled: f32,
    pub reserved: [i32; 3],
}

/// Opaque library handle.
#[repr(C)]
pub struct ihs_handle {
    _private: [u8; 0],
    _marker: PhantomData<(*mut u8, PhantomPinned)>,
}

/// FP64 energies, momentum and angular momentum, clock and solver counters (`ihs_stats_get`).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_stats {
    pub ke: f64,
    pub ie: f64,
    pub pe: f64,
    pub p: [f64; 3],
    pub L: [f64; 3],
    pub t: f64,
    pub dt: f64,
    pub steps: i64,
    pub force_evals: i64,
    pub mean_h_iters: f64,
    pub mean_h_evals: f64,
    pub mean_neighbours: f64,
    pub nbr_overflow: u32,
    pub walk_overflow: u32,
    pub h_capped: u32,
    pub h_unconverged: u32,
    pub nonfinite: u32,
    pub tie_runs: u32,
    pub grav_cell_interactions: i64,
    pub grav_pp_interactions: i64,
    pub nbr_candidates: i64,
    pub walk_nodes: i64,
}

/// One friends-of-friends clump (accretion.rs:29-41 `Clump`).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_clump {
    pub first: i64,
    pub n_members: i64,
    pub mass: f64,
    pub com_pos: [f64; 3],
    pub com_vel: [f64; 3],
    pub rho: f64,
    pub radius: f64,
    pub internal_ke: f64,
    pub self_pe: f64,
    pub earth_mass: f64,
    pub bound: i32,
    pub outside_roche: i32,
}

/// impact-run `Measure` (ie/tools/impact-run/src/main.rs:715-731) plus the values it is built from.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_measure {
    pub earth_frac: f64,
    pub disk_kg: f64,
    pub e_disk: f64,
    pub t_disk: f64,
    pub e_rem: f64,
    pub t_rem: f64,
    pub e_esc: f64,
    pub t_esc: f64,
    pub r_remnant: f64,
    pub clump_kg: f64,
    pub clump_earth_frac: f64,
    pub clump_n: i64,
    pub n_clumps: i64,
    pub n_bound: i64,
    pub m_total: f64,
    pub m_remnant: f64,
    pub com: [f64; 3],
    pub v_com: [f64; 3],
    pub link: f64,
    pub n_disk: i64,
}

/// `disk_stats_json` (gpu_sph.rs:399-469) as numbers, plus the same quantities in SI units.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct ihs_disk_summary {
    pub disk: f64,
    pub earth_pct: f64,
    pub remnant_km: f64,
    pub escaped: f64,
    pub moon: f64,
    pub e_disk: f64,
    pub t_disk: f64,
    pub esc: f64,
    pub r_remnant: f64,
    pub m_remnant: f64,
    pub biggest: f64,
    pub link: f64,
    pub n_disk: i64,
    pub n_clumps: i64,
}

// Layouts of the header's structs (checked against the C compiler's by the repository's tests).
const _: () = assert!(std::mem::size_of::<ihs_particle>() == 48);
const _: () = assert!(std::mem::size_of::<ihs_eos>() == 48);
const _: () = assert!(std::mem::size_of::<ihs_config>() == 96);
const _: () = assert!(std::mem::size_of::<ihs_stats>() == 184);
const _: () = assert!(std::mem::size_of::<ihs_clump>() == 120);
const _: () = assert!(std::mem::size_of::<ihs_measure>() == 192);
const _: () = assert!(std::mem::size_of::<ihs_disk_summary>() == 112);

#[link(name = "ihs_sph")]
extern "C" {
    pub fn ihs_default_config(cfg: *mut ihs_config);
    pub fn ihs_abi_version() -> c_int;
    pub fn ihs_create(device: c_int, cfg: *const ihs_config, out: *mut *mut ihs_handle) -> c_int;
    pub fn ihs_destroy(h: *mut ihs_handle) -> c_int;
    pub fn ihs_last_error(h: *const ihs_handle) -> *const c_char;
    pub fn ihs_upload(h: *mut ihs_handle, p: *const ihs_particle, n: i64, eos: *const ihs_eos, n_mat: i32,
                      softening: f32) -> c_int;
    pub fn ihs_upload_batch(h: *mut ihs_handle, p: *const ihs_particle, counts: *const i64, nseg: i32,
                            eos: *const ihs_eos, n_mat: i32, softening: *const f32) -> c_int;
    pub fn ihs_segment_clock(h: *mut ihs_handle, seg: i32, t: *mut f64, dt: *mut f64, steps: *mut i64) -> c_int;
    pub fn ihs_set_dt(h: *mut ihs_handle, dt: f64, damp: f32) -> c_int;
    pub fn ihs_set_av(h: *mut ihs_handle, alpha: f32, beta: f32) -> c_int;
    pub fn ihs_set_omega(h: *mut ihs_handle, omega: f32) -> c_int;
    pub fn ihs_step_kdk(h: *mut ihs_handle, substeps: i32) -> c_int;
    pub fn ihs_relax(h: *mut ihs_handle, steps: i32) -> c_int;
    pub fn ihs_force_eval(h: *mut ihs_handle) -> c_int;
    pub fn ihs_readback(h: *mut ihs_handle, out: *mut ihs_particle, cap: i64, blocking: i32) -> i64;
    pub fn ihs_begin_readback(h: *mut ihs_handle) -> c_int;
    pub fn ihs_take_readback(h: *mut ihs_handle, out: *mut ihs_particle, cap: i64) -> i64;
    pub fn ihs_count(h: *const ihs_handle) -> i64;
    pub fn ihs_sync(h: *mut ihs_handle) -> c_int;
    pub fn ihs_stats_get(h: *mut ihs_handle, out: *mut ihs_stats, reset: i32) -> c_int;
    pub fn ihs_segment_energy(h: *mut ihs_handle, seg: i32, out9: *mut f64) -> c_int;
    pub fn ihs_gravity(h: *mut ihs_handle, pos3: *const f64, mass: *const f64, n: i64, softening: f64, theta: f64,
                       mode: i32, acc3: *mut f64, pot: *mut f64) -> c_int;
    pub fn ihs_debug_tree(h: *mut ihs_handle, sorted_codes: *mut u32, order: *mut u32, ccodes: *mut u32,
                          left: *mut u32, right: *mut u32, parent: *mut u32) -> i64;
    pub fn ihs_debug_field(h: *mut ihs_handle, field: i32, out: *mut c_void) -> c_int;
    pub fn ihs_direct_f64(h: *mut ihs_handle, ids: *const u32, nt: i64, acc3: *mut f64) -> c_int;
    pub fn ihs_direct_f64_noalloc(h: *mut ihs_handle, ids: *const u32, nt: i64, acc3: *mut f64) -> c_int;
    pub fn ihs_profile(h: *mut ihs_handle, enable: i32) -> c_int;
    pub fn ihs_profile_read(h: *mut ihs_handle, ms: *mut f64, calls: *mut i64, reset: i32) -> c_int;
    pub fn ihs_mlp_set_weights(h: *mut ihs_handle, w: *const f32, nparam: i64) -> c_int;
    pub fn ihs_mlp_get_weights(h: *mut ihs_handle, w: *mut f32, nparam: i64) -> c_int;
    pub fn ihs_mlp_record(h: *mut ihs_handle, max_samples: i64) -> c_int;
    pub fn ihs_mlp_samples(h: *mut ihs_handle, out: *mut f32, cap: i64) -> i64;
    pub fn ihs_mlp_train(h: *mut ihs_handle, rows_host: *const f32, n_rows: i64, steps: i32, batch: i32, lr: f32,
                         seed: u64, loss: *mut f32) -> c_int;
    pub fn ihs_mlp_gradient(h: *mut ihs_handle, x: *const f32, y: *const f32, n: i64, grad: *mut f32) -> c_int;
    pub fn ihs_mlp_predict(h: *mut ihs_handle, x: *const f32, n: i64, y: *mut f32) -> c_int;
    pub fn ihs_measure_impact(h: *mut ihs_handle, segment: i32, n_earth: i64, out: *mut ihs_measure) -> c_int;
    pub fn ihs_disk_stats(h: *mut ihs_handle, segment: i32, out: *mut ihs_disk_summary) -> c_int;
    pub fn ihs_disk_clumps(h: *mut ihs_handle, segment: i32, require_bound: i32, frame: *mut f64,
                           clumps: *mut ihs_clump, cap: i64, n_clumps: *mut i64) -> c_int;
    pub fn ihs_body_bulk(h: *mut ihs_handle, segment: i32, prov: u32, out5: *mut f64) -> c_int;
    pub fn ihs_assemble(h: *mut ihs_handle, segment: i32, placements: *const f64, nbody: i32, softening: *mut f32,
                        dt: *mut f32) -> c_int;
    pub fn ihs_checkpoint_bytes(h: *mut ihs_handle) -> i64;
    pub fn ihs_checkpoint_save(h: *mut ihs_handle, buf: *mut c_void, cap: i64) -> i64;
    pub fn ihs_checkpoint_load(h: *mut ihs_handle, buf: *const c_void, bytes: i64) -> c_int;
    pub fn ihs_find_clumps(h: *mut ihs_handle, n: i64, pos3: *const f64, vel3: *const f64, mass: *const f64,
                           rho: *const f64, linking_length: f64, g: f64, softening: f64, central3: *const f64,
                           central_mass: f64, central_radius: f64, clumps: *mut ihs_clump, cap: i64,
                           n_clumps: *mut i64, label: *mut i64) -> c_int;
}
