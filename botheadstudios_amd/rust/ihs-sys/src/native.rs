//! `GpuSphNative`: the reference's `GpuSph` surface (ie/crates/engine/src/gpu_sph.rs:742-922) over libihs_sph, so
//! its callers (the `OrbitDemo` phase machine, lib.rs:2956-3179, and impact-run, main.rs:394-434) swap types. There
//! is no wgpu encoder: every `encode_*` call enqueues on the handle's HIP stream and returns at once, which is what
//! `encode_*` + `queue.submit` did. Errors come back as `IhsError` with the library's message; `.expect(..)` gives
//! the reference's panic-on-failure behaviour.

use std::ffi::CStr;
use std::fmt;
use std::os::raw::c_void;
use std::ptr;

use crate::*;

/// A failed call: the library's status code and its `ihs_last_error` message.
#[derive(Clone, Debug, PartialEq)]
pub struct IhsError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for IhsError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "libihs_sph error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for IhsError {}

/// One device handle: a HIP device, a stream, and the device-resident particle state.
pub struct GpuSphNative {
    h: *mut ihs_handle,
}

// A handle may move between threads; it is used from one thread at a time (&mut self for every call that runs work).
unsafe impl Send for GpuSphNative {}

impl GpuSphNative {
    /// `GpuSph::new` (gpu_sph.rs:767): the library's defaults with this capacity, on device 0.
    pub fn new(capacity: u32) -> Result<Self, IhsError> {
        let mut cfg = ihs_config::default();
        unsafe { ihs_default_config(&mut cfg) };
        cfg.capacity = capacity as i32;
        Self::with_config(0, &cfg)
    }

    /// A handle with an explicit configuration (walk setting, precision, smoothing-length mode).
    pub fn with_config(device: i32, cfg: &ihs_config) -> Result<Self, IhsError> {
        let mut h: *mut ihs_handle = ptr::null_mut();
        let rc = unsafe { ihs_create(device, cfg, &mut h) };
        if rc != IHS_OK || h.is_null() {
            // Without a handle there is no message to ask for: name the failure from the status.
            let message = match rc {
                IHS_ERR_ARG => "ihs_create: invalid configuration or device index",
                IHS_ERR_HIP => "ihs_create: no HIP device, or a device allocation failed",
                _ => "ihs_create failed",
            };
            return Err(IhsError { code: rc, message: message.to_string() });
        }
        Ok(Self { h })
    }

    fn check(&self, rc: i64) -> Result<i64, IhsError> {
        if rc >= 0 {
            return Ok(rc);
        }
        let message = unsafe { CStr::from_ptr(ihs_last_error(self.h)) }.to_string_lossy().into_owned();
        Err(IhsError { code: rc as i32, message })
    }

    /// `GpuSph::upload` (gpu_sph.rs:839-850). More particles than the capacity is an error, not a clamp.
    pub fn upload(&mut self, particles: &[ihs_particle], eos: &[ihs_eos], softening: f32) -> Result<(), IhsError> {
        let rc = unsafe {
            ihs_upload(self.h, particles.as_ptr(), particles.len() as i64, eos.as_ptr(), eos.len() as i32, softening)
        };
        self.check(rc as i64).map(|_| ())
    }

    /// Independent scenarios in one handle (impact-run's ensemble loop, main.rs:645-671): `counts[s]` particles
    /// each, concatenated in `particles`, with their own softening.
    pub fn upload_batch(&mut self, particles: &[ihs_particle], counts: &[i64], eos: &[ihs_eos],
                        softening: &[f32]) -> Result<(), IhsError> {
        assert_eq!(counts.len(), softening.len(), "one softening per scenario");
        assert_eq!(counts.iter().sum::<i64>(), particles.len() as i64, "counts must cover the particles");
        let rc = unsafe {
            ihs_upload_batch(self.h, particles.as_ptr(), counts.as_ptr(), counts.len() as i32, eos.as_ptr(),
                             eos.len() as i32, softening.as_ptr())
        };
        self.check(rc as i64).map(|_| ())
    }

    /// `GpuSph::set_dt` (gpu_sph.rs:853-857); dt <= 0 selects the device-side Courant dt (no read-back).
    pub fn set_dt(&mut self, dt: f32, damp: f32) -> Result<(), IhsError> {
        self.check(unsafe { ihs_set_dt(self.h, dt as f64, damp) } as i64).map(|_| ())
    }

    /// `GpuSph::set_av` (gpu_sph.rs:864-868).
    pub fn set_av(&mut self, alpha: f32, beta: f32) -> Result<(), IhsError> {
        self.check(unsafe { ihs_set_av(self.h, alpha, beta) } as i64).map(|_| ())
    }

    /// Rotating-frame relaxation (`SphParams::omega`, gpu_sph.rs:68-76).
    pub fn set_omega(&mut self, omega: f32) -> Result<(), IhsError> {
        self.check(unsafe { ihs_set_omega(self.h, omega) } as i64).map(|_| ())
    }

    /// `GpuSph::encode_relax` (gpu_sph.rs:893-898): enqueued, returns at once.
    pub fn encode_relax(&mut self, steps: u32) -> Result<(), IhsError> {
        self.check(unsafe { ihs_relax(self.h, steps as i32) } as i64).map(|_| ())
    }

    /// `GpuSph::encode_kdk` (gpu_sph.rs:901-908): enqueued, returns at once.
    pub fn encode_kdk(&mut self, substeps: u32) -> Result<(), IhsError> {
        self.check(unsafe { ihs_step_kdk(self.h, substeps as i32) } as i64).map(|_| ())
    }

    /// `GpuSph::count` (gpu_sph.rs:870).
    pub fn count(&self) -> u32 {
        unsafe { ihs_count(self.h) }.max(0) as u32
    }

    /// `GpuSph::begin_readback` (gpu_store.rs:126-145): a snapshot of the state behind the queued work, copied
    /// into pinned host memory without waiting. A no-op while a snapshot is in flight.
    pub fn begin_readback(&mut self) -> Result<(), IhsError> {
        self.check(unsafe { ihs_begin_readback(self.h) } as i64).map(|_| ())
    }

    /// `GpuSph::take_readback` (gpu_store.rs:147-152): the snapshot of the last `begin_readback` once copied,
    /// however many steps were queued after it; `None` while pending or when none was begun.
    pub fn take_readback(&mut self) -> Result<Option<Vec<ihs_particle>>, IhsError> {
        let mut out = vec![ihs_particle::default(); self.count() as usize];
        let rc = unsafe { ihs_take_readback(self.h, out.as_mut_ptr(), out.len() as i64) };
        if rc == IHS_EAGAIN as i64 {
            return Ok(None);
        }
        let n = self.check(rc)? as usize;
        out.truncate(n);
        Ok(Some(out))
    }

    /// The current state in upload order, after the queued work (blocking).
    pub fn readback(&mut self) -> Result<Vec<ihs_particle>, IhsError> {
        let mut out = vec![ihs_particle::default(); self.count() as usize];
        let n = self.check(unsafe { ihs_readback(self.h, out.as_mut_ptr(), out.len() as i64, 1) })? as usize;
        out.truncate(n);
        Ok(out)
    }

    /// Waits for the queued work.
    pub fn sync(&mut self) -> Result<(), IhsError> {
        self.check(unsafe { ihs_sync(self.h) } as i64).map(|_| ())
    }

    /// FP64 energies, momentum, angular momentum and solver counters (totals over a batch).
    pub fn stats(&mut self, reset: bool) -> Result<ihs_stats, IhsError> {
        let mut s = ihs_stats::default();
        self.check(unsafe { ihs_stats_get(self.h, &mut s, reset as i32) } as i64)?;
        Ok(s)
    }

    /// One scenario's simulated time, last dt and step count.
    pub fn segment_clock(&mut self, seg: i32) -> Result<(f64, f64, i64), IhsError> {
        let (mut t, mut dt, mut steps) = (0.0, 0.0, 0i64);
        self.check(unsafe { ihs_segment_clock(self.h, seg, &mut t, &mut dt, &mut steps) } as i64)?;
        Ok((t, dt, steps))
    }

    /// `total_energy` (gpu_sph.rs:624-649) of one scenario: KE, IE, PE, p xyz, L xyz.
    pub fn segment_energy(&mut self, seg: i32) -> Result<[f64; 9], IhsError> {
        let mut e = [0.0f64; 9];
        self.check(unsafe { ihs_segment_energy(self.h, seg, e.as_mut_ptr()) } as i64)?;
        Ok(e)
    }

    /// `GravityField::accelerations` (gpu_gravity.rs:556-578) on host bodies: `IHS_GRAVITY_TREE` or `_DIRECT`.
    pub fn gravity(&mut self, pos3: &[f64], mass: &[f64], softening: f64, theta: f64,
                   mode: i32) -> Result<Vec<f64>, IhsError> {
        assert_eq!(pos3.len(), 3 * mass.len(), "pos3 holds three coordinates per body");
        let mut acc = vec![0.0f64; pos3.len()];
        let rc = unsafe {
            ihs_gravity(self.h, pos3.as_ptr(), mass.as_ptr(), mass.len() as i64, softening, theta, mode,
                        acc.as_mut_ptr(), ptr::null_mut())
        };
        self.check(rc as i64)?;
        Ok(acc)
    }

    /// Sets the learned smoothing-length seed's weights (`IHS_MLP_NPARAM` floats) and enables it.
    pub fn set_mlp_weights(&mut self, w: &[f32]) -> Result<(), IhsError> {
        self.check(unsafe { ihs_mlp_set_weights(self.h, w.as_ptr(), w.len() as i64) } as i64).map(|_| ())
    }

    /// impact-run `measure(body, n_earth)` + `moon_candidate` (main.rs:732-878) on the device state.
    pub fn measure(&mut self, segment: i32, n_earth: usize) -> Result<ihs_measure, IhsError> {
        let mut m = ihs_measure::default();
        self.check(unsafe { ihs_measure_impact(self.h, segment, n_earth as i64, &mut m) } as i64)?;
        Ok(m)
    }

    /// `disk_stats_json` (gpu_sph.rs:399-469) from the device state: the same JSON text.
    pub fn disk_stats_json(&mut self, segment: i32) -> Result<String, IhsError> {
        let mut d = ihs_disk_summary::default();
        self.check(unsafe { ihs_disk_stats(self.h, segment, &mut d) } as i64)?;
        Ok(format!("{{\"disk\":{:.3},\"earth_pct\":{:.0},\"remnant_km\":{:.0},\"escaped\":{:.3},\"moon\":{:.3}}}",
                   d.disk, d.earth_pct, d.remnant_km, d.escaped, d.moon))
    }

    /// `body_bulk` (gpu_sph.rs:279-316) of the particles with this provenance: centre xyz, mass, radius.
    pub fn body_bulk(&mut self, segment: i32, prov: u32) -> Result<[f64; 5], IhsError> {
        let mut out = [0.0f64; 5];
        self.check(unsafe { ihs_body_bulk(self.h, segment, prov, out.as_mut_ptr()) } as i64)?;
        Ok(out)
    }

    /// `assemble_from_relaxed_n` (gpu_sph.rs:318-339) in place; `placements` holds 9 doubles per body (offset,
    /// velocity, spin). Returns the softening and the shock-safe dt.
    pub fn assemble(&mut self, segment: i32, placements: &[f64]) -> Result<(f32, f32), IhsError> {
        assert_eq!(placements.len() % 9, 0, "nine doubles per body");
        let (mut soft, mut dt) = (0f32, 0f32);
        let rc = unsafe {
            ihs_assemble(self.h, segment, placements.as_ptr(), (placements.len() / 9) as i32, &mut soft, &mut dt)
        };
        self.check(rc as i64)?;
        Ok((soft, dt))
    }

    /// A checkpoint image of the device state (resume with `checkpoint_load` on a handle created from the
    /// `ihs_config` stored at byte 40 of the image).
    pub fn checkpoint_save(&mut self) -> Result<Vec<u8>, IhsError> {
        let bytes = self.check(unsafe { ihs_checkpoint_bytes(self.h) })? as usize;
        let mut buf = vec![0u8; bytes];
        let n = self.check(unsafe { ihs_checkpoint_save(self.h, buf.as_mut_ptr() as *mut c_void, bytes as i64) })?;
        buf.truncate(n as usize);
        Ok(buf)
    }

    pub fn checkpoint_load(&mut self, image: &[u8]) -> Result<(), IhsError> {
        let rc = unsafe { ihs_checkpoint_load(self.h, image.as_ptr() as *const c_void, image.len() as i64) };
        self.check(rc as i64).map(|_| ())
    }

    /// The raw handle, for calls this wrapper does not cover.
    pub fn raw(&mut self) -> *mut ihs_handle {
        self.h
    }
}

impl Drop for GpuSphNative {
    fn drop(&mut self) {
        unsafe {
            ihs_destroy(self.h);
        }
    }
}
