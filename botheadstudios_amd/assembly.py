"""IC assembly around the GPU relaxation, with the relaxed state kept on the device.

Mirrors ie/crates/engine/src/gpu_sph.rs:198-393:
  - `far_apart_pair(target, impactor, separation_m)` (gpu_sph.rs:231-255) and `build_far_apart_n(bodies,
    separation)` (gpu_sph.rs:197-213): the bodies recentred on their own centres of mass, placed far apart along +x
    at rest, each tagged with its source index as `prov`; returns (particles, softening, relax dt);
  - `relax(sph, steps, dt)`: `GpuSph::encode_relax` (gpu_sph.rs:893-898) under the reference's relax discipline
    (artificial viscosity off while relaxing, gpu_sph.rs:246-249);
  - `body_bulk(sph, prov)` (gpu_sph.rs:279-316) on the device state;
  - `BodyPlacement`, `assemble_from_relaxed_n(sph, placements)` (gpu_sph.rs:318-339),
    `assemble_from_relaxed_at(...)` (gpu_sph.rs:353-370) and `assemble_from_relaxed_with(...)` (gpu_sph.rs:379-393):
    the relaxed bodies placed on the collision in place (no readback, no second upload); they return the
    reference's softening and shock-safe dt.

Bodies are `RelaxBody` records: 48-byte particles (any origin, at rest), the body's softening and the EOS table
its `mat` indexes. `from_ics_body` builds one from `ics.make_body`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .ics import SPH_PARTICLE_DTYPE
from .materials import tillotson_c2

G = 6.674e-11  # orbit.rs:18


@dataclass
class RelaxBody:
    """One body before relaxation (the role of HydroBody in gpu_sph.rs:231-255)."""
    particles: np.ndarray  # 48-byte records
    softening: float
    eos: np.ndarray        # (n_mat, 10) Tillotson rows, indexed by particles["mat"]

    def com(self) -> np.ndarray:
        """gpu_sph.rs:651-658: sequential mass-weighted mean position."""
        p = self.particles["pos"].astype(np.float64)
        m = self.particles["mass"].astype(np.float64)
        c = np.zeros(3)
        for i in range(3):
            c[i] = np.cumsum(p[:, i] * m)[-1] if len(m) else 0.0
        return c / np.cumsum(m)[-1]

    def relax_dt(self, cfl: float) -> float:
        """hydrostatic.rs:377-386: cfl * min h / max(c_max, 1), c_max over the body's materials at (rho0, u of
        the first particle)."""
        p = self.particles
        min_h = float(p["h"].astype(np.float64).min())
        u0 = float(p["u"][0]) if len(p) else 1.0
        c_max = 1.0
        for m in np.unique(p["mat"]):
            row = self.eos[int(m)]
            c_max = max(c_max, math.sqrt(tillotson_c2(row, float(row[0]), u0)))
        return cfl * min_h / max(c_max, 1.0)


def from_ics_body(body, particle_mass: float, u0: float, eos: np.ndarray, softening: float | None = None) -> RelaxBody:
    """A RelaxBody from `ics.make_body` (softening defaults to half the finest spacing, gpu_sph.rs:176-178)."""
    n = body.pos.shape[0]
    p = np.zeros(n, SPH_PARTICLE_DTYPE)
    p["pos"] = body.pos
    p["h"] = body.h
    p["u"] = u0
    p["mass"] = particle_mass
    p["mat"] = body.mat
    p["rho"] = body.rho
    if softening is None:
        softening = 0.5 * (particle_mass / float(np.max(body.rho))) ** (1.0 / 3.0)
    return RelaxBody(p, float(softening), np.asarray(eos, np.float64))


def _push(out: list, body: RelaxBody, prov: int, offset) -> None:
    """gpu_sph.rs:663-680 / 696-721: the body translated by `offset`, at rest, tagged `prov`."""
    q = body.particles.copy()
    pos = body.particles["pos"].astype(np.float64) + np.asarray(offset, np.float64)
    q["pos"] = pos.astype(np.float32)
    q["vel"] = 0.0
    q["prov"] = prov
    out.append(q)


def far_apart_pair(target: RelaxBody, impactor: RelaxBody, separation_m: float):
    """gpu_sph.rs:231-255: the target (prov 0) recentred at the origin, the impactor (prov 1) recentred at
    `separation_m` along +x, both at rest. Returns (particles, softening, relax dt at cfl 0.2)."""
    out: list = []
    _push(out, target, 0, -target.com())
    _push(out, impactor, 1, -impactor.com() + np.array([separation_m, 0.0, 0.0]))
    softening = np.float32(min(target.softening, impactor.softening))
    relax_dt = np.float32(min(target.relax_dt(0.2), impactor.relax_dt(0.2)))
    return np.concatenate(out), float(softening), float(relax_dt)


def build_far_apart_n(bodies: list, separation: float):
    """gpu_sph.rs:197-213 for already-built bodies: body k (prov k) at k * separation * 2 * r_max along +x, where
    r_max is the largest body radius (farthest particle from its centre of mass). Positions are offset, not
    recentred, as SphAssembly::add_body does. Returns (particles, softening, relax dt at cfl 0.2)."""
    def radius(b: RelaxBody) -> float:
        d = b.particles["pos"].astype(np.float64) - b.com()
        return float(np.sqrt((d * d).sum(1)).max())

    r_max = max(radius(b) for b in bodies)
    spacing = separation * 2.0 * r_max
    out: list = []
    for k, b in enumerate(bodies):
        _push(out, b, k, np.array([k * spacing, 0.0, 0.0]))
    softening = np.float32(min(b.softening for b in bodies))
    relax_dt = np.float32(min(b.relax_dt(0.2) for b in bodies))
    return np.concatenate(out), float(softening), float(relax_dt)


def relax(sph, steps: int, dt: float, damp: float = 0.995, av=(1.0, 2.0)) -> None:
    """`steps` damped relaxation steps (each: one force evaluation, then v = (v + a dt) damp, x += v dt) with
    artificial viscosity off; the viscosity is restored to `av` afterwards."""
    sph.set_av(0.0, 0.0)
    sph.set_dt(dt, damp)
    sph.encode_relax(int(steps))
    sph.set_av(*av)


def body_bulk(sph, prov: int, segment: int = 0):
    """gpu_sph.rs:279-316 on the device state: (centre, mass, radius enclosing 98% of the mass)."""
    out = np.zeros(5, np.float64)
    sph._check(sph._L.ihs_body_bulk(sph._h, int(segment), int(prov), out.ctypes.data))
    return out[:3].copy(), float(out[3]), float(out[4])


@dataclass
class BodyPlacement:
    """gpu_sph.rs:258-266: where body `prov` goes, its bulk velocity and its angular velocity (any axis)."""
    offset: tuple = (0.0, 0.0, 0.0)
    vel: tuple = (0.0, 0.0, 0.0)
    spin: tuple = (0.0, 0.0, 0.0)


def assemble_from_relaxed_n(sph, placements: list, segment: int = 0):
    """gpu_sph.rs:318-339 in place on the device state. Returns (softening, dt); the softening becomes the
    scenario's, the dt is the one to pass to set_dt."""
    pl = np.zeros((len(placements), 9), np.float64)
    for k, p in enumerate(placements):
        pl[k, 0:3] = p.offset
        pl[k, 3:6] = p.vel
        pl[k, 6:9] = p.spin
    soft = C.c_float(0.0)
    dt = C.c_float(0.0)
    sph._check(sph._L.ihs_assemble(sph._h, int(segment), pl.ctypes.data, len(placements), C.byref(soft),
                                    C.byref(dt)))
    return soft.value, dt.value


def assemble_from_relaxed_at(sph, target_spin: float, impactor_offset, impactor_vel, impactor_spin: float,
                             segment: int = 0):
    """gpu_sph.rs:353-370: the target (prov 0) at the origin spun about +z, the impactor (prov 1) at the given
    offset and velocity, spun about +z."""
    return assemble_from_relaxed_n(sph, [BodyPlacement((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, target_spin)),
                                         BodyPlacement(tuple(impactor_offset), tuple(impactor_vel),
                                                       (0.0, 0.0, impactor_spin))], segment)


def assemble_from_relaxed_with(sph, v_esc_multiple: float, start_separation: float, impact_parameter: float,
                               target_spin: float = 0.0, segment: int = 0):
    """gpu_sph.rs:379-393: the declared canonical impact from the relaxed bodies' measured bulk radii and masses:
    the impactor starts at start_separation x contact, aimed impact_parameter x r_target off axis, inbound at
    v_esc_multiple x the mutual escape speed at contact. Returns (softening, dt, geometry dict)."""
    _, m_e, r_e = body_bulk(sph, 0, segment)
    _, m_t, r_t = body_bulk(sph, 1, segment)
    contact = r_e + r_t
    v_esc = v_esc_multiple * math.sqrt(2.0 * G * (m_e + m_t) / contact)
    d0, b = start_separation * contact, impact_parameter * r_e
    soft, dt = assemble_from_relaxed_at(sph, target_spin, (d0, b, 0.0), (-v_esc, 0.0, 0.0), 0.0, segment)
    return soft, dt, dict(m_target=m_e, r_target=r_e, m_impactor=m_t, r_impactor=r_t, contact=contact, v_esc=v_esc,
                          d0=d0, b=b)
