"""Disk / remnant measurement and accretion clumps over the C ABI, on the device-resident state.

Mirrors, with the reference's names and return shapes:
  - `measure(sph, n_earth)` -> `Measure`: ie/tools/impact-run/src/main.rs:712-778 (`Measure`, `measure`) with the
    Moon candidate of `moon_candidate` (main.rs:802-878);
  - `disk_stats_json(sph)`: ie/crates/engine/src/gpu_sph.rs:399-469, the same JSON text;
  - `moonlet_bodies`, `largest_moonlet_orbit`, `disk_moonlets`: gpu_sph.rs:478-640;
  - `find_clumps`, `accrete`, `Clump`, `Body`, `Accreted`: ie/crates/engine/src/accretion.rs:29-188.

The measurement runs on the particles as they sit in HBM after the last step (no readback). It differs from the
reference in two deliberate ways: sums are double-double in a fixed order (so results do not depend on how the
particles are stored; the reference's sequential sums round differently by at most their own rounding error), and
among clumps of equal mass the one with the lowest member index wins (the reference leaves that to hash-map order).
Members are reported as a label per body rather than as lists.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

G = 6.674e-11      # orbit.rs:18
M_MOON = 7.342e22  # main.rs:713, gpu_sph.rs:401


@dataclass
class Measure:
    """main.rs:715-731, plus the intermediate values (m_total, m_remnant, com, v_com, link, n_disk)."""
    earth_frac: float
    disk_kg: float
    e_disk: float
    t_disk: float
    e_rem: float
    t_rem: float
    e_esc: float
    t_esc: float
    r_remnant: float
    clump_kg: float
    clump_earth_frac: float
    clump_n: int
    n_clumps: int
    n_bound: int
    m_total: float = 0.0
    m_remnant: float = 0.0
    com: tuple = (0.0, 0.0, 0.0)
    v_com: tuple = (0.0, 0.0, 0.0)
    link: float = 0.0
    n_disk: int = 0


@dataclass
class Clump:
    """accretion.rs:29-41. `members` holds the member indices (into the bodies passed to find_clumps, or into the
    scenario's particles for the disk functions) when labels were requested; `first` is the lowest of them."""
    first: int
    n_members: int
    mass: float
    com_pos: np.ndarray
    com_vel: np.ndarray
    rho: float
    radius: float
    internal_ke: float
    self_pe: float
    bound: bool
    outside_roche: bool
    earth_mass: float = 0.0
    members: np.ndarray | None = None

    def accretes(self) -> bool:
        """accretion.rs:44-47: more than one member, bound, outside Roche."""
        return self.n_members >= 2 and self.bound and self.outside_roche


@dataclass
class Body:
    """accretion.rs:50-57."""
    pos: np.ndarray
    vel: np.ndarray
    mass: float
    rho: float
    radius: float


@dataclass
class Accreted:
    """accretion.rs:60-64."""
    bodies: list = field(default_factory=list)
    consumed: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))


def _clump(c: _lib.IhsClump) -> Clump:
    return Clump(first=int(c.first), n_members=int(c.n_members), mass=float(c.mass),
                 com_pos=np.array(c.com_pos[:], np.float64), com_vel=np.array(c.com_vel[:], np.float64),
                 rho=float(c.rho), radius=float(c.radius), internal_ke=float(c.internal_ke),
                 self_pe=float(c.self_pe), bound=bool(c.bound), outside_roche=bool(c.outside_roche),
                 earth_mass=float(c.earth_mass))


def measure(sph, n_earth: int, segment: int = 0) -> Measure:
    """impact-run `measure(body, n_earth)`: target particles are those with upload index < n_earth."""
    m = _lib.IhsMeasure()
    sph._check(sph._L.ihs_measure_impact(sph._h, int(segment), int(n_earth), C.byref(m)))
    return Measure(earth_frac=m.earth_frac, disk_kg=m.disk_kg, e_disk=m.e_disk, t_disk=m.t_disk, e_rem=m.e_rem,
                   t_rem=m.t_rem, e_esc=m.e_esc, t_esc=m.t_esc, r_remnant=m.r_remnant, clump_kg=m.clump_kg,
                   clump_earth_frac=m.clump_earth_frac, clump_n=int(m.clump_n), n_clumps=int(m.n_clumps),
                   n_bound=int(m.n_bound), m_total=m.m_total, m_remnant=m.m_remnant, com=tuple(m.com[:]),
                   v_com=tuple(m.v_com[:]), link=m.link, n_disk=int(m.n_disk))


def print_measure(m: Measure, m_earth: float, m_theia: float, v_esc: float) -> str:
    """The report of main.rs:780-793, returned as text."""
    lines = [
        f"\n=== DEFORMABLE-EARTH IMPACT (M_e={m_earth:.2e}, M_t={m_theia:.2e}, v={v_esc:.0f} m/s, "
        f"R_remnant={m.r_remnant / 1e3:.0f} km) ===",
        f"  ORBITING DISK (perigee > remnant): Earth {m.e_disk:.3e} | Theia {m.t_disk:.3e} kg = "
        f"{m.disk_kg / M_MOON:.3f} M_moon  → {m.earth_frac:.0f}% EARTH",
        f"  remnant: Earth {m.e_rem:.3e} | Theia {m.t_rem:.3e} kg · escaped: Earth {m.e_esc:.3e} | "
        f"Theia {m.t_esc:.3e} kg",
        f"  → Earth material {'DID' if m.e_disk > 0.0 else 'did NOT'} reach orbit",
        f"  ACCRETION: {m.n_clumps} disk clumps, {m.n_bound} self-bound",
    ]
    if m.clump_kg > 0.0:
        lines.append(f"  MOON CANDIDATE: largest bound+outside-Roche clump = {m.clump_kg:.3e} kg = "
                     f"{m.clump_kg / M_MOON:.3f} M_moon ({m.clump_n} particles) → {m.clump_earth_frac:.0f}% EARTH")
    else:
        lines.append("  MOON CANDIDATE: none yet (no bound clump outside Roche — disk still dispersed; needs more "
                     "time / N)")
    return "\n".join(lines)


def disk_stats(sph, segment: int = 0) -> _lib.IhsDiskSummary:
    d = _lib.IhsDiskSummary()
    sph._check(sph._L.ihs_disk_stats(sph._h, int(segment), C.byref(d)))
    return d


def disk_stats_json(sph, segment: int = 0) -> str:
    """gpu_sph.rs:399-469: {"disk","earth_pct","remnant_km","escaped","moon"} with the reference's precisions."""
    if sph.count() == 0:
        return "null"
    d = disk_stats(sph, segment)
    return (f'{{"disk":{d.disk:.3f},"earth_pct":{d.earth_pct:.0f},"remnant_km":{d.remnant_km:.0f},'
            f'"escaped":{d.escaped:.3f},"moon":{d.moon:.3f}}}')


def disk_clumps(sph, require_bound: bool = False, segment: int = 0):
    """(frame dict, clumps) of the scenario's orbiting disk, as gpu_sph.rs:478-640 select and measure it."""
    frame = np.zeros(13, np.float64)
    n = C.c_int64(0)
    cap = max(1, int(sph.count()))
    buf = (_lib.IhsClump * cap)()
    sph._check(sph._L.ihs_disk_clumps(sph._h, int(segment), int(bool(require_bound)), frame.ctypes.data, buf, cap,
                                       C.byref(n)))
    keys = ("m_total", "com", "v_com", "r_remnant", "m_remnant", "mu", "disk_mass", "link", "n_disk")
    f = dict(zip(keys, (frame[0], frame[1:4].copy(), frame[4:7].copy(), frame[7], frame[8], frame[9], frame[10],
                        frame[11], int(frame[12]))))
    return f, [_clump(buf[i]) for i in range(min(n.value, cap))]


def moonlet_bodies(sph, segment: int = 0):
    """gpu_sph.rs:478-512: (Earth-relative-frame COM position, radius, mass) of accreting clumps of >= 3 members."""
    if sph.count() < 2:
        return []
    _, clumps = disk_clumps(sph, False, segment)
    return [(c.com_pos, c.radius, c.mass) for c in clumps if c.accretes() and c.n_members >= 3]


def largest_moonlet_orbit(sph, segment: int = 0):
    """gpu_sph.rs:517-564: (r, v, energy, a, mass, ecc, theta) of the largest accreting clump's orbit about the
    remnant, or None."""
    if sph.count() < 2:
        return None
    f, clumps = disk_clumps(sph, False, segment)
    cand = [c for c in clumps if c.accretes() and c.n_members >= 3]
    if not cand:
        return None
    big = max(cand, key=lambda c: c.mass)
    mu = f["mu"]
    rel_p = big.com_pos - f["com"]
    rel_v = big.com_vel - f["v_com"]
    r = math.sqrt(rel_p[0] * rel_p[0] + rel_p[1] * rel_p[1] + rel_p[2] * rel_p[2])
    v = math.sqrt(rel_v[0] * rel_v[0] + rel_v[1] * rel_v[1] + rel_v[2] * rel_v[2])
    energy = 0.5 * v * v - mu / max(r, 1.0)
    a = -mu / (2.0 * energy) if energy < 0.0 else math.inf
    hv = (rel_p[1] * rel_v[2] - rel_v[1] * rel_p[2], rel_p[2] * rel_v[0] - rel_v[2] * rel_p[0],
          rel_p[0] * rel_v[1] - rel_v[0] * rel_p[1])
    h = math.sqrt(hv[0] * hv[0] + hv[1] * hv[1] + hv[2] * hv[2])
    ecc = math.sqrt(max(1.0 + 2.0 * energy * h * h / (mu * mu), 0.0))
    theta = math.atan2(rel_p[1], rel_p[0])
    return (r, v, energy, a, big.mass, ecc, theta)


def disk_moonlets(sph, earth_radius: float, segment: int = 0):
    """gpu_sph.rs:566-623: [(a, mass)] moonlets at 3.8 Earth radii: one per accreting clump, else the whole bound-
    or-passing disk as one, else none."""
    if sph.count() < 2:
        return []
    f, clumps = disk_clumps(sph, False, segment)
    if f["n_disk"] < 2:
        return []
    moonlets = [(3.8 * earth_radius, c.mass) for c in clumps if c.accretes()]
    if moonlets:
        return moonlets
    return [(3.8 * earth_radius, f["disk_mass"])] if f["disk_mass"] > 0.0 else []


def find_clumps(sph, pos, vel, mass, rho, linking_length: float, g: float, softening: float, central_pos,
                central_mass: float, central_radius: float, with_members: bool = True):
    """accretion.rs:73-155 on host bodies, computed on the handle's device. Raises on ragged inputs (the
    reference asserts)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
    rho = np.ascontiguousarray(rho, np.float64).reshape(-1)
    if vel.shape[0] != n or mass.shape[0] != n or rho.shape[0] != n:
        raise ValueError("accretion: ragged particle arrays")
    cen = np.ascontiguousarray(central_pos, np.float64).reshape(3)
    cap = max(1, n)
    buf = (_lib.IhsClump * cap)()
    nc = C.c_int64(0)
    lab = np.zeros(max(1, n), np.int64)
    sph._check(sph._L.ihs_find_clumps(sph._h, n, pos.ctypes.data, vel.ctypes.data, mass.ctypes.data,
                                       rho.ctypes.data, float(linking_length), float(g), float(softening),
                                       cen.ctypes.data, float(central_mass), float(central_radius), buf, cap,
                                       C.byref(nc), lab.ctypes.data if with_members else None))
    clumps = [_clump(buf[i]) for i in range(nc.value)]
    if with_members and n:
        order = np.argsort(lab[:n], kind="stable")
        sl = lab[:n][order]
        starts = np.searchsorted(sl, np.arange(nc.value))
        ends = np.searchsorted(sl, np.arange(nc.value), side="right")
        for i, c in enumerate(clumps):
            c.members = order[starts[i]:ends[i]]
    return clumps


def accrete(sph, pos, vel, mass, rho, linking_length: float, g: float, softening: float, central_pos,
            central_mass: float, central_radius: float) -> Accreted:
    """accretion.rs:165-188: every accreting clump becomes one body at its centre of mass; consumed is sorted."""
    clumps = find_clumps(sph, pos, vel, mass, rho, linking_length, g, softening, central_pos, central_mass,
                         central_radius, with_members=True)
    out = Accreted()
    consumed = []
    for c in clumps:
        if c.accretes():
            out.bodies.append(Body(pos=c.com_pos, vel=c.com_vel, mass=c.mass, rho=c.rho, radius=c.radius))
            consumed.append(c.members)
    out.consumed = np.sort(np.concatenate(consumed)) if consumed else np.zeros(0, np.int64)
    return out
