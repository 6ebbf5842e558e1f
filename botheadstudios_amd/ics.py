"""Two-body giant-impact initial conditions: seeded jittered cubic lattices cut to spheres.

The reference seeds bodies on Fibonacci-sphere shells (hydrostatic.rs:30-38,97-140); the impact
configurations this library is measured on ask instead for a seeded jittered cubic lattice with a
30% iron core and a granite mantle, so this builder is new. It reuses the reference's
deterministic mixer for the jitter (ie/tools/impact-run/src/main.rs:179-189, `hash_jitter`) and
its collision geometry conventions (start at 1.6 x contact, ie/crates/engine/src/terra/world_def.rs:344,
speed corrected for infall by energy conservation as ie/crates/engine/src/lib.rs:2814-2818 does).

Every particle has the same mass m = M_total / N. Material table order is [granite, iron].
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .materials import TILLOTSON, eos_table

G = 6.674e-11  # orbit.rs:18
M_EARTH = 5.972e24
ETA_48 = (48.0 * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)  # h = eta * (m/rho)^(1/3) -> 48 neighbours in r < h
MAT_GRANITE = 0
MAT_IRON = 1
MATERIALS = ("granite", "iron")

# The 48-byte interchange record (gpu_sph.rs:25-37 SphParticle).
SPH_PARTICLE_DTYPE = np.dtype([
    ("pos", np.float32, 3), ("h", np.float32), ("vel", np.float32, 3), ("u", np.float32),
    ("mass", np.float32), ("mat", np.uint32), ("rho", np.float32), ("prov", np.uint32),
])
assert SPH_PARTICLE_DTYPE.itemsize == 48


def hash_jitter(run: int, i: np.ndarray, axis: int) -> np.ndarray:
    """impact-run:179-189 splitmix64-style mixer -> values in [-1, 1]; vectorised over i."""
    with np.errstate(over="ignore"):
        i = np.asarray(i, dtype=np.uint64)
        x = (np.uint64(run) * np.uint64(0x9E3779B97F4A7C15)) ^ (i * np.uint64(0xD1B54A32D192ED03)) \
            ^ (np.uint64(axis) * np.uint64(0xCA5A826395121157)) ^ np.uint64(0x2545F4914F6CDD1D)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x.astype(np.float64) / float(np.iinfo(np.uint64).max) * 2.0 - 1.0


def _lattice_ball(spacing: float, radius: float) -> np.ndarray:
    """Cubic lattice points (i+1/2)*spacing within `radius` + 2 cells, sorted by radius (stable)."""
    k = int(math.ceil(radius / spacing)) + 2
    ax = (np.arange(-k, k, dtype=np.float64) + 0.5) * spacing
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    r = np.sqrt((p * p).sum(1))
    keep = r <= radius + 2.0 * spacing
    p, r = p[keep], r[keep]
    o = np.argsort(r, kind="stable")
    return p[o], r[o]


@dataclass
class Body:
    pos: np.ndarray   # (n,3) f64, centred on its own origin before placement
    mat: np.ndarray   # (n,) u32
    h: np.ndarray     # (n,) f64
    rho: np.ndarray   # (n,) f64
    radius: float
    mass: float


def make_body(n: int, m: float, core_frac: float = 0.3) -> Body:
    """n equal-mass particles: the n_core iron lattice points nearest the centre, then the granite
    lattice points of the shell outside the core sphere, nearest first."""
    rho_fe, rho_gr = TILLOTSON["iron"][0], TILLOTSON["granite"][0]
    n_core = int(round(n * core_frac))
    n_mantle = n - n_core
    d_fe = (m / rho_fe) ** (1.0 / 3.0)
    d_gr = (m / rho_gr) ** (1.0 / 3.0)
    r_core = (3.0 * n_core * d_fe ** 3 / (4.0 * math.pi)) ** (1.0 / 3.0)
    r_out = (3.0 * (n_core * d_fe ** 3 + n_mantle * d_gr ** 3) / (4.0 * math.pi)) ** (1.0 / 3.0)
    core, _ = _lattice_ball(d_fe, r_core)
    core = core[:n_core]
    mp, mr = _lattice_ball(d_gr, r_out)
    mp = mp[mr >= r_core][:n_mantle]
    if core.shape[0] != n_core or mp.shape[0] != n_mantle:
        raise RuntimeError("lattice too small for the requested particle count")
    pos = np.concatenate([core, mp])
    mat = np.concatenate([np.full(n_core, MAT_IRON, np.uint32), np.full(n_mantle, MAT_GRANITE, np.uint32)])
    h = np.concatenate([np.full(n_core, ETA_48 * d_fe), np.full(n_mantle, ETA_48 * d_gr)])
    rho = np.concatenate([np.full(n_core, rho_fe), np.full(n_mantle, rho_gr)])
    return Body(pos, mat, h, rho, r_out, n * m)


@dataclass
class ImpactIC:
    pos: np.ndarray
    vel: np.ndarray
    u: np.ndarray
    h: np.ndarray
    m: np.ndarray
    mat: np.ndarray
    rho: np.ndarray
    prov: np.ndarray
    eos: np.ndarray          # (n_mat, 10) f64, order MATERIALS
    softening: float
    meta: dict

    @property
    def n(self) -> int:
        return self.pos.shape[0]

    def particles48(self) -> np.ndarray:
        """The AoS 48-byte records the C ABI takes (f32 fields)."""
        p = np.zeros(self.n, SPH_PARTICLE_DTYPE)
        p["pos"] = self.pos
        p["h"] = self.h
        p["vel"] = self.vel
        p["u"] = self.u
        p["mass"] = self.m
        p["mat"] = self.mat
        p["rho"] = self.rho
        p["prov"] = self.prov
        return p


def impact_ic(n: int, m_target: float = 0.9, m_impactor: float = 0.1, angle_deg: float = 45.0,
              v_over_vesc: float = 1.1, seed: int = 0x1A7E0001, u0: float = 1.0e6,
              sep_contacts: float = 1.6, jitter: float = 0.25) -> ImpactIC:
    """Two-body impact in the centre-of-mass frame. Masses in Earth masses."""
    mt, mi = m_target * M_EARTH, m_impactor * M_EARTH
    n_t = int(round(n * mt / (mt + mi)))
    n_i = n - n_t
    m = (mt + mi) / n
    bt = make_body(n_t, m)
    bi = make_body(n_i, m)
    pos = np.concatenate([bt.pos, bi.pos])
    # Jitter U[-jitter*spacing, +jitter*spacing) per axis from the reference mixer.
    idx = np.arange(n, dtype=np.uint64)
    spacing = (m / np.where(np.concatenate([bt.mat, bi.mat]) == MAT_IRON, TILLOTSON["iron"][0],
                            TILLOTSON["granite"][0])) ** (1.0 / 3.0)
    for ax in range(3):
        pos[:, ax] += jitter * spacing * hash_jitter(seed, idx, ax)
    M_t, M_i = n_t * m, n_i * m
    M = M_t + M_i
    r_c = bt.radius + bi.radius
    v_esc = math.sqrt(2.0 * G * M / r_c)
    v_c = v_over_vesc * v_esc
    d0 = sep_contacts * r_c
    v0 = math.sqrt(v_c ** 2 - 2.0 * G * M * (1.0 / r_c - 1.0 / d0))
    ang = math.radians(angle_deg)
    ell = math.sin(ang) * r_c * v_c          # specific angular momentum at contact
    v_t = ell / d0
    v_r = math.sqrt(max(v0 ** 2 - v_t ** 2, 0.0))
    r_rel = np.array([-d0, 0.0, 0.0])         # impactor relative to target
    v_rel = np.array([v_r, v_t, 0.0])
    vel = np.zeros_like(pos)
    pos[:n_t] += -(M_i / M) * r_rel
    vel[:n_t] = -(M_i / M) * v_rel
    pos[n_t:] += (M_t / M) * r_rel
    vel[n_t:] = (M_t / M) * v_rel
    h = np.concatenate([bt.h, bi.h])
    mat = np.concatenate([bt.mat, bi.mat])
    rho = np.concatenate([bt.rho, bi.rho])
    prov = np.concatenate([np.zeros(n_t, np.uint32), np.ones(n_i, np.uint32)])
    soft = 0.25 * float(h.min())  # gpu_sph.rs:335: softening = min_h / 4
    meta = dict(n=n, n_target=n_t, n_impactor=n_i, particle_mass=m, r_target=bt.radius, r_impactor=bi.radius,
                v_esc=v_esc, v_contact=v_c, d0=d0, v0=v0, angle_deg=angle_deg, seed=seed,
                m_target=m_target, m_impactor=m_impactor, v_over_vesc=v_over_vesc)
    return ImpactIC(pos, vel, np.full(n, u0), h, np.full(n, m), mat, rho, prov, eos_table(MATERIALS), soft, meta)


# The five measured configurations (sizes and geometry); seeds are 0x1A7E0000 + k.
CONFIGS = {
    1: dict(n=16384, m_target=0.9, m_impactor=0.1, angle_deg=45.0, v_over_vesc=1.1, seed=0x1A7E0001),
    2: dict(n=262144, m_target=0.9, m_impactor=0.1, angle_deg=45.0, v_over_vesc=1.1, seed=0x1A7E0002),
    3: dict(n=1048576, m_target=0.9, m_impactor=0.15, angle_deg=30.0, v_over_vesc=1.2, seed=0x1A7E0003),
    4: dict(n=1048576, m_target=0.9, m_impactor=0.1, angle_deg=45.0, v_over_vesc=1.1, seed=0x1A7E0004),
}


def config_ic(k: int, n: int | None = None) -> ImpactIC:
    """Configuration k's impact (optionally at another particle count with the same geometry)."""
    c = dict(CONFIGS[k])
    if n is not None:
        c["n"] = n
    return impact_ic(**c)


def sweep_scenarios(count: int = 64, seed: int = 0x1A7E0005, n_lo: int = 4096, n_hi: int = 131072):
    """Batched-sweep parameters: N log-uniform in [n_lo, n_hi], impactor 0.05-0.2 Earth masses,
    angle 0-60 degrees, speed 1.0-1.5 v_esc (one hash_jitter stream per field)."""
    k = np.arange(count, dtype=np.uint64)
    u = [(hash_jitter(seed, k, ax) + 1.0) * 0.5 for ax in range(4)]
    n = np.exp(np.log(n_lo) + u[0] * (np.log(n_hi) - np.log(n_lo))).round().astype(np.int64)
    return [dict(n=int(n[i]), m_target=0.9, m_impactor=float(0.05 + 0.15 * u[1][i]),
                 angle_deg=float(60.0 * u[2][i]), v_over_vesc=float(1.0 + 0.5 * u[3][i]),
                 seed=int(seed * 1000 + i)) for i in range(count)]

# ---------------------------------------------------------------------------------------------
# Post-impact scenes for the disk / remnant measurement.


def cold_blob(center, bulk, radius, n, m_i, rho):
    """The accretion tests' blob (accretion.rs:199-213): n bodies on a Fibonacci spiral inside `radius`, all moving
    at `bulk`."""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    i = np.arange(n, dtype=np.float64)
    rr = radius * np.cbrt((i + 0.5) / n)
    y = 1.0 - 2.0 * (i + 0.5) / n
    rad = np.sqrt(np.maximum(1.0 - y * y, 0.0))
    th = golden * i
    p = np.asarray(center, np.float64) + np.stack([np.cos(th) * rad * rr, y * rr, np.sin(th) * rad * rr], axis=1)
    v = np.tile(np.asarray(bulk, np.float64), (n, 1))
    return p, v, np.full(n, m_i), np.full(n, rho)


def disk_scene(n_ball: int = 30000, n_disk: int = 1500, n_esc: int = 200, n_blob: int = 300, n_hot: int = 200,
               seed: int = 1, m: float = 1.0e20, R: float = 6.4e6, h: float = 3.2e5):
    """A synthetic post-impact snapshot with known orbit classes (the measurement's test and benchmark workload), as
    48-byte records (target = index < n_target, prov 0):
       - a remnant ball of n_ball bodies (uniform in radius R, near rest): remnant;
       - n_disk bodies on circular orbits at 3R..6R in a thin annulus: disk;
       - a cold blob of n_blob bodies at 4.5R on a circular orbit: disk, one bound clump outside Roche;
       - a hot blob of n_hot bodies at 5.5R (internal speeds +-1.5 km/s normal to the orbit): disk, one unbound
         clump;
       - n_esc bodies moving radially outward at 3 escape speeds from 2R..4R: escaped.
    The ball and the first half of the annulus are target material. Returns (records, n_target, classes)."""
    rng = np.random.default_rng(seed)
    M = n_ball * m
    mu = G * M

    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    ball = unit(n_ball) * (R * np.cbrt(rng.random(n_ball)))[:, None]
    vball = rng.normal(scale=5.0, size=(n_ball, 3))
    r = R * (3.0 + 3.0 * rng.random(n_disk))
    ph = 2 * math.pi * rng.random(n_disk)
    disk = np.stack([r * np.cos(ph), r * np.sin(ph), 0.02 * R * rng.normal(size=n_disk)], axis=1)
    vc = np.sqrt(mu / r)
    vdisk = np.stack([-vc * np.sin(ph), vc * np.cos(ph), np.zeros(n_disk)], axis=1)
    bp, bv, _, _ = cold_blob((4.5 * R, 0.0, 0.0), (0.0, math.sqrt(mu / (4.5 * R)), 0.0), 2.0e5, n_blob, m, 3000.0)
    hp, hv, _, _ = cold_blob((0.0, -5.5 * R, 0.0), (math.sqrt(mu / (5.5 * R)), 0.0, 0.0), 1.0e6, n_hot, m, 3000.0)
    hv[:, 2] = np.where(np.arange(n_hot) % 2 == 0, 1500.0, -1500.0)  # out of the plane: still a disk orbit
    re = R * (2.0 + 2.0 * rng.random(n_esc))
    ue = unit(n_esc)
    esc = ue * re[:, None]
    vesc = ue * (3.0 * np.sqrt(2.0 * mu / re))[:, None]
    pos = np.concatenate([ball, disk[: n_disk // 2], disk[n_disk // 2:], bp, hp, esc])
    vel = np.concatenate([vball, vdisk[: n_disk // 2], vdisk[n_disk // 2:], bv, hv, vesc])
    cls = np.concatenate([np.zeros(n_ball, np.int32), np.ones(n_disk + n_blob + n_hot, np.int32),
                          np.full(n_esc, 2, np.int32)])
    n = pos.shape[0]
    n_target = n_ball + n_disk // 2
    out = np.zeros(n, SPH_PARTICLE_DTYPE)
    out["pos"] = pos.astype(np.float32)
    out["vel"] = vel.astype(np.float32)
    out["h"] = np.float32(h)
    out["u"] = np.float32(1.0e5)
    out["mass"] = np.float32(m)
    out["rho"] = np.float32(3000.0)
    out["prov"] = (np.arange(n) >= n_target).astype(np.uint32)
    return out, n_target, cls
