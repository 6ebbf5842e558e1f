"""ctypes wrapper over the measurement half of liboracle_sph.so — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the product package never
does. measure_oracle.cpp restates the reference's Rust f64 code (file:line citations there); measure_oracle.h lists
what pins each function.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as _o

PARTICLE_DTYPE = np.dtype([
    ("pos", np.float32, 3), ("h", np.float32), ("vel", np.float32, 3), ("u", np.float32),
    ("mass", np.float32), ("mat", np.uint32), ("rho", np.float32), ("prov", np.uint32),
])


class OrMeasure(C.Structure):
    _fields_ = [
        ("earth_frac", C.c_double), ("disk_kg", C.c_double), ("e_disk", C.c_double), ("t_disk", C.c_double),
        ("e_rem", C.c_double), ("t_rem", C.c_double), ("e_esc", C.c_double), ("t_esc", C.c_double),
        ("r_remnant", C.c_double), ("clump_kg", C.c_double), ("clump_earth_frac", C.c_double),
        ("clump_n", C.c_int64), ("n_clumps", C.c_int64), ("n_bound", C.c_int64),
        ("m_total", C.c_double), ("m_remnant", C.c_double), ("com", C.c_double * 3), ("v_com", C.c_double * 3),
        ("link", C.c_double), ("n_disk", C.c_int64),
    ]


class OrClump(C.Structure):
    _fields_ = [
        ("first", C.c_int64), ("n_members", C.c_int64), ("mass", C.c_double), ("com_pos", C.c_double * 3),
        ("com_vel", C.c_double * 3), ("rho", C.c_double), ("radius", C.c_double), ("internal_ke", C.c_double),
        ("self_pe", C.c_double), ("earth_mass", C.c_double), ("bound", C.c_int32), ("outside_roche", C.c_int32),
    ]


_declared = False


def _lib():
    global _declared
    L = _o.lib()
    if not _declared:
        vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        L.or_sum_oi.argtypes = [vp, i64]
        L.or_sum_oi.restype = f64
        L.or_perigee.argtypes = [i32, vp, vp, f64, vp]
        L.or_perigee.restype = C.c_int
        L.or_measure_ir.argtypes = [vp, i64, i64, C.POINTER(OrMeasure), vp]
        L.or_measure_ir.restype = None
        L.or_engine_frame.argtypes = [vp, i64, i32, vp, vp, vp]
        L.or_engine_frame.restype = None
        L.or_find_clumps.argtypes = [i64, vp, vp, vp, vp, f64, f64, f64, vp, f64, f64, vp, i64, vp]
        L.or_find_clumps.restype = i64
        L.or_body_bulk.argtypes = [vp, i64, C.c_uint32, vp]
        L.or_body_bulk.restype = None
        L.or_assemble_n.argtypes = [vp, i64, vp, i32, vp, vp, vp]
        L.or_assemble_n.restype = i64
        _declared = True
    return L


def as_particles(p) -> np.ndarray:
    a = np.ascontiguousarray(p)
    if a.dtype.itemsize != 48:
        raise ValueError("expected 48-byte particle records")
    return a.view(PARTICLE_DTYPE)


def sum_oi(terms) -> float:
    t = np.ascontiguousarray(terms, np.float64)
    return _lib().or_sum_oi(t.ctypes.data, t.size)


def perigee(which: str, rel_p, rel_v, mu: float):
    """which = "impact_run" (main.rs:232-247) or "orbit" (orbit.rs:96-116); None for no periapsis."""
    p = np.ascontiguousarray(rel_p, np.float64)
    v = np.ascontiguousarray(rel_v, np.float64)
    out = C.c_double(0.0)
    ok = _lib().or_perigee(0 if which == "impact_run" else 1, p.ctypes.data, v.ctypes.data, float(mu),
                           C.byref(out))
    return out.value if ok else None


def measure_ir(parts, n_earth: int, with_classes: bool = False):
    P = as_particles(parts)
    m = OrMeasure()
    cls = np.zeros(max(1, P.shape[0]), np.int32)
    _lib().or_measure_ir(P.ctypes.data, P.shape[0], int(n_earth), C.byref(m),
                         cls.ctypes.data if with_classes else None)
    d = {f: getattr(m, f) for f, _ in OrMeasure._fields_}
    d["com"] = np.array(m.com[:])
    d["v_com"] = np.array(m.v_com[:])
    return (d, cls[:P.shape[0]]) if with_classes else d


def engine_frame(parts, require_bound: bool, with_classes: bool = False):
    P = as_particles(parts)
    frame = np.zeros(12, np.float64)
    et = np.zeros(3, np.float64)
    cls = np.zeros(max(1, P.shape[0]), np.int32)
    _lib().or_engine_frame(P.ctypes.data, P.shape[0], int(bool(require_bound)), frame.ctypes.data, et.ctypes.data,
                           cls.ctypes.data if with_classes else None)
    f = {"m_total": frame[0], "com": frame[1:4].copy(), "v_com": frame[4:7].copy(), "r_remnant": frame[7],
         "m_remnant": frame[8], "mu": frame[9], "disk_mass": frame[10], "link": frame[11], "e_disk": et[0],
         "t_disk": et[1], "esc": et[2]}
    return (f, cls[:P.shape[0]]) if with_classes else f


def find_clumps(pos, vel, mass, rho, linking_length, g, softening, central_pos, central_mass, central_radius):
    """(clumps as dicts ordered by lowest member, label per body)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
    rho = np.ascontiguousarray(rho, np.float64).reshape(-1)
    cen = np.ascontiguousarray(central_pos, np.float64).reshape(3)
    buf = (OrClump * max(1, n))()
    lab = np.zeros(max(1, n), np.int64)
    k = _lib().or_find_clumps(n, pos.ctypes.data, vel.ctypes.data, mass.ctypes.data, rho.ctypes.data,
                              float(linking_length), float(g), float(softening), cen.ctypes.data, float(central_mass),
                              float(central_radius), buf, max(1, n), lab.ctypes.data)
    out = []
    for i in range(k):
        c = buf[i]
        d = {f: getattr(c, f) for f, _ in OrClump._fields_}
        d["com_pos"] = np.array(c.com_pos[:])
        d["com_vel"] = np.array(c.com_vel[:])
        out.append(d)
    return out, lab[:n]


def disk_set(parts, require_bound: bool):
    """The disk bodies the engine functions hand to find_clumps (gpu_sph.rs:439-453): pos, vel, mass, rho floored
    at 1, in index order, with the frame."""
    P = as_particles(parts)
    f, cls = engine_frame(P, require_bound, with_classes=True)
    sel = cls == 1
    d = P[sel]
    pos = d["pos"].astype(np.float64)
    vel = d["vel"].astype(np.float64)
    mass = d["mass"].astype(np.float64)
    rho = np.fmax(d["rho"], np.float32(1.0)).astype(np.float64)  # f32::max(1.0): NaN gives 1
    return f, np.nonzero(sel)[0], pos, vel, mass, rho


def body_bulk(parts, prov: int):
    """(centre, mass, radius) of gpu_sph.rs:279-316."""
    P = as_particles(parts)
    out = np.zeros(5, np.float64)
    _lib().or_body_bulk(P.ctypes.data, P.shape[0], int(prov), out.ctypes.data)
    return out[:3].copy(), float(out[3]), float(out[4])


def assemble_n(parts, placements):
    """gpu_sph.rs:318-339. placements: (nbody, 9) offset, vel, spin. Returns (records, softening, dt)."""
    P = as_particles(parts)
    pl = np.ascontiguousarray(placements, np.float64).reshape(-1, 9)
    out = np.zeros(max(1, P.shape[0]), PARTICLE_DTYPE)
    soft = C.c_float(0.0)
    dt = C.c_float(0.0)
    k = _lib().or_assemble_n(P.ctypes.data, P.shape[0], pl.ctypes.data, pl.shape[0], out.ctypes.data, C.byref(soft),
                             C.byref(dt))
    return out[:k], soft.value, dt.value
