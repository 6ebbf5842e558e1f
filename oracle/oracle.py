"""ctypes wrapper over liboracle_sph.so — the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product package never does. Every function restates the reference's CPU arithmetic; the C++
source (sph_oracle.cpp) cites the reference file:line for each.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sph.so")
_lib = None

G = 6.674e-11  # orbit.rs:18


def build() -> str:
    """Compile the oracle with its Makefile (g++, no reference sources involved)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_native = None


def native_lib():
    """The same oracle compiled with -march=native for this host (built on first use; the CPU baseline's timing
    library). Falls back to the portable build when the compiler cannot target this host."""
    global _native
    if _native is None:
        path = os.path.join(_HERE, "liboracle_sph_native.so")
        try:
            subprocess.run(["make", "-s", "-C", _HERE, "liboracle_sph_native.so"], check=True, capture_output=True)
            _native = C.CDLL(path)
            _declare(_native)
        except (OSError, subprocess.CalledProcessError):
            _native = lib()
    return _native


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OrState(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("xyz", C.c_void_p),
        ("vel", C.c_void_p),
        ("u", C.c_void_p),
        ("h", C.c_void_p),
        ("m", C.c_void_p),
        ("mat", C.c_void_p),
        ("eos", C.c_void_p),
        ("n_mat", C.c_int32),
        ("softening", C.c_double),
        ("av_alpha", C.c_double),
        ("av_beta", C.c_double),
        ("theta", C.c_double),
        ("gravity_mode", C.c_int32),
        ("cs_fd", C.c_int32),
        ("brute", C.c_int32),
        ("leaf_cap", C.c_int32),
        ("nthreads", C.c_int32),
        ("h_solve", C.c_int32),
        ("h_pad", C.c_int32),
        ("h_eta", C.c_double),
        ("h_tol", C.c_double),
    ]


def _declare(L):
    d, i64, i32, f32 = C.c_double, C.c_int64, C.c_int32, C.c_float
    L.or_pressure.argtypes = [_dp, d, d]
    L.or_pressure.restype = d
    L.or_sound_speed_sq.argtypes = [_dp, d, d]
    L.or_sound_speed_sq.restype = d
    L.or_sound_speed_wgsl.argtypes = [_dp, d, d]
    L.or_sound_speed_wgsl.restype = d
    L.or_pressure_f32.argtypes = [_fp, f32, f32]
    L.or_pressure_f32.restype = f32
    L.or_sph_w.argtypes = [d, d]
    L.or_sph_w.restype = d
    L.or_sph_dw.argtypes = [d, d]
    L.or_sph_dw.restype = d
    L.or_expand_bits.argtypes = [C.c_uint32]
    L.or_expand_bits.restype = C.c_uint32
    L.or_bbox_f32.argtypes = [_fp, i64, _fp, _fp]
    L.or_morton_f32.argtypes = [_fp, i64, _fp, _fp, _up]
    L.or_sort_pairs.argtypes = [_up, i64, _up, _up]
    L.or_cluster_codes.argtypes = [_up, i64, i64, _up]
    L.or_karras.argtypes = [_up, i64, _up, _up, _up]
    L.or_karras_ex.argtypes = [_up, i64, i32, _up, _up, _up]
    L.or_upsweep_f32.argtypes = [_fp, i64, i64, _up, _up, i64, _fp, _fp, _fp]
    L.or_bh_walk_f32.argtypes = [_fp, _up, i64, i64, _up, _up, _fp, _fp, _fp, i64, f32, f32, _fp]
    L.or_bh_walk_count.argtypes = [_fp, i64, i64, _up, _up, _fp, _fp, _fp, i64, f32, f32, _lp, i64, _lp, C.c_void_p]
    L.or_direct_f64.argtypes = [_dp, _dp, i64, d, C.c_void_p, i64, _dp, C.c_void_p, C.c_int]
    L.or_octree_f64.argtypes = [_dp, _dp, i64, d, d, C.c_int, _dp]
    L.or_grid_pairs.argtypes = [_dp, i64, d, d, C.c_void_p, i64]
    L.or_grid_pairs.restype = i64
    L.or_density.argtypes = [C.POINTER(OrState), _dp]
    L.or_forces.argtypes = [C.POINTER(OrState), _dp, _dp, _dp]
    L.or_courant_dt.argtypes = [C.POINTER(OrState), _dp, d]
    L.or_courant_dt.restype = d
    L.or_kdk.argtypes = [C.POINTER(OrState), d, d, i32]
    L.or_kdk.restype = d
    L.or_relax.argtypes = [C.POINTER(OrState), d, d, i32]
    L.or_energy.argtypes = [C.POINTER(OrState), _dp]
    L.or_hsolve.argtypes = [_dp, _dp, i64, d, d, i32, d, _dp, C.c_void_p, C.c_void_p, C.c_int]
    L.or_hsolve.restype = i64
    L.or_hresidual.argtypes = [_dp, _dp, i64, d, _dp, _dp, C.c_void_p, C.c_int]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---- EOS / kernel (scalars) ----
def pressure(eos10, rho, u):
    return lib().or_pressure(_f64(eos10), float(rho), float(u))


def sound_speed_sq(eos10, rho, u):
    return lib().or_sound_speed_sq(_f64(eos10), float(rho), float(u))


def sound_speed_wgsl(eos10, rho, u):
    return lib().or_sound_speed_wgsl(_f64(eos10), float(rho), float(u))


def pressure_f32(eos10, rho, u):
    return lib().or_pressure_f32(_f32(eos10), float(rho), float(u))


def sph_w(r, h):
    return lib().or_sph_w(float(r), float(h))


def sph_dw(r, h):
    return lib().or_sph_dw(float(r), float(h))


# ---- LBVH leg ----
def expand_bits(v):
    return lib().or_expand_bits(int(v))


def bbox_f32(xyz):
    xyz = _f32(xyz)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().or_bbox_f32(xyz, xyz.shape[0], lo, hi)
    return lo, hi


def morton_f32(xyz, lo=None, hi=None):
    xyz = _f32(xyz)
    if lo is None:
        lo, hi = bbox_f32(xyz)
    codes = np.zeros(xyz.shape[0], np.uint32)
    lib().or_morton_f32(xyz, xyz.shape[0], _f32(lo), _f32(hi), codes)
    return codes


def sort_pairs(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    sc = np.zeros_like(codes)
    order = np.zeros_like(codes)
    lib().or_sort_pairs(codes, codes.shape[0], sc, order)
    return sc, order


def cluster_codes(sorted_codes, k):
    n = sorted_codes.shape[0]
    L = (n + k - 1) // k
    cc = np.zeros(L, np.uint32)
    lib().or_cluster_codes(np.ascontiguousarray(sorted_codes, np.uint32), n, k, cc)
    return cc


def karras(ccodes, fix_dups=False):
    """fix_dups=False: the reference's split (bh_gravity.wgsl:159-160) verbatim; True: the
    duplicate-run split consistent with the index-augmented ranges (what the GPU builds)."""
    ccodes = np.ascontiguousarray(ccodes, np.uint32)
    L = ccodes.shape[0]
    nn = 2 * L - 1
    left = np.zeros(nn, np.uint32)
    right = np.zeros(nn, np.uint32)
    parent = np.zeros(nn, np.uint32)
    lib().or_karras_ex(ccodes, L, int(bool(fix_dups)), left, right, parent)
    return left, right, parent


class Lbvh:
    """The reference's CPU Morton leg + Karras tree + moments (gpu_gravity.rs:318-478)."""

    def __init__(self, xyz32, mass32, k, fix_dups=True):
        xyz32 = _f32(xyz32)
        n = xyz32.shape[0]
        self.n, self.k = n, k
        self.lo, self.hi = bbox_f32(xyz32)
        self.codes = morton_f32(xyz32, self.lo, self.hi)
        self.sorted_codes, self.order = sort_pairs(self.codes)
        self.ccodes = cluster_codes(self.sorted_codes, k)
        self.L = self.ccodes.shape[0]
        self.left, self.right, self.parent = karras(self.ccodes, fix_dups)
        body = np.zeros((n, 4), np.float32)
        body[:, :3] = xyz32
        body[:, 3] = mass32
        self.sbody = np.ascontiguousarray(body[self.order])
        nn = 2 * self.L - 1
        self.com4 = np.zeros((nn, 4), np.float32)
        self.bmin = np.zeros((nn, 3), np.float32)
        self.bmax = np.zeros((nn, 3), np.float32)
        lib().or_upsweep_f32(self.sbody, n, k, self.left, self.right, self.L, self.com4, self.bmin, self.bmax)

    def walk_counts(self, theta, soft, targets, with_acc=False):
        """Per sorted target: (accepted nodes, leaf-particle interactions, opened nodes) of the oracle walk, and with
        with_acc its FP32 acceleration (G included)."""
        t = np.ascontiguousarray(targets, np.int64)
        counts = np.zeros((t.shape[0], 3), np.int64)
        acc = np.zeros((t.shape[0], 3), np.float32) if with_acc else None
        lib().or_bh_walk_count(self.sbody, self.n, self.k, self.left, self.right, self.com4, self.bmin, self.bmax,
                               self.L, float(theta), float(np.float32(soft * soft)), t, t.shape[0], counts,
                               acc.ctypes.data if with_acc else None)
        return (counts, acc) if with_acc else counts

    def walk(self, theta, soft):
        acc = np.zeros((self.n, 3), np.float32)
        lib().or_bh_walk_f32(self.sbody, self.order, self.n, self.k, self.left, self.right, self.com4,
                             self.bmin, self.bmax, self.L, float(theta), float(np.float32(soft * soft)), acc)
        return acc


# ---- FP64 gravity ----
def direct_f64(xyz, m, soft, targets=None, nthreads=0, with_pot=False):
    xyz = _f64(xyz)
    m = _f64(m)
    n = xyz.shape[0]
    if targets is not None:
        t = np.ascontiguousarray(targets, np.int64)
        nt = t.shape[0]
        tp = t.ctypes.data
    else:
        t, nt, tp = None, n, None
    acc = np.zeros((nt, 3), np.float64)
    pot = np.zeros(nt, np.float64) if with_pot else None
    lib().or_direct_f64(xyz, m, n, float(soft), tp, nt, acc, pot.ctypes.data if with_pot else None, nthreads)
    return (acc, pot) if with_pot else acc


def octree_f64(xyz, m, theta, soft, leaf_cap=1):
    xyz = _f64(xyz)
    acc = np.zeros_like(xyz)
    lib().or_octree_f64(xyz, _f64(m), xyz.shape[0], float(theta), float(soft), int(leaf_cap), acc)
    return acc


def grid_pairs(xyz, cell, radius=0.0):
    xyz = _f64(xyz)
    n = xyz.shape[0]
    need = lib().or_grid_pairs(xyz, n, float(cell), float(radius), None, 0)
    cnt = abs(need)
    pairs = np.zeros((max(cnt, 1), 2), np.int64)
    got = lib().or_grid_pairs(xyz, n, float(cell), float(radius), pairs.ctypes.data, cnt)
    assert got == cnt
    return pairs[:cnt]


# ---- HydroBody ----
class Hydro:
    """FP64 SPH state for the oracle (hydrostatic.rs HydroBody / sph-verify CpuState)."""

    def __init__(self, xyz, vel, u, h, m, mat, eos_table, softening, av_alpha=1.0, av_beta=2.0, theta=0.5,
                 gravity_mode=1, cs_fd=1, brute=False, leaf_cap=1, nthreads=0, h_solve=False, h_eta=2.2545,
                 h_tol=1e-4, native=False):
        self.xyz = _f64(xyz).copy()
        self.vel = _f64(vel).copy()
        self.u = _f64(u).copy()
        self.h = _f64(h).copy()
        self.m = _f64(m).copy()
        self.mat = np.ascontiguousarray(mat, np.int32).copy()
        self.eos = _f64(eos_table).copy()
        self._L = native_lib() if native else lib()
        self.st = OrState()
        s = self.st
        s.n = self.xyz.shape[0]
        s.xyz, s.vel, s.u = self.xyz.ctypes.data, self.vel.ctypes.data, self.u.ctypes.data
        s.h, s.m, s.mat = self.h.ctypes.data, self.m.ctypes.data, self.mat.ctypes.data
        s.eos = self.eos.ctypes.data
        s.n_mat = self.eos.shape[0]
        s.softening = float(softening)
        s.av_alpha, s.av_beta, s.theta = float(av_alpha), float(av_beta), float(theta)
        s.gravity_mode, s.cs_fd, s.brute = int(gravity_mode), int(cs_fd), int(bool(brute))
        s.leaf_cap, s.nthreads = int(leaf_cap), int(nthreads)
        s.h_solve, s.h_eta, s.h_tol = int(bool(h_solve)), float(h_eta), float(h_tol)

    def density(self):
        rho = np.zeros(self.st.n)
        self._L.or_density(C.byref(self.st), rho)
        return rho

    def forces(self, rho=None):
        if rho is None:
            rho = self.density()
        acc = np.zeros((self.st.n, 3))
        dudt = np.zeros(self.st.n)
        self._L.or_forces(C.byref(self.st), _f64(rho), acc, dudt)
        return acc, dudt

    def courant_dt(self, cfl, rho=None):
        if rho is None:
            rho = self.density()
        return self._L.or_courant_dt(C.byref(self.st), _f64(rho), float(cfl))

    def kdk(self, steps, dt=0.0, cfl=0.1):
        return self._L.or_kdk(C.byref(self.st), float(dt), float(cfl), int(steps))

    def relax(self, steps, dt, damp):
        """hydrostatic.rs:283-291 relax_step, `steps` times (AV off, u fixed)."""
        self._L.or_relax(C.byref(self.st), float(dt), float(damp), int(steps))

    def energy(self):
        out = np.zeros(9)
        self._L.or_energy(C.byref(self.st), out)
        return out


# ---- Newton h-solve (new; no reference counterpart) ----
def hsolve(xyz, m, h0, eta, tol=1e-4, max_iter=50, h_max=None, nthreads=0):
    xyz = _f64(xyz)
    m = _f64(m)
    h = _f64(h0).copy()
    if h_max is None:
        h_max = 4.0 * float(h.max())
    iters = np.zeros(xyz.shape[0], np.int32)
    resid = np.zeros(xyz.shape[0], np.float64)
    bad = lib().or_hsolve(xyz, m, xyz.shape[0], float(eta), float(tol), int(max_iter), float(h_max), h,
                          iters.ctypes.data, resid.ctypes.data, nthreads)
    return h, iters, resid, bad


def hresidual(xyz, m, h, eta, nthreads=0):
    xyz = _f64(xyz)
    n = xyz.shape[0]
    resid = np.zeros(n)
    nngb = np.zeros(n, np.int32)
    lib().or_hresidual(xyz, _f64(m), n, float(eta), _f64(h), resid, nngb.ctypes.data, nthreads)
    return resid, nngb
