"""Host-side mirror of the reference's GPU SPH stepper over the C ABI.

`GpuSph` follows `GpuSph` in ie/crates/engine/src/gpu_sph.rs:742-922 method for method
(new / upload / set_dt / set_av / encode_relax / encode_kdk / count / begin_readback / take_readback),
and `GravityField.accelerations` follows ie/crates/engine/src/gpu_gravity.rs:556-578. Differences
from the reference, by design:
  - errors raise `IhsError` (the reference clamps silently and reports wgpu errors asynchronously);
  - `set_dt(dt <= 0)` selects the device-side adaptive Courant dt (impact-run:403-410 read it back
    to the host every step);
  - the smoothing length can be solved per step (Newton) or seeded by the learned model.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .ics import SPH_PARTICLE_DTYPE

PRECISION = {"fp32": 0, "split": 1}
H_MODE = {"fixed": 0, "newton": 1, "learned": 2}
FIELDS = {"h": (0, np.float32, 1), "rho": (1, np.float32, 1), "rpc": (2, np.float32, 4), "asph": (3, np.float32, 4),
          "agrav": (4, np.float32, 4), "hiter": (5, np.uint32, 1), "ncount": (6, np.uint32, 1),
          "feat": (7, np.float32, 4), "pos64": (8, np.float64, 3), "hs": (9, np.float32, 1)}
IHS_EAGAIN = -11


class IhsError(RuntimeError):
    pass


def eos_records(eos) -> np.ndarray:
    """(n_mat, 10) Tillotson table -> the 48-byte SphEos records."""
    e = np.asarray(eos, dtype=np.float64)
    if e.ndim == 2 and e.shape[1] == 12 and e.dtype == np.float32:
        return np.ascontiguousarray(e)
    out = np.zeros((e.shape[0], 12), np.float32)
    out[:, :10] = e[:, :10]
    return out


def default_config() -> _lib.IhsConfig:
    cfg = _lib.IhsConfig()
    _lib.load().ihs_default_config(C.byref(cfg))
    return cfg


class GpuSph:
    def __init__(self, capacity: int, device: int = 0, leaf_size: int = 16, precision: str = "fp32",
                 h_mode: str = "newton", theta: float = 0.5, cfl: float = 0.1, n_ngb: float = 48.0,
                 h_tol: float = 1e-4, h_max_iter: int = 40, h_margin: float = 1.25, h_cap_factor: float = 4.0,
                 nbr_capacity: int = 192, use_graph: bool = True, h_margin_seeded: float = 1.03,
                 group_max: int = 64, quadrupole: bool = True, reuse_positions: bool = False,
                 fused_walk: bool = True, theta_mono: float = 0.0,
                 quad_tol: float = 0.0, theta_settled: float = 0.0):
        self._L = _lib.load()
        cfg = default_config()
        cfg.capacity = int(capacity)
        cfg.leaf_size = int(leaf_size)
        cfg.precision = PRECISION[precision]
        cfg.h_mode = H_MODE[h_mode]
        cfg.theta, cfg.cfl, cfg.n_ngb, cfg.h_tol = float(theta), float(cfl), float(n_ngb), float(h_tol)
        cfg.h_max_iter, cfg.h_margin, cfg.h_cap_factor = int(h_max_iter), float(h_margin), float(h_cap_factor)
        cfg.nbr_capacity, cfg.use_graph = int(nbr_capacity), int(bool(use_graph))
        cfg.h_margin_seeded, cfg.group_max = float(h_margin_seeded), int(group_max)
        cfg.quadrupole = int(bool(quadrupole))
        cfg.reuse_positions = int(bool(reuse_positions))
        cfg.fused_walk = int(bool(fused_walk))
        cfg.theta_mono = float(theta_mono)
        cfg.quad_tol = float(quad_tol)
        cfg.theta_settled = float(theta_settled)
        self.config = cfg
        h = C.c_void_p()
        rc = self._L.ihs_create(int(device), C.byref(cfg), C.byref(h))
        if rc != 0:
            raise IhsError(f"ihs_create failed ({rc}); check the device index and configuration")
        self._h = h

    @classmethod
    def from_config(cls, cfg: _lib.IhsConfig, device: int = 0) -> "GpuSph":
        """A handle created from a raw ihs_config (for example the one stored in a checkpoint)."""
        self = cls.__new__(cls)
        self._L = _lib.load()
        self.config = cfg
        h = C.c_void_p()
        rc = self._L.ihs_create(int(device), C.byref(cfg), C.byref(h))
        if rc != 0:
            raise IhsError(f"ihs_create failed ({rc}); check the device index and configuration")
        self._h = h
        return self

    # ---- lifetime ----
    def close(self):
        if getattr(self, "_h", None):
            self._L.ihs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            msg = self._L.ihs_last_error(self._h)
            raise IhsError(f"ihs error {rc}: {msg.decode() if msg else ''}")
        return rc

    # ---- GpuSph surface ----
    def upload(self, particles: np.ndarray, eos, softening: float):
        p = np.ascontiguousarray(particles, dtype=SPH_PARTICLE_DTYPE)
        e = eos_records(eos)
        self._check(self._L.ihs_upload(self._h, p.ctypes.data, p.shape[0], e.ctypes.data, e.shape[0],
                                       float(softening)))

    def upload_batch(self, scenarios, eos, softenings):
        """Several independent scenarios in one state (list of 48-byte particle arrays, one softening each): every
        pass of a step covers all of them in one launch, each with its own tree and Courant dt."""
        parts = [np.ascontiguousarray(p_, dtype=SPH_PARTICLE_DTYPE) for p_ in scenarios]
        allp = np.ascontiguousarray(np.concatenate(parts))
        counts = np.array([len(p_) for p_ in parts], np.int64)
        soft = np.ascontiguousarray(softenings, np.float32)
        assert soft.shape == counts.shape
        e = eos_records(eos)
        self._check(self._L.ihs_upload_batch(self._h, allp.ctypes.data, counts.ctypes.data, len(counts),
                                             e.ctypes.data, e.shape[0], soft.ctypes.data))
        self._counts = counts

    def segment_clock(self, seg: int) -> dict:
        t, dt, steps = C.c_double(), C.c_double(), C.c_int64()
        self._check(self._L.ihs_segment_clock(self._h, int(seg), C.byref(t), C.byref(dt), C.byref(steps)))
        return {"t": t.value, "dt": dt.value, "steps": steps.value}

    def set_dt(self, dt: float, damp: float = 1.0):
        self._check(self._L.ihs_set_dt(self._h, float(dt), float(damp)))

    def set_av(self, alpha: float, beta: float):
        self._check(self._L.ihs_set_av(self._h, float(alpha), float(beta)))

    def set_omega(self, omega: float):
        self._check(self._L.ihs_set_omega(self._h, float(omega)))

    def encode_kdk(self, substeps: int):
        self._check(self._L.ihs_step_kdk(self._h, int(substeps)))

    def encode_relax(self, steps: int):
        self._check(self._L.ihs_relax(self._h, int(steps)))

    def force_eval(self):
        self._check(self._L.ihs_force_eval(self._h))

    def count(self) -> int:
        return int(self._L.ihs_count(self._h))

    def sync(self):
        self._check(self._L.ihs_sync(self._h))

    def begin_readback(self):
        """gpu_store.rs:126-145: snapshot the state as it will be once the queued steps have run, without
        waiting (a no-op while a snapshot is in flight)."""
        self._check(self._L.ihs_begin_readback(self._h))

    def take_readback(self):
        """gpu_store.rs:147-152: that snapshot once its copy has completed, else None (pending or idle). Steps
        queued after begin_readback do not change it."""
        out = np.zeros(self.count(), SPH_PARTICLE_DTYPE)
        rc = self._L.ihs_take_readback(self._h, out.ctypes.data, out.shape[0])
        if rc == IHS_EAGAIN:
            return None
        n = self._check(rc)
        return out[:n]

    def readback(self) -> np.ndarray:
        out = np.zeros(self.count(), SPH_PARTICLE_DTYPE)
        self._check(self._L.ihs_readback(self._h, out.ctypes.data, out.shape[0], 1))
        return out

    # ---- diagnostics ----
    def stats(self, reset: bool = False) -> dict:
        s = _lib.IhsStats()
        self._check(self._L.ihs_stats_get(self._h, C.byref(s), int(bool(reset))))
        return {
            "ke": s.ke, "ie": s.ie, "pe": s.pe, "etot": s.ke + s.ie + s.pe, "p": list(s.p), "L": list(s.L), "t": s.t,
            "dt": s.dt, "steps": s.steps, "force_evals": s.force_evals, "mean_h_iters": s.mean_h_iters,
            "mean_h_evals": s.mean_h_evals, "mean_neighbours": s.mean_neighbours, "nbr_overflow": s.nbr_overflow,
            "walk_overflow": s.walk_overflow, "h_capped": s.h_capped, "h_unconverged": s.h_unconverged,
            "nonfinite": s.nonfinite, "tie_runs": s.tie_runs, "grav_cell_interactions": s.grav_cell_interactions,
            "grav_pp_interactions": s.grav_pp_interactions, "nbr_candidates": s.nbr_candidates,
            "walk_nodes": s.walk_nodes,
        }

    def segment_energy(self, seg: int = 0) -> dict:
        """One scenario's FP64 KE, IE, PE (1/2 sum m phi), momentum and angular momentum about the origin."""
        e = np.zeros(9, np.float64)
        self._check(self._L.ihs_segment_energy(self._h, int(seg), e.ctypes.data))
        return {"ke": e[0], "ie": e[1], "pe": e[2], "etot": e[0] + e[1] + e[2], "p": list(e[3:6]),
                "L": list(e[6:9])}

    def profile(self, enable: bool = True):
        """Per-stage HIP-event timing (steps launch without the graph while enabled)."""
        self._check(self._L.ihs_profile(self._h, int(bool(enable))))

    def profile_read(self, reset: bool = False) -> dict:
        ms = np.zeros(len(_lib.STAGES), np.float64)
        calls = np.zeros(len(_lib.STAGES), np.int64)
        self._check(self._L.ihs_profile_read(self._h, ms.ctypes.data, calls.ctypes.data, int(bool(reset))))
        return {name: (float(ms[i]), int(calls[i])) for i, name in enumerate(_lib.STAGES)}

    def field(self, name: str) -> np.ndarray:
        fid, dt, width = FIELDS[name]
        n = self.count()
        out = np.zeros((n, width) if width > 1 else n, dt)
        self._check(self._L.ihs_debug_field(self._h, fid, out.ctypes.data))
        return out

    def debug_tree(self) -> dict:
        n = self.count()
        k = self.config.leaf_size
        L = (n + k - 1) // k
        sc = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        cc = np.zeros(L, np.uint32)
        left = np.zeros(2 * L - 1, np.uint32)
        right = np.zeros(2 * L - 1, np.uint32)
        parent = np.zeros(2 * L - 1, np.uint32)
        got = self._check(self._L.ihs_debug_tree(self._h, sc.ctypes.data, order.ctypes.data, cc.ctypes.data,
                                                 left.ctypes.data, right.ctypes.data, parent.ctypes.data))
        assert got == L
        return dict(sorted_codes=sc, order=order, ccodes=cc, left=left, right=right, parent=parent)

    def direct_f64(self, ids) -> np.ndarray:
        """FP64 direct-sum gravitational accelerations (n_t, 3) at the given upload-order ids (force-error probe)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        acc = np.zeros((ids.shape[0], 3), np.float64)
        self._check(self._L.ihs_direct_f64(self._h, ids.ctypes.data, ids.shape[0], acc.ctypes.data))
        return acc

    def direct_f64_noalloc(self, ids) -> np.ndarray:
        """direct_f64 through the allocation-free entry point (at most 65,536 targets per call)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        acc = np.zeros((ids.shape[0], 3), np.float64)
        self._check(self._L.ihs_direct_f64_noalloc(self._h, ids.ctypes.data, ids.shape[0], acc.ctypes.data))
        return acc

    def gravity_error(self, n_targets: int = 8192, seed: int = 7) -> dict:
        """Error of the last evaluation's tree-walk gravity against the FP64 direct sum on sampled targets.

        rms: the RMS over targets of each target's relative error |a - a_ref| / |a_ref| (the bound the autotuner
        and the tests enforce; it weighs targets with weak gravity fully, so it is the stricter of the two norms
        on these bodies). rms_ref_norm: gpu-bh-verify's norm (ie/tools/gpu-bh-verify/src/main.rs:164-181),
        sqrt(sum |a - a_ref|^2 / sum |a_ref|^2). max: the worst target's relative error."""
        n = self.count()
        rng = np.random.default_rng(seed)
        ids = np.sort(rng.choice(n, size=min(n_targets, n), replace=False)).astype(np.uint32)
        ref = self.direct_f64(ids)
        got = self.field("agrav")[ids, :3].astype(np.float64)
        e2 = ((got - ref) ** 2).sum(1)
        r2 = (ref ** 2).sum(1)
        err = np.sqrt(e2) / np.maximum(np.sqrt(r2), 1e-300)
        return {"rms": float(np.sqrt(np.mean(err ** 2))), "rms_ref_norm": float(np.sqrt(e2.sum() / max(r2.sum(), 1e-300))),
                "max": float(err.max()), "targets": int(len(ids)), "ids": ids, "ref": ref}

    # ---- learned smoothing-length seed ----
    def set_mlp_weights(self, w):
        w = np.ascontiguousarray(w, np.float32).ravel()
        self._check(self._L.ihs_mlp_set_weights(self._h, w.ctypes.data, w.shape[0]))

    def load_mlp(self, path: str):
        """Loads trained weights written by `mlp.save_weights` and enables the seed."""
        from .mlp import load_weights
        self.set_mlp_weights(load_weights(path))

    def mlp_weights(self) -> np.ndarray:
        w = np.zeros(_lib.MLP_NPARAM, np.float32)
        self._check(self._L.ihs_mlp_get_weights(self._h, w.ctypes.data, w.shape[0]))
        return w

    def mlp_record(self, max_samples: int):
        """Start (max_samples > 0) or stop (0) recording (features, target) rows at post-drift Newton solves."""
        self._check(self._L.ihs_mlp_record(self._h, int(max_samples)))

    def mlp_samples(self) -> np.ndarray:
        k = self._check(self._L.ihs_mlp_samples(self._h, None, 0))
        out = np.zeros((k, _lib.MLP_SAMPLE), np.float32)
        if k:
            got = self._check(self._L.ihs_mlp_samples(self._h, out.ctypes.data, k))
            out = out[:got]
        return out

    def mlp_train(self, steps: int, batch: int, lr: float, seed: int = 0, rows: np.ndarray | None = None) -> np.ndarray:
        """Adam on the device over the recorded rows (or `rows`, shape (n, 8)); returns the per-step loss."""
        loss = np.zeros(max(int(steps), 1), np.float32)
        if rows is not None:
            rows = np.ascontiguousarray(rows, np.float32)
            assert rows.ndim == 2 and rows.shape[1] == _lib.MLP_SAMPLE
            ptr, n = rows.ctypes.data, rows.shape[0]
        else:
            ptr, n = None, 0
        self._check(self._L.ihs_mlp_train(self._h, ptr, n, int(steps), int(batch), float(lr),
                                          int(seed) & (2 ** 64 - 1), loss.ctypes.data))
        return loss[:int(steps)]

    def mlp_gradient(self, x, y) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        assert x.ndim == 2 and x.shape[1] == _lib.MLP_IN and y.shape == (x.shape[0],)
        g = np.zeros(_lib.MLP_NPARAM + 1, np.float32)
        self._check(self._L.ihs_mlp_gradient(self._h, x.ctypes.data, y.ctypes.data, x.shape[0], g.ctypes.data))
        return g[:-1], float(g[-1])

    def mlp_predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 2 and x.shape[1] == _lib.MLP_IN
        y = np.zeros(x.shape[0], np.float32)
        self._check(self._L.ihs_mlp_predict(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    # ---- GravityField ----
    def gravity(self, pos, mass, softening: float, theta: float = 0.5, mode: str = "tree", with_pot: bool = False):
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        n = pos.shape[0]
        acc = np.zeros((n, 3), np.float64)
        pot = np.zeros(n, np.float64) if with_pot else None
        self._check(self._L.ihs_gravity(self._h, pos.ctypes.data, mass.ctypes.data, n, float(softening), float(theta),
                                        1 if mode == "tree" else 0, acc.ctypes.data,
                                        pot.ctypes.data if with_pot else None))
        return (acc, pot) if with_pot else acc


def gravity_route(n: int, knee: int) -> str:
    """gpu_gravity.rs:556-578: the exact direct sum below the knee, the LBVH tree at or above it."""
    return "tree" if n >= knee else "direct"


class GravityField:
    """gpu_gravity.rs:532-626: direct sum below the knee, LBVH tree at or above it (native: synchronous).

    The reference's knee (TREE_KNEE) was measured on its own box; `autotune.gravity_knee` measures this device's,
    stored in data/tuned.json with the leaf size of the faster tree, and that pair is used when present."""

    TREE_KNEE = 24_000  # gpu_gravity.rs:50 (Apple M4 Max)
    THETA = 0.5         # gpu_gravity.rs:55

    def __init__(self, capacity: int, device: int = 0, leaf_size: int | None = None, knee: int | None = None):
        from .autotune import tuned_knee

        measured = tuned_knee()
        self.knee = int(knee if knee is not None else (measured[0] if measured else self.TREE_KNEE))
        if leaf_size is None:
            leaf_size = measured[1] if measured else 16
        self._sph = GpuSph(capacity, device=device, leaf_size=leaf_size, h_mode="fixed", nbr_capacity=1,
                           use_graph=False)

    def accelerations(self, pos, mass, softening: float):
        if len(pos) == 0:
            return np.zeros((0, 3))
        return self._sph.gravity(pos, mass, softening, self.THETA, gravity_route(len(pos), self.knee))

    def close(self):
        self._sph.close()
