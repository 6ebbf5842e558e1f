"""Autotuner for the tree walk: multipole order, leaf size K, walk-group size and opening angle theta.

The step's cost is dominated by the Barnes-Hut walk. Its cost and its force error both depend on:
- the leaf bucket K (exact particle-particle sums inside opened leaves vs monopoles of small subtrees);
- the group size (how many particles share one walk);
- theta;
- the multipole order of accepted cells (monopole, as the reference, or monopole + quadrupole);
- with quadrupoles, which accepted cells may act as monopoles: those that also pass the opening test at a smaller
  angle theta_mono, or whose quadrupole term is below quad_tol times their monopole term (`tune_monopole`, a second
  stage at the chosen leaf, group and theta).

The reference fixes theta = 0.5 with K = 1 (ie/crates/engine/src/gpu_gravity.rs:55,61) at an RMS force error of
0.70% (gpu-bh-verify:552-569). Here the force error is a bound, RMS relative error <= 1e-3 against the FP64
direct sum on 8,192 sampled targets, so the tuner picks the fastest setting whose measured error stays
below the bound with a margin.

Each candidate is measured on the device:
- ms per KDK step of the plain Newton step, from direct launches;
- the walk's RMS/max error from `GpuSph.gravity_error`.

Results are stored per particle-count bucket in botheadstudios_amd/data/tuned.json, which bench.py and callers
read through `tuned_params(n)`.

    python -m botheadstudios_amd.autotune --cfg 4          # tune for the 1M-particle impact
    python -m botheadstudios_amd.autotune --cfg 4 --mono   # only the monopole-cell stage, at the tuned walk
    python -m botheadstudios_amd.autotune --knee           # the standalone gravity's direct-vs-tree knee
"""
from __future__ import annotations

import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "data", "tuned.json")
DEFAULT = {"leaf": 16, "group": 64, "theta": 0.5, "quadrupole": True, "theta_mono": 0.0, "quad_tol": 0.0,
           "theta_settled": 0.0}
QUAD_THETAS = (1.2, 1.1, 1.0, 0.95, 0.9, 0.875, 0.85, 0.825, 0.8, 0.75, 0.7, 0.6, 0.5, 0.45, 0.4)
MONO_THETAS = (0.5, 0.45, 0.42, 0.4, 0.375, 0.35)


MONO_GRID = ((0.0, 0.0), (0.0, 0.03), (0.5, 0.03), (0.55, 0.03), (0.6, 0.03), (0.5, 0.05), (0.6, 0.0))


def measure(ic, leaf: int, theta: float, group: int = 64, quadrupole: bool = True, steps: int = 3, warmup: int = 1,
            n_targets: int = 8192, device: int = 0, theta_mono: float = 0.0, quad_tol: float = 0.0,
            theta_settled: float = 0.0) -> dict:
    """ms/step (plain Newton, FP32) and the walk's force error for one setting on an impact IC. With theta_settled,
    also the error of the walk that keeps the lists (an evaluation at the unchanged positions after the steps);
    err_rms is then the larger of the two."""
    from .gpu_sph import GpuSph

    with GpuSph(ic.n, device=device, leaf_size=leaf, theta=theta, group_max=group, quadrupole=quadrupole,
                h_mode="newton", use_graph=True, theta_mono=theta_mono, quad_tol=quad_tol,
                theta_settled=theta_settled) as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)
        sph.encode_kdk(warmup)
        sph.sync()
        t0 = time.perf_counter()
        sph.encode_kdk(steps)
        sph.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        st = sph.stats()
        err = sph.gravity_error(n_targets)
        if theta_settled > 0:
            sph.force_eval()  # positions and h unchanged: the lists are kept and the walk runs at theta_settled
            e2 = sph.gravity_error(n_targets)
            err = {"rms": max(err["rms"], e2["rms"]), "max": max(err["max"], e2["max"]), "rms_settled": e2["rms"]}
    return {"leaf": leaf, "group": group, "theta": theta, "quadrupole": bool(quadrupole), "theta_mono": theta_mono,
            "quad_tol": quad_tol, "theta_settled": theta_settled, "ms_per_step": ms, "err_rms": err["rms"],
            "err_rms_settled": err.get("rms_settled"),
            "err_max": err["max"], "ok": st["nbr_overflow"] == 0 and st["walk_overflow"] == 0 and st["nonfinite"] == 0}


def tune(ic, err_target: float = 1e-3, margin: float = 0.92, leaves=(1, 2, 4, 8, 16), groups=(128, 64),
         multipoles=((True, QUAD_THETAS), (False, MONO_THETAS)), steps: int = 3, log=print) -> dict:
    """Fastest (multipole order, leaf, group, theta) with err_rms <= margin * err_target.

    For each (order, leaf, group), theta is lowered from the largest candidate until the bound holds. A smaller
    theta is only slower, so the search stops there."""
    table = []
    for quad, thetas in multipoles:
        for leaf in leaves:
            for group in groups:
                if group < leaf:
                    continue
                for th in thetas:
                    m = measure(ic, leaf, th, group, quad, steps=steps)
                    table.append(m)
                    log(f"  {'quad' if quad else 'mono'} K={leaf:2d} group={group} theta={th:.3f}: "
                        f"{m['ms_per_step']:.2f} ms/step, err rms {m['err_rms']:.2e} max {m['err_max']:.2e}")
                    if m["ok"] and m["err_rms"] <= margin * err_target:
                        break
    feasible = [m for m in table if m["ok"] and m["err_rms"] <= margin * err_target]
    best = min(feasible, key=lambda m: m["ms_per_step"]) if feasible else None
    return {"n": ic.n, "err_target": err_target, "margin": margin, "best": best, "table": table}


def tune_monopole(ic, base: dict, err_target: float = 1e-3, margin: float = 0.92, grid=MONO_GRID, steps: int = 3,
                  reps: int = 2, log=print) -> dict:
    """Second stage at a quadrupole walk's (leaf, group, theta): the fastest (theta_mono, quad_tol) whose error stays
    within margin * err_target. Each setting is timed reps times, interleaved, and its best time kept."""
    table = {}
    for _ in range(reps):
        for tm, qt in grid:
            if tm > base["theta"]:
                continue
            m = measure(ic, base["leaf"], base["theta"], base["group"], True, steps=steps, theta_mono=tm, quad_tol=qt)
            k = (tm, qt)
            if k not in table or m["ms_per_step"] < table[k]["ms_per_step"]:
                table[k] = m
            log(f"  theta {base['theta']:.3f} theta_mono {tm:.2f} quad_tol {qt:.3f}: {m['ms_per_step']:.3f} ms/step, "
                f"err rms {m['err_rms']:.2e} max {m['err_max']:.2e}")
    rows = list(table.values())
    feasible = [m for m in rows if m["ok"] and m["err_rms"] <= margin * err_target]
    best = min(feasible, key=lambda m: m["ms_per_step"]) if feasible else None
    return {"best": best, "table": rows}


SETTLED_THETAS = (0.9, 0.85, 0.8, 0.75, 0.7, 0.65, 0.6)


def tune_settled(ic, base: dict, err_target: float = 1e-3, margin: float = 0.92, thetas=SETTLED_THETAS,
                 steps: int = 3, reps: int = 2, log=print) -> dict:
    """Third stage: the opening angle of the walk that keeps the lists (the first evaluation of a KDK step), at the
    tuned walk. Without the neighbour search's opened nodes that walk is less accurate at the same angle; the
    fastest theta_settled whose error (and the list-building walk's) stays within margin * err_target wins."""
    table = {}
    for _ in range(reps):
        for th in thetas:
            if th > base["theta"]:
                continue
            m = measure(ic, base["leaf"], base["theta"], base["group"], base["quadrupole"], steps=steps,
                        theta_mono=base.get("theta_mono", 0.0), quad_tol=base.get("quad_tol", 0.0), theta_settled=th)
            if th not in table or m["ms_per_step"] < table[th]["ms_per_step"]:
                table[th] = m
            log(f"  theta_settled {th:.3f}: {m['ms_per_step']:.3f} ms/step, err rms {m['err_rms']:.2e} "
                f"(kept lists {m['err_rms_settled']:.2e})")
    rows = list(table.values())
    feasible = [m for m in rows if m["ok"] and m["err_rms"] <= margin * err_target]
    best = min(feasible, key=lambda m: m["ms_per_step"]) if feasible else None
    if best is not None and best["ms_per_step"] >= base.get("ms_per_step", float("inf")):
        best = None  # keeping the lists must pay for itself
    return {"best": best, "table": rows}


def bucket(n: int) -> str:
    """Particle-count bucket key (powers of two)."""
    return str(1 << max(0, int(round(math.log2(max(n, 1))))))


def save(result: dict, path: str = TUNED):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[bucket(result["n"])] = result
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def tuned_params(n: int, path: str = TUNED) -> dict:
    """The tuned {leaf, group, theta, quadrupole, theta_mono, quad_tol} for n particles (nearest tuned bucket), or
    the defaults."""
    if not os.path.exists(path):
        return dict(DEFAULT)
    with open(path) as f:
        data = json.load(f)
    cands = [(abs(math.log2(int(k)) - math.log2(max(n, 1))), v) for k, v in data.items() if v.get("best")]
    if not cands:
        return dict(DEFAULT)
    best = min(cands, key=lambda c: c[0])[1]["best"]
    return {"leaf": int(best["leaf"]), "group": int(best["group"]), "theta": float(best["theta"]),
            "quadrupole": bool(best.get("quadrupole", False)), "theta_mono": float(best.get("theta_mono", 0.0)),
            "quad_tol": float(best.get("quad_tol", 0.0)), "theta_settled": float(best.get("theta_settled", 0.0))}


def walk_kwargs(params: dict, leaf: bool = True) -> dict:
    """GpuSph keyword arguments of a tuned walk setting (leaf_size left out with leaf=False)."""
    kw = dict(group_max=params["group"], theta=params["theta"], quadrupole=params["quadrupole"],
              theta_mono=params.get("theta_mono", 0.0), quad_tol=params.get("quad_tol", 0.0),
              theta_settled=params.get("theta_settled", 0.0))
    if leaf:
        kw["leaf_size"] = params["leaf"]
    return kw


KNEE_NS = (1500, 3000, 6000, 12000, 24000, 48000, 96000, 192000, 384000)


def gravity_knee(ns=KNEE_NS, leaves=(1, 16), theta: float = 0.5, reps: int = 5, seed: int = 11, device: int = 0,
                 log=print) -> dict:
    """The direct-vs-tree knee of the standalone gravity (GravityField, ihs_gravity) measured on this device.

    The reference reads its knee off the same comparison on its own box (gpu_gravity.rs:40-50: N = 24,000 on an
    Apple M4 Max); both paths here include the host-to-device copy of the bodies and the copy back of the
    accelerations, as a GravityField call does. Clouds are uniform balls of equal masses. For each leaf size the
    tree's and the direct sum's ms per call are measured at every N; the leaf with the fastest tree at the largest N
    is kept, and the knee is the smallest N from which that tree is faster at every larger measured N."""
    import numpy as np

    from .gpu_sph import GpuSph

    rng = np.random.default_rng(seed)
    table = []
    for leaf in leaves:
        with GpuSph(max(ns), device=device, leaf_size=leaf, h_mode="fixed", nbr_capacity=1, use_graph=False) as s:
            for n in ns:
                v = rng.normal(size=(n, 3))
                pos = v / np.linalg.norm(v, axis=1, keepdims=True) * (6.4e6 * np.cbrt(rng.random(n)))[:, None]
                mass = np.full(n, 6.0e24 / n)
                soft = 0.25 * 6.4e6 / np.cbrt(n)
                row = {"n": n, "leaf": leaf}
                for mode in ("direct", "tree"):
                    s.gravity(pos, mass, soft, theta, mode)
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        s.gravity(pos, mass, soft, theta, mode)
                    row[f"{mode}_ms"] = 1e3 * (time.perf_counter() - t0) / reps
                table.append(row)
                log(f"leaf {leaf} N={n}: direct {row['direct_ms']:.3f} ms, tree {row['tree_ms']:.3f} ms")
    nmax = max(ns)
    leaf = min(leaves, key=lambda k: next(r["tree_ms"] for r in table if r["leaf"] == k and r["n"] == nmax))
    rows = sorted((r for r in table if r["leaf"] == leaf), key=lambda r: r["n"])
    knee = nmax * 2
    for i in range(len(rows) - 1, -1, -1):
        if rows[i]["tree_ms"] < rows[i]["direct_ms"]:
            knee = rows[i]["n"]
        else:
            break
    return {"knee": int(knee), "leaf": int(leaf), "theta": theta, "table": table}


def save_knee(result: dict, path: str = TUNED):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data["gravity_knee"] = result
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def tuned_knee(path: str = TUNED):
    """(knee, leaf) of the standalone gravity measured on this device type, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("gravity_knee")
    return (int(k["knee"]), int(k["leaf"])) if k else None


def main(argv=None):
    import argparse

    from . import ics

    ap = argparse.ArgumentParser()
    ap.add_argument("--knee", action="store_true", help="measure the standalone gravity's direct-vs-tree knee")
    ap.add_argument("--cfg", type=int, default=4)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--err", type=float, default=1e-3)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=TUNED)
    ap.add_argument("--mono", action="store_true", help="only the monopole-cell stage, at the stored tuned walk")
    ap.add_argument("--settled", action="store_true",
                    help="only the kept-list walk's opening angle, at the stored tuned walk")
    a = ap.parse_args(argv)
    if a.knee:
        res = gravity_knee()
        save_knee(res, a.out)
        print(json.dumps({"knee": res["knee"], "leaf": res["leaf"]}), flush=True)
        return
    ic = ics.config_ic(a.cfg, n=a.n)
    t0 = time.time()
    if a.mono or a.settled:
        with open(a.out) as f:
            res = json.load(f)[bucket(ic.n)]
    else:
        res = tune(ic, err_target=a.err, steps=a.steps)
    if not a.settled and res["best"] and res["best"].get("quadrupole"):
        mono = tune_monopole(ic, res["best"], err_target=a.err, margin=res.get("margin", 0.92), steps=a.steps)
        res["monopole_table"] = mono["table"]
        if mono["best"]:
            res["best"] = mono["best"]
    if res["best"]:
        base = dict(res["best"], theta_settled=0.0)
        if a.settled:  # the list-building step's time at the same moment, for the comparison
            base.update(measure(ic, base["leaf"], base["theta"], base["group"], base["quadrupole"], steps=a.steps,
                                theta_mono=base.get("theta_mono", 0.0), quad_tol=base.get("quad_tol", 0.0)))
        st = tune_settled(ic, base, err_target=a.err, margin=res.get("margin", 0.92), steps=a.steps)
        res["settled_table"] = st["table"]
        res["best"] = st["best"] or base
    save(res, a.out)
    print(json.dumps({"best": res["best"], "n": ic.n, "seconds": round(time.time() - t0, 1)}), flush=True)


if __name__ == "__main__":
    main()
