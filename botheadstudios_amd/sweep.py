"""Batched impact-parameter sweep: many scenarios in one state against the same scenarios one at a time.

Scenarios come from `ics.sweep_scenarios`: N log-uniform in [4,096, 131,072], impactor 0.05-0.2 Earth masses,
angle 0-60 degrees, 1.0-1.5 v_esc. The reference runs such ensembles one scenario after another
(ie/tools/impact-run/src/main.rs:645-671). Here `GpuSph.upload_batch` puts every scenario in one state, so that
each pass of a KDK step covers all of them in one launch, in one captured graph. Each scenario keeps its own
tree, walks, softening and Courant dt, and the result is bit-identical to running it alone (tests/test_gpu_batch.py).

    python -m botheadstudios_amd.sweep --count 64 --steps 10
"""
from __future__ import annotations

import json
import time

import numpy as np


def build(count: int = 64, seed: int | None = None):
    from . import ics

    kw = {} if seed is None else {"seed": seed}
    return [ics.impact_ic(**s) for s in ics.sweep_scenarios(count, **kw)]


def _handle(capacity, params, h_mode, precision, weights):
    from . import autotune
    from .gpu_sph import GpuSph

    sph = GpuSph(capacity, leaf_size=1, precision=precision, h_mode=h_mode,
                 **autotune.walk_kwargs(params, leaf=False))
    if weights is not None and h_mode == "learned":
        sph.load_mlp(weights)
    return sph


def run_batched(scen, steps: int, warmup: int, params: dict, h_mode: str = "learned", precision: str = "fp32",
                weights=None) -> dict:
    ntot = sum(ic.n for ic in scen)
    sph = _handle(ntot, params, h_mode, precision, weights)
    try:
        sph.upload_batch([ic.particles48() for ic in scen], scen[0].eos, [ic.softening for ic in scen])
        sph.set_dt(0.0)
        sph.encode_kdk(warmup)
        sph.sync()
        t0 = time.perf_counter()
        sph.encode_kdk(steps)
        sph.sync()
        sec = time.perf_counter() - t0
        st = sph.stats()
        clocks = [sph.segment_clock(k) for k in range(len(scen))]
    finally:
        sph.close()
    return {"sec": sec, "particle_steps_per_s": ntot * steps / sec, "ms_per_step": 1e3 * sec / steps,
            "nonfinite": st["nonfinite"], "h_capped": st["h_capped"], "h_unconverged": st["h_unconverged"],
            "dt_min": min(c["dt"] for c in clocks), "dt_max": max(c["dt"] for c in clocks)}


def run_sequential(scen, steps: int, warmup: int, params: dict, h_mode: str = "learned", precision: str = "fp32",
                   weights=None) -> dict:
    """The same scenarios one at a time: each advances `steps` (captured) steps and completes before the next starts;
    set-up and warm-up are outside the timed region, as for the batch."""
    handles = []
    try:
        for ic in scen:
            sph = _handle(ic.n, params, h_mode, precision, weights)
            handles.append(sph)
            sph.upload(ic.particles48(), ic.eos, ic.softening)
            sph.set_dt(0.0)
            sph.encode_kdk(warmup)
            sph.sync()
        t0 = time.perf_counter()
        for sph in handles:
            sph.encode_kdk(steps)
            sph.sync()
        sec = time.perf_counter() - t0
    finally:
        for sph in handles:
            sph.close()
    ntot = sum(ic.n for ic in scen)
    return {"sec": sec, "particle_steps_per_s": ntot * steps / sec, "ms_per_step": 1e3 * sec / steps}


def measure(count: int = 64, steps: int = 10, warmup: int = 2, h_mode: str = "learned", precision: str = "fp32",
            log=print) -> dict:
    from . import autotune, mlp

    scen = build(count)
    params = autotune.tuned_params(int(np.mean([ic.n for ic in scen])))
    weights = mlp.default_weights_path()
    mode = h_mode if weights is not None else "newton"
    log(f"sweep: {count} scenarios, {sum(ic.n for ic in scen)} particles, {mode}/{precision}, walk {params}")
    b = run_batched(scen, steps, warmup, params, mode, precision, weights)
    log(f"  batched    {b['ms_per_step']:.2f} ms/step, {b['particle_steps_per_s']:.3e} particle-steps/s")
    q = run_sequential(scen, steps, warmup, params, mode, precision, weights)
    log(f"  one by one {q['ms_per_step']:.2f} ms/step, {q['particle_steps_per_s']:.3e} particle-steps/s")
    return {"scenarios": count, "particles": int(sum(ic.n for ic in scen)),
            "n_min": int(min(ic.n for ic in scen)), "n_max": int(max(ic.n for ic in scen)), "steps": steps,
            "h_mode": mode, "precision": precision, "walk": params, "batched": b, "sequential": q,
            "speedup": b["particle_steps_per_s"] / q["particle_steps_per_s"]}


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--h-mode", default="learned")
    ap.add_argument("--precision", default="fp32")
    a = ap.parse_args(argv)
    res = measure(a.count, a.steps, a.warmup, a.h_mode, a.precision)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
