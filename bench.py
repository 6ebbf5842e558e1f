"""Benchmark: SPH + self-gravity giant-impact KDK steps at 1,048,576 particles on one MI355X.

One "step" is the reference's leapfrog step: two full force evaluations (Morton sort, LBVH, neighbour lists,
smoothing-length solve, density/EOS, SPH forces, Barnes-Hut gravity) with a half-kick + drift between them
and a half-kick after (ie/shaders/sph_step.wgsl:213-237, ie/crates/engine/src/hydrostatic.rs:345-359), with
the Courant dt formed on the device each step. Inputs: the two-body impact (0.9 + 0.1 Earth masses, 45 deg,
1.1 v_esc, 30% iron cores, granite mantles, seeded jittered lattices) at N = 1,048,576, already resident in
HBM when the timed region starts.

Prints one JSON line. `value` is the particle-steps/s of the headline variant (learned-h seed + split
precision when a trained network is available, otherwise the plain Newton FP32 step, named in config).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

METRIC = "SPH+gravity particle-steps/s at 1M particles, learned-h vs plain, equal tolerance"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOPS_PER_INTERACTION = 20.0  # softened monopole / particle-particle interaction (standard count)
# Monopole + traceless quadrupole cell: 50 flops for the acceleration (see interact_quad in k_gravity.hip:
# 3 sub, 3 FMA r^2, rsqrt, 5 mul for M r^-k, 15 for q.d, 5 for d.q.d, 3 for the radial factor, 12 for the
# accumulation) plus 4 for the potential, which only the second evaluation of a step computes: 52 on average.
FLOPS_PER_QUAD_CELL = 52.0
FUSED_WALK = True  # the gravity walk writes the neighbour lists (ihs_config.fused_walk, the library default)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--leaf", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--theta", type=float, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--group", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--quadrupole", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--variants", default="plain,learned,reuse",
                    help="plain: Newton FP32; learned: learned seed + split precision; reuse: both again with the "
                         "first evaluation of each step reusing the previous one's position-only work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched 64-scenario sweep line")
    ap.add_argument("--sweep-steps", type=int, default=10)
    ap.add_argument("--no-measure", action="store_true", help="skip the disk / remnant measurement line")
    ap.add_argument("--measure-only", action="store_true",
                    help="run only the disk / remnant measurement leg (prints its object as the JSON line)")
    ap.add_argument("--measure-reps", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (hipMemcpy): the known-size kernel that calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    return ap.parse_args()


def copy_probe(mib: int):
    """One device-to-device copy of mib MiB (the runtime's copy kernel reads and writes exactly that much)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    size = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), size) or hip.hipMalloc(C.byref(b), size):
        raise RuntimeError("hipMalloc failed for the copy probe")
    try:
        if hip.hipMemset(a, 1, size) or hip.hipMemcpy(b, a, size, 3) or hip.hipDeviceSynchronize():
            raise RuntimeError("copy probe failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    log(f"copy probe: {mib} MiB device to device")


def measured_traffic(n, args, variant):
    """HBM bytes per k_gravity_walk launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when that
    summary was taken on this workload, walk setting and variant (else None)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_gravity_walk_1M_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        line = next((l for l in f if l.startswith("traffic: ")), None)
    if line is None:
        return None
    rec = json.loads(line[len("traffic: "):])
    same = (rec.get("n") == n and rec.get("leaf") == args.leaf and rec.get("group") == args.group and
            abs(rec.get("theta", -1) - args.theta) < 1e-6 and rec.get("quadrupole") == bool(args.quadrupole) and
            rec.get("variant") == variant and rec.get("fused_walk", False) == FUSED_WALK)
    return rec["bytes_per_launch"] if same else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_variant(ic, precision, h_mode, args, weights=None, reuse=False):
    from botheadstudios_amd.gpu_sph import GpuSph

    sph = GpuSph(ic.n, leaf_size=args.leaf, precision=precision, h_mode=h_mode, theta=args.theta,
                 group_max=args.group, quadrupole=args.quadrupole, reuse_positions=reuse)
    try:
        if weights is not None:
            sph.load_mlp(weights)
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)      # adaptive Courant dt on the device
        sph.set_av(1.0, 2.0)
        sph.encode_kdk(args.warmup)
        sph.sync()
        sph.stats(reset=True)
        t0 = time.perf_counter()
        sph.encode_kdk(args.steps)
        sph.sync()
        t1 = time.perf_counter()
        st = sph.stats(reset=True)
        # Per-stage HIP-event timing over a separate profiled pass (direct launches, same kernels).
        sph.profile(True)
        sph.encode_kdk(args.profile_steps)
        sph.sync()
        prof = sph.profile_read(reset=True)
        pst = sph.stats(reset=True)
        sph.profile(False)
        # Force error of the last evaluation's tree walk against the FP64 direct sum on 8192 sampled targets.
        gerr = sph.gravity_error(8192)
        for k in ("nbr_overflow", "walk_overflow", "nonfinite"):
            if st[k] or pst[k]:
                raise RuntimeError(f"{precision}/{h_mode}: {k} = {st[k]}/{pst[k]}")
        sec = t1 - t0
        return dict(sec=sec, ms_per_step=1e3 * sec / args.steps, value=ic.n * args.steps / sec, stats=st,
                    profile=prof, profile_stats=pst, gravity_error=gerr)
    finally:
        sph.close()


def roofline(res, n, args, variant):
    """Dominant stage -> achieved vs peak. Gravity (FP32 VALU): 20 flops per particle-particle or monopole cell
    interaction, 52 per quadrupole cell interaction."""
    prof = res["profile"]
    stage = max(prof, key=lambda k: prof[k][0])
    ms, calls = prof[stage]
    avg_ms = ms / max(calls, 1)
    pst = res["profile_stats"]
    if stage == "gravity":
        inter = pst["grav_cell_interactions"] + pst["grav_pp_interactions"]
        cell_flops = FLOPS_PER_QUAD_CELL if args.quadrupole else FLOPS_PER_INTERACTION
        flops = (cell_flops * pst["grav_cell_interactions"] +
                 FLOPS_PER_INTERACTION * pst["grav_pp_interactions"]) / max(calls, 1)
        achieved = flops / (avg_ms * 1e-3) / 1e12
        return dict(kernel="k_gravity_walk", bound="mfma", achieved=achieved, peak=FP32_PEAK_TFLOPS,
                    unit="TFLOP/s", frac=achieved / FP32_PEAK_TFLOPS, traffic=measured_traffic(n, args, variant),
                    avg_ms=avg_ms,
                    per_launch=dict(interactions=inter / max(calls, 1), flops=flops,
                                    cell_per_particle=pst["grav_cell_interactions"] / max(calls, 1) / n,
                                    pp_per_particle=pst["grav_pp_interactions"] / max(calls, 1) / n))
    # HBM-bound stages: algorithmic bytes per particle per launch (SoA compulsory traffic).
    bytes_per = {"sort": 32 + 64, "gather": 2 * 56, "tree": 16, "nbr": 20, "hsolve": 28, "density": 60,
                 "forces": 60, "kick": 104}.get(stage, 64)
    achieved = bytes_per * n / (avg_ms * 1e-3) / 1e9
    return dict(kernel=stage, bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / HBM_PEAK_GBS, traffic=None, avg_ms=avg_ms)


def cpu_baseline(args):
    """The oracle's FP64 restatement of HydroBody::step (hydrostatic.rs:345-359: grid neighbours, octree
    Barnes-Hut theta 0.5 with leaf bucket 8, Courant dt) with the same Newton h-solve (48 neighbours, 1e-4) before
    every density pass, 1 thread, on the same impact geometry at a bounded particle count."""
    from botheadstudios_amd import ics
    from oracle import oracle

    ic = ics.config_ic(4, n=args.cpu_n)
    o = oracle.Hydro(ic.pos, ic.vel, ic.u, ic.h, ic.m, ic.mat.astype(np.int32), ic.eos,
                     ic.softening, theta=0.5, gravity_mode=0, cs_fd=0, brute=False, leaf_cap=8, nthreads=1,
                     h_solve=True, h_eta=ics.ETA_48, h_tol=1e-4)
    t0 = time.perf_counter()
    o.kdk(args.cpu_steps, dt=0.0, cfl=0.1)
    sec = time.perf_counter() - t0
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        pass
    return dict(value=args.cpu_n * args.cpu_steps / sec, unit="particle-steps/s", cores=1, kind="port",
                sample=f"oracle FP64 HydroBody KDK step with the Newton h-solve (octree leaf 8, theta 0.5) on the same impact "
                       f"geometry at N={args.cpu_n}, {args.cpu_steps} steps, {sec:.1f} s, 1 thread "
                       f"({model or platform.processor()}, nproc={os.cpu_count()})")


def measure_leg(args):
    """Disk / remnant measurement after an impact: impact-run `measure` with its Moon candidate
    (ie/tools/impact-run/src/main.rs:732-878) and gpu_sph.rs `disk_stats_json` (gpu_sph.rs:399-469), on a synthetic
    post-impact snapshot of 1,000,000 particles resident in HBM (a remnant of 900,000, a 60,000-body disk, a cold
    3,000-body moonlet, a hot 2,000-body clump, 35,000 escaping), against the oracle's restatement of the
    reference's CPU code on the same snapshot."""
    from botheadstudios_amd import ics
    from botheadstudios_amd import measure as M
    from botheadstudios_amd.gpu_sph import GpuSph
    from botheadstudios_amd.materials import eos_table

    m = 1.0e20 * 30000 / 900000
    p, nt, cls = ics.disk_scene(n_ball=900000, n_disk=60000, n_blob=3000, n_hot=2000, n_esc=35000, seed=9,
                                h=1.2e5, m=m)
    out = {"workload": f"synthetic post-impact snapshot, N={len(p)}, {int((cls == 1).sum())} disk bodies"}
    with GpuSph(len(p), leaf_size=16, h_mode="fixed") as s:
        s.upload(p, eos_table(["granite", "iron"]), 1.0e4)
        for _ in range(2):
            g = M.measure(s, nt)
            d = M.disk_stats(s)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(args.measure_reps):
            g = M.measure(s, nt)
        t1 = time.perf_counter()
        for _ in range(args.measure_reps):
            d = M.disk_stats(s)
        t2 = time.perf_counter()
    out["ms_measure"] = (t1 - t0) * 1e3 / args.measure_reps
    out["ms_disk_stats"] = (t2 - t1) * 1e3 / args.measure_reps
    out["result"] = {"n_disk": g.n_disk, "disk_kg": g.disk_kg, "earth_frac": g.earth_frac, "r_remnant": g.r_remnant,
                     "n_clumps": g.n_clumps, "n_bound": g.n_bound, "clump_n": g.clump_n, "clump_kg": g.clump_kg,
                     "disk_stats_moon": d.moon}
    if not args.no_cpu_baseline:
        from oracle import measure_oracle as mo
        t0 = time.perf_counter()
        o = mo.measure_ir(p, nt)
        sec = time.perf_counter() - t0
        out["matches_oracle"] = all(getattr(g, k) == o[k] for k in ("n_disk", "n_clumps", "n_bound", "clump_n")) and \
            all(abs(getattr(g, k) - o[k]) <= 1e-12 * abs(o[k]) for k in ("disk_kg", "r_remnant", "clump_kg"))
        out["cpu_baseline"] = dict(value=sec * 1e3, unit="ms per measure", cores=1, kind="port",
                                   sample="oracle restatement of impact-run measure + moon_candidate (sorted-Kahan "
                                          "sums, O(k^2) friends-of-friends) on the same snapshot, 1 thread")
        out["speedup_vs_cpu"] = sec * 1e3 / out["ms_measure"]
    return out


def main():
    args = parse()
    if args.measure_only:
        log("disk / remnant measurement ...")
        print(json.dumps({"measure": measure_leg(args)}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        sys.exit(2)
    from botheadstudios_amd import autotune, ics
    from botheadstudios_amd import mlp as mlp_mod

    ic = ics.config_ic(4, n=args.n)
    tuned = autotune.tuned_params(ic.n)
    args.leaf = tuned["leaf"] if args.leaf is None else args.leaf
    args.theta = tuned["theta"] if args.theta is None else args.theta
    args.group = tuned["group"] if args.group is None else args.group
    args.quadrupole = tuned["quadrupole"] if args.quadrupole is None else bool(args.quadrupole)
    log(f"walk parameters: leaf {args.leaf}, group {args.group}, theta {args.theta}, quadrupole {args.quadrupole}")
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    variants = {}
    wanted = args.variants.split(",")
    if "plain" in wanted:
        log("plain: Newton h, FP32 ...")
        variants["plain_fp32"] = run_variant(ic, "fp32", "newton", args)
        log(f"  {variants['plain_fp32']['ms_per_step']:.3f} ms/step")
    weights = mlp_mod.default_weights_path()
    if "learned" in wanted and weights is not None:
        log("learned: MLP-seeded Newton h, split precision ...")
        variants["learned_split"] = run_variant(ic, "split", "learned", args, weights=weights)
        log(f"  {variants['learned_split']['ms_per_step']:.3f} ms/step")
    if "reuse" in wanted:
        log("plain, reusing position-only work ...")
        variants["plain_fp32_reuse"] = run_variant(ic, "fp32", "newton", args, reuse=True)
        log(f"  {variants['plain_fp32_reuse']['ms_per_step']:.3f} ms/step")
        if weights is not None:
            log("learned + split, reusing position-only work ...")
            variants["learned_split_reuse"] = run_variant(ic, "split", "learned", args, weights=weights, reuse=True)
            log(f"  {variants['learned_split_reuse']['ms_per_step']:.3f} ms/step")
    head = "learned_split" if "learned_split" in variants else "plain_fp32"
    res = variants[head]
    line = {
        "metric": METRIC, "value": res["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if head == "plain_fp32" else "fp32+fp64",
        "data": "synthetic (seeded jittered-lattice two-body impact)",
        "config": {"workload": f"two-body giant impact, N={ic.n}, 0.9+0.1 Earth masses, 45 deg, 1.1 v_esc; "
                               f"KDK step = 2 force evaluations; variant={head}",
                   "n": ic.n, "leaf": args.leaf, "group": args.group, "theta": args.theta, "quadrupole": args.quadrupole, "n_ngb": 48, "h_tol": 1e-4, "variant": head,
                   "fused_walk": FUSED_WALK},
        "roofline": roofline(res, ic.n, args, head),
        "variants": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"],
                         "mean_h_evals": v["stats"]["mean_h_evals"], "mean_h_iters": v["stats"]["mean_h_iters"],
                         "mean_neighbours": v["stats"]["mean_neighbours"],
                         "gravity_rms_rel_err": v["gravity_error"]["rms"], "gravity_max_rel_err": v["gravity_error"]["max"],
                         "h_capped": v["stats"]["h_capped"], "h_unconverged": v["stats"]["h_unconverged"],
                         "nbr_candidates_per_particle": v["stats"]["nbr_candidates"] / max(1, v["stats"]["force_evals"]) / ic.n,
                         "walk_nodes_per_group_eval": v["stats"]["walk_nodes"] / max(1, v["stats"]["force_evals"]),
                         "stage_ms_per_step": {s: round(ms / max(args.profile_steps, 1), 4)
                                               for s, (ms, c) in v["profile"].items()}}
                     for k, v in variants.items()},
    }
    if "plain_fp32" in variants and "learned_split" in variants:
        line["learned_vs_plain"] = variants["learned_split"]["value"] / variants["plain_fp32"]["value"]
    if "plain_fp32_reuse" in variants and "learned_split_reuse" in variants:
        line["learned_vs_plain_reuse"] = (variants["learned_split_reuse"]["value"] /
                                          variants["plain_fp32_reuse"]["value"])
    if not args.no_sweep:
        # Batched impact-parameter sweep (64 scenarios, N log-uniform in [4096, 131072]) in one launch per pass,
        # against the same scenarios one at a time. Reported beside the headline, not as it.
        from botheadstudios_amd import sweep as sweep_mod
        try:
            line["sweep"] = sweep_mod.measure(64, args.sweep_steps, 2, log=log)
        except Exception as e:  # the headline line must still print
            line["sweep"] = {"error": repr(e)}
    if not args.no_measure:
        log("disk / remnant measurement ...")
        try:
            line["measure"] = measure_leg(args)
        except Exception as e:  # the headline line must still print
            line["measure"] = {"error": repr(e)}
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
