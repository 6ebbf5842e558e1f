"""Benchmark: SPH + self-gravity giant-impact KDK steps at 1,048,576 particles on one MI355X.

One "step" is the reference's leapfrog step: two full force evaluations (Morton sort, LBVH, neighbour lists,
smoothing-length solve, density/EOS, SPH forces, Barnes-Hut gravity) with a half-kick + drift between them
and a half-kick after (ie/shaders/sph_step.wgsl:213-237, ie/crates/engine/src/hydrostatic.rs:345-359), with
the Courant dt formed on the device each step. Inputs: the two-body impact (0.9 + 0.1 Earth masses, 45 deg,
1.1 v_esc, 30% iron cores, granite mantles, seeded jittered lattices) at N = 1,048,576, already resident in
HBM when the timed region starts.

Prints one JSON line. `value` is the particle-steps/s of the headline variant (learned-h seed + split
precision when a trained network is available, otherwise the plain Newton FP32 step, named in config). Beside it:
- variants: plain Newton FP32; the learned seed with split precision; the same pair reusing the position-only work
  of a step's first evaluation; a continuity-seeded Newton solve (the network's weights zeroed, so its output is 0
  and the seed is the continuity estimate, at the same tight search radius); the pair at the configuration's
  literal walk setting (monopole, theta 0.5, leaf 16); and the pair timed after contact, when the bodies are
  merging and the smoothing lengths change fastest;
- roofline of the dominant kernel (the Barnes-Hut walk, FP32 VALU-bound) with its algorithmic flops priced on the
  oracle walk's interaction counts at the same theta and leaf on the same tree, and the same for every pass;
- the CPU baseline: the oracle's FP64 restatement of HydroBody::step on the CPU-runnable configuration (N = 16,384,
  built -march=native on this host, 1 thread), with its energy and angular-momentum drift beside a GPU run of the
  same initial conditions.
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
FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak (packed FP32 issue)
FP32_UNPACKED_TFLOPS = 78.6     # the same without packed issue, which this build disables for all device code
FLOPS_PER_INTERACTION = 20.0  # softened monopole / particle-particle interaction (standard count)
FLOPS_PER_OPEN = 10.0         # an opened node's MAC test (algorithmic pricing of a Barnes-Hut walk)
N_TRUE = 47.0                 # pairs inside the support per particle: the h-solve's target (48 with the particle)
WALK_TARGETS = 8192           # sampled targets of the oracle walk that prices the gravity pass
# Monopole + traceless quadrupole cell: 50 flops for the acceleration (see interact_quad in k_gravity.hip:
# 3 sub, 3 FMA r^2, rsqrt, 5 mul for M r^-k, 15 for q.d, 5 for d.q.d, 3 for the radial factor, 12 for the
# accumulation) plus 4 for the potential, which only the second evaluation of a step computes: 52 on average.
FLOPS_PER_QUAD_CELL = 52.0
FUSED_WALK = True  # the gravity walk writes the neighbour lists (ihs_config.fused_walk, the library default)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--leaf", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--theta", type=float, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--group", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--quadrupole", type=int, default=None, help="default: the autotuned value for this N")
    ap.add_argument("--theta-mono", type=float, default=None,
                    help="cells passing the opening test at this angle act as monopoles (0: off); default: the "
                         "autotuned value for this N")
    ap.add_argument("--theta-settled", type=float, default=None,
                    help="opening angle of the walk that keeps the neighbour lists at unchanged positions (0: every "
                         "evaluation rebuilds them); default: the autotuned value for this N")
    ap.add_argument("--quad-tol", type=float, default=None,
                    help="cells whose quadrupole term is below this fraction of their monopole act as monopoles "
                         "(0: off); default: the autotuned value for this N")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--variants", default="plain,learned,reuse,continuity,literal,contact",
                    help="plain: Newton FP32; learned: learned seed + split precision; reuse: both again with the "
                         "first evaluation of each step reusing the previous one's position-only work; continuity: "
                         "the seeded solve with the network zeroed (continuity estimate); literal: plain and learned "
                         "at monopole, theta 0.5, leaf 16; contact: plain and learned timed after contact")
    ap.add_argument("--variant-steps", type=int, default=50, help="timed steps of the variants beside the headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-steps", type=int, default=10, help="100 is the configuration's full run (about 90 s)")
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
            abs(rec.get("theta_mono", 0.0) - args.theta_mono) < 1e-6 and
            abs(rec.get("quad_tol", 0.0) - args.quad_tol) < 1e-9 and
            abs(rec.get("theta_settled", 0.0) - args.theta_settled) < 1e-9 and
            rec.get("variant") == variant and rec.get("fused_walk", False) == FUSED_WALK)
    return rec["bytes_per_launch"] if same else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def contact_time(ic) -> float:
    """Time from the initial conditions until the bodies' centres are one contact distance apart, on the two-point-
    mass orbit the initial conditions were placed on (RK4 of the relative motion)."""
    from botheadstudios_amd.ics import G
    m = ic.meta
    M = float(ic.m.sum())
    r_c = m["r_target"] + m["r_impactor"]
    ell = np.sin(np.radians(m["angle_deg"])) * r_c * m["v_contact"]
    v_t = ell / m["d0"]
    v_r = np.sqrt(max(m["v0"] ** 2 - v_t ** 2, 0.0))
    y = np.array([-m["d0"], 0.0, v_r, v_t])

    def f(y):
        r3 = (y[0] ** 2 + y[1] ** 2) ** 1.5
        return np.array([y[2], y[3], -G * M * y[0] / r3, -G * M * y[1] / r3])

    t, dt = 0.0, 0.5
    while np.hypot(y[0], y[1]) > r_c and t < 1e6:
        k1 = f(y); k2 = f(y + 0.5 * dt * k1); k3 = f(y + 0.5 * dt * k2); k4 = f(y + dt * k3)
        y = y + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        t += dt
    return t


def run_variant(ic, precision, h_mode, args, weights=None, reuse=False, steps=None, walk=None, zero_seed=False,
                until_t=None, keep_state=False):
    """One variant: upload, (optionally) untimed steps until the simulated time reaches until_t, warm-up, `steps`
    timed KDK steps (hipGraph replays), then a profiled pass with per-stage HIP events and the force error."""
    from botheadstudios_amd.gpu_sph import GpuSph

    w = {**dict(leaf=args.leaf, theta=args.theta, group=args.group, quadrupole=args.quadrupole,
                theta_mono=args.theta_mono, quad_tol=args.quad_tol, theta_settled=args.theta_settled),
         **(walk or {})}
    steps = args.steps if steps is None else steps
    sph = GpuSph(ic.n, leaf_size=w["leaf"], precision=precision, h_mode=h_mode, theta=w["theta"],
                 group_max=w["group"], quadrupole=w["quadrupole"], reuse_positions=reuse,
                 theta_mono=w.get("theta_mono", 0.0), quad_tol=w.get("quad_tol", 0.0),
                 theta_settled=w.get("theta_settled", 0.0))
    try:
        if weights is not None:
            sph.load_mlp(weights)
            if zero_seed:  # output 0: the seed is the continuity estimate h_prev exp(div v dt / 3)
                sph.set_mlp_weights(np.zeros(sph.mlp_weights().shape, np.float32))
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)      # adaptive Courant dt on the device
        sph.set_av(1.0, 2.0)
        t_start = None
        if until_t is not None:
            while sph.segment_clock(0)["t"] < until_t:
                sph.encode_kdk(50)
            t_start = sph.segment_clock(0)["t"]
        sph.encode_kdk(args.warmup)
        sph.sync()
        sph.stats(reset=True)
        t0 = time.perf_counter()
        sph.encode_kdk(steps)
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
        # The first evaluation of each step keeps the lists and walks at theta_settled: its own force error, from an
        # evaluation at the same (unchanged) positions.
        gerr_settled = None
        if w.get("theta_settled", 0.0) > 0:
            sph.force_eval()
            gerr_settled = sph.gravity_error(8192)
        for k in ("nbr_overflow", "walk_overflow", "nonfinite"):
            if st[k] or pst[k]:
                raise RuntimeError(f"{precision}/{h_mode}: {k} = {st[k]}/{pst[k]}")
        sec = t1 - t0
        out = dict(sec=sec, steps=steps, ms_per_step=1e3 * sec / steps, value=ic.n * steps / sec, stats=st,
                   profile=prof, profile_stats=pst, gravity_error=gerr, gravity_error_settled=gerr_settled, walk=w, t_start=t_start,
                   t_end=sph.segment_clock(0)["t"])
        if keep_state:
            back = sph.readback()
            out["pos32"] = np.ascontiguousarray(back["pos"], np.float32)
            out["mass32"] = np.ascontiguousarray(back["mass"], np.float32)
        return out
    finally:
        sph.close()


def oracle_walk_work(pos32, mass32, leaf, theta, soft, targets=WALK_TARGETS, seed=3):
    """Algorithmic work of the Barnes-Hut pass: the reference's walk (bh_gravity.wgsl:337-394, monopole, per-particle
    MAC) restated in the oracle, on the same tree (the LBVH is bit-exact with the device's) at the same theta and
    leaf, on sampled targets. flops per particle = 20 (accepted nodes + leaf-particle interactions) + 10 opened."""
    from oracle import oracle
    tree = oracle.Lbvh(pos32, mass32, leaf)
    rng = np.random.default_rng(seed)
    tg = np.sort(rng.choice(tree.n, size=min(targets, tree.n), replace=False))
    c = tree.walk_counts(theta, soft, tg).mean(0)
    return dict(nodes=float(c[0]), pp=float(c[1]), opened=float(c[2]), targets=int(len(tg)),
                flops=FLOPS_PER_INTERACTION * float(c[0] + c[1]) + FLOPS_PER_OPEN * float(c[2]))


def equal_accuracy_walk(pos32, mass32, leaf, soft, ids, ref, target_rms, seed=3):
    """The oracle walk (monopole, per-particle MAC) priced at the opening angle where it reaches the device walk's
    force error: the RMS over the same targets of |a - a_ref| / |a_ref| against the same FP64 direct sums.
    Bisection on theta over [0.05, 1.2] (the error grows with theta)."""
    from oracle import oracle
    tree = oracle.Lbvh(pos32, mass32, leaf)
    inv = np.empty(tree.n, np.int64)
    inv[tree.order] = np.arange(tree.n)
    tg = inv[np.asarray(ids, np.int64)]
    rn = np.linalg.norm(ref, axis=1)

    def at(theta):
        c, a = tree.walk_counts(theta, soft, tg, with_acc=True)
        err = np.linalg.norm(a.astype(np.float64) - ref, axis=1) / np.maximum(rn, 1e-300)
        return float(np.sqrt(np.mean(err ** 2))), c.mean(0)

    lo, hi = 0.05, 1.2
    for _ in range(9):
        mid = 0.5 * (lo + hi)
        if at(mid)[0] <= target_rms:
            lo = mid
        else:
            hi = mid
    rms, c = at(lo)
    return dict(theta=lo, rms=rms, target_rms=target_rms, nodes=float(c[0]), pp=float(c[1]), opened=float(c[2]),
                targets=int(len(tg)),
                flops=FLOPS_PER_INTERACTION * float(c[0] + c[1]) + FLOPS_PER_OPEN * float(c[2]))


def stage_work(stage, res, n, walk_flops):
    """Algorithmic (bytes, flops) per particle per force evaluation of a pass (SoA compulsory traffic; FP32 work
    priced per interaction), S = 56 B of carried state."""
    st = res["profile_stats"]
    k = st["mean_h_evals"] or 1.0
    L_over_N = 1.0 / res["walk"]["leaf"]
    return {"sort": (32 + 64, 0.0), "gather": (2 * 56, 0.0), "tree": (16 + 80 * L_over_N, 30 * L_over_N),
            "nbr": (20, 0.0), "gravity": (24, walk_flops), "hsolve": (28, k * 25 * N_TRUE),
            "density": (28 + 20, 25 * N_TRUE + 300), "forces": (60, 70 * N_TRUE), "kick": (104, 40)}.get(stage,
                                                                                                           (0, 0.0))


def passes(res, n, walk_flops):
    """Every pass of the step against its roofline: time per force evaluation, algorithmic bytes and flops,
    the fraction of HBM and of FP32 VALU peak, and the bound (the larger fraction)."""
    out = {}
    evals = 2.0 * max(res.get("profile_steps", 1), 1)
    for stage, (ms, calls) in res["profile"].items():
        if calls == 0:
            continue
        t = ms / evals * 1e-3
        b, f = stage_work(stage, res, n, walk_flops)
        gbs = b * n / t / 1e9
        tfs = f * n / t / 1e12
        hb, vf = gbs / HBM_PEAK_GBS, tfs / FP32_PEAK_TFLOPS
        out[stage] = dict(ms_per_eval=round(ms / evals, 4), bytes_per_particle=b, flops_per_particle=round(f, 1),
                          gbs=round(gbs, 1), tflops=round(tfs, 2), hbm_frac=round(hb, 4), valu_frac=round(vf, 4),
                          bound="valu" if vf > hb else "hbm", frac=round(max(hb, vf), 4))
    return out


def roofline(res, n, args, variant, walk_work, equal_work=None):
    """The dominant kernel, k_gravity_walk, against the FP32 VALU peak. achieved = the algorithmic flops of one
    launch (oracle-walk interactions at the same theta and leaf, 20 flops per interaction and 10 per opened node, for
    every particle) / the launch's average duration (HIP events on the handle's stream). The walk's own executed
    interactions (group MAC, quadrupole cells at 52 flops, the nodes the fused neighbour test opens) are reported
    beside it as `executed`."""
    prof = res["profile"]
    ms, calls = prof["gravity"]
    avg_ms = ms / max(calls, 1)
    pst = res["profile_stats"]
    alg = walk_work["flops"] * n
    achieved = alg / (avg_ms * 1e-3) / 1e12
    cell_flops = FLOPS_PER_QUAD_CELL if res["walk"]["quadrupole"] else FLOPS_PER_INTERACTION
    exe = (cell_flops * pst["grav_cell_interactions"] + FLOPS_PER_INTERACTION * pst["grav_pp_interactions"]) / \
        max(calls, 1)
    exe_tf = exe / (avg_ms * 1e-3) / 1e12
    return dict(kernel="k_gravity_walk", bound="valu", achieved=achieved, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                frac=achieved / FP32_PEAK_TFLOPS, traffic=measured_traffic(n, args, variant), avg_ms=avg_ms,
                algorithmic=dict(flops_per_launch=alg, per_particle=walk_work,
                                 definition="oracle walk (bh_gravity.wgsl:337-394) on the same tree at the same theta "
                                            "and leaf: 20 flops per accepted node or leaf-particle interaction, 10 per "
                                            "opened node"),
                frac_of_unpacked_peak=achieved / FP32_UNPACKED_TFLOPS,
                equal_accuracy=None if equal_work is None else dict(
                    per_particle=equal_work, tflops=equal_work["flops"] * n / (avg_ms * 1e-3) / 1e12,
                    frac=equal_work["flops"] * n / (avg_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                    definition="the oracle walk priced as above at the theta where it reaches this walk's RMS force "
                               "error on the same targets"),
                executed=dict(flops_per_launch=exe, tflops=exe_tf, frac=exe_tf / FP32_PEAK_TFLOPS,
                              cell_per_particle=pst["grav_cell_interactions"] / max(calls, 1) / n,
                              pp_per_particle=pst["grav_pp_interactions"] / max(calls, 1) / n,
                              definition=f"the walk's own interactions: {cell_flops:.0f} flops per accepted "
                                         "quadrupole cell, 20 per monopole (leaf particle, or cell that acts as a "
                                         "monopole)"))


def _drift(e0, e1):
    """Relative drifts of total energy (KE + IE + PE) and of the angular-momentum z component."""
    E0, E1 = e0[0] + e0[1] + e0[2], e1[0] + e1[1] + e1[2]
    return abs(E1 - E0) / abs(E0), abs(e1[8] - e0[8]) / abs(e0[8])


def cpu_baseline(args):
    """The oracle's FP64 restatement of HydroBody::step (hydrostatic.rs:345-359: grid neighbours, octree
    Barnes-Hut theta 0.5 with leaf bucket 8, Courant dt 0.1) with the same Newton h-solve (48 neighbours, 1e-4)
    before every density pass, on the CPU-runnable configuration (N = 16,384, 0.9 + 0.1 Earth masses, 45 deg,
    1.1 v_esc), built -O3 -march=native on this host and run on 1 thread (the reference is serial). Its energy and
    angular-momentum drift over the run is reported beside a GPU run of the same initial conditions and steps
    (plain Newton FP32, leaf 8, theta 0.5)."""
    from botheadstudios_amd import ics
    from botheadstudios_amd.gpu_sph import GpuSph
    from oracle import oracle

    ic = ics.config_ic(1, n=args.cpu_n)
    o = oracle.Hydro(ic.pos, ic.vel, ic.u, ic.h, ic.m, ic.mat.astype(np.int32), ic.eos,
                     ic.softening, theta=0.5, gravity_mode=0, cs_fd=0, brute=False, leaf_cap=8, nthreads=1,
                     h_solve=True, h_eta=ics.ETA_48, h_tol=1e-4, native=True)
    native = o._L is not oracle.lib()
    e0 = o.energy()
    t0 = time.perf_counter()
    t_sim = o.kdk(args.cpu_steps, dt=0.0, cfl=0.1)
    sec = time.perf_counter() - t0
    e1 = o.energy()
    dE, dL = _drift(e0, e1)
    with GpuSph(ic.n, leaf_size=8, theta=0.5, h_mode="newton", precision="fp32") as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)
        sph.force_eval()
        g0 = sph.stats()
        sph.encode_kdk(args.cpu_steps)
        g1 = sph.stats()
    gv = lambda s_: [s_["ke"], s_["ie"], s_["pe"], *s_["p"], *s_["L"]]
    gE, gL = _drift(gv(g0), gv(g1))
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        pass
    return dict(value=args.cpu_n * args.cpu_steps / sec, unit="particle-steps/s", cores=1, kind="port",
                sample=f"oracle FP64 HydroBody KDK step with the Newton h-solve (octree leaf 8, theta 0.5, Courant 0.1) "
                       f"on configuration 1 (N={args.cpu_n}), {args.cpu_steps} steps to t = {t_sim:.1f} s in {sec:.1f} s, "
                       f"1 thread, {'-march=native' if native else 'portable build'} "
                       f"({model or platform.processor()}, nproc={os.cpu_count()})",
                steps=args.cpu_steps, t_sim=t_sim, energy_drift=dE, angmom_drift=dL,
                gpu_same_ic=dict(steps=g1["steps"], t_sim=g1["t"], energy_drift=gE, angmom_drift=gL,
                                 setting="plain Newton FP32, leaf 8, theta 0.5, quadrupole cells"))


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
    args.theta_mono = tuned["theta_mono"] if args.theta_mono is None else args.theta_mono
    args.quad_tol = tuned["quad_tol"] if args.quad_tol is None else args.quad_tol
    args.theta_settled = tuned["theta_settled"] if args.theta_settled is None else args.theta_settled
    log(f"walk parameters: leaf {args.leaf}, group {args.group}, theta {args.theta}, quadrupole {args.quadrupole}, "
        f"theta_mono {args.theta_mono}, quad_tol {args.quad_tol}, theta_settled {args.theta_settled}")
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    variants = {}
    wanted = args.variants.split(",")
    weights = mlp_mod.default_weights_path()
    vsteps = args.variant_steps

    failed = {}

    def add(name, *a, **kw):
        log(f"{name} ...")
        try:
            variants[name] = run_variant(ic, *a, **kw)
        except Exception as e:  # a variant beside the headline must not stop the line from printing
            if name in ("plain_fp32", "learned_split"):
                raise
            failed[name] = repr(e)
            log(f"  failed: {e!r}")
            return
        variants[name]["profile_steps"] = args.profile_steps
        log(f"  {variants[name]['ms_per_step']:.3f} ms/step")

    if "plain" in wanted:
        add("plain_fp32", "fp32", "newton", args)
    if "learned" in wanted and weights is not None:
        add("learned_split", "split", "learned", args, weights=weights, keep_state=True)
    if "reuse" in wanted:
        add("plain_fp32_reuse", "fp32", "newton", args, reuse=True, steps=vsteps)
        if weights is not None:
            add("learned_split_reuse", "split", "learned", args, weights=weights, reuse=True, steps=vsteps)
    if "continuity" in wanted and weights is not None:
        add("continuity_split", "split", "learned", args, weights=weights, zero_seed=True, steps=vsteps)
    if "literal" in wanted:
        lit = dict(leaf=16, theta=0.5, group=64, quadrupole=False, theta_mono=0.0, quad_tol=0.0, theta_settled=0.0)
        add("plain_fp32_literal", "fp32", "newton", args, walk=lit, steps=vsteps)
        if weights is not None:
            add("learned_split_literal", "split", "learned", args, weights=weights, walk=lit, steps=vsteps)
    if "contact" in wanted:
        tc = contact_time(ic)
        log(f"contact at t = {tc:.0f} s")
        add("plain_fp32_contact", "fp32", "newton", args, until_t=tc, steps=vsteps)
        if weights is not None:
            add("learned_split_contact", "split", "learned", args, weights=weights, until_t=tc, steps=vsteps)
    head = "learned_split" if "learned_split" in variants else "plain_fp32"
    res = variants[head]
    if "pos32" not in res:
        res = variants[head] = run_variant(ic, "fp32", "newton", args, keep_state=True)
        res["profile_steps"] = args.profile_steps
    log("pricing the gravity pass on the oracle walk ...")
    walk_work = oracle_walk_work(res["pos32"], res["mass32"], res["walk"]["leaf"], res["walk"]["theta"],
                                 ic.softening)
    ge = res["gravity_error"]
    log("... and at equal force error ...")
    equal_work = equal_accuracy_walk(res["pos32"], res["mass32"], res["walk"]["leaf"], ic.softening, ge["ids"],
                                     ge["ref"], ge["rms"])
    line = {
        "metric": METRIC, "value": res["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if head == "plain_fp32" else "fp32+fp64",
        "data": "synthetic (seeded jittered-lattice two-body impact)",
        "config": {"workload": f"two-body giant impact, N={ic.n}, 0.9+0.1 Earth masses, 45 deg, 1.1 v_esc; "
                               f"KDK step = 2 force evaluations; variant={head}",
                   "n": ic.n, "leaf": args.leaf, "group": args.group, "theta": args.theta, "quadrupole": args.quadrupole,
                   "theta_mono": args.theta_mono, "quad_tol": args.quad_tol, "theta_settled": args.theta_settled,
                   "n_ngb": 48, "h_tol": 1e-4, "variant": head,
                   "fused_walk": FUSED_WALK},
        "roofline": roofline(res, ic.n, args, head, walk_work, equal_work),
        "passes": passes(res, ic.n, walk_work["flops"]),
        "variants": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"], "steps": v["steps"],
                         "walk": v["walk"], "t_start": v["t_start"], "t_end": v["t_end"],
                         "mean_h_evals": v["stats"]["mean_h_evals"], "mean_h_iters": v["stats"]["mean_h_iters"],
                         "mean_neighbours": v["stats"]["mean_neighbours"],
                         "gravity_rms_rel_err": v["gravity_error"]["rms"],
                         "gravity_rms_ref_norm": v["gravity_error"]["rms_ref_norm"],
                         "gravity_max_rel_err": v["gravity_error"]["max"],
                         "gravity_rms_rel_err_kept_lists": (v["gravity_error_settled"] or {}).get("rms"),
                         "gravity_max_rel_err_kept_lists": (v["gravity_error_settled"] or {}).get("max"),
                         "h_capped": v["stats"]["h_capped"], "h_unconverged": v["stats"]["h_unconverged"],
                         "nbr_candidates_per_particle": v["stats"]["nbr_candidates"] / max(1, v["stats"]["force_evals"]) / ic.n,
                         "walk_nodes_per_group_eval": v["stats"]["walk_nodes"] / max(1, v["stats"]["force_evals"]),
                         "stage_ms_per_step": {s: round(ms / max(args.profile_steps, 1), 4)
                                               for s, (ms, c) in v["profile"].items()}}
                     for k, v in variants.items()},
    }
    if failed:
        line["failed_variants"] = failed
    ratios = {"learned_vs_plain": ("learned_split", "plain_fp32"),
              "learned_vs_plain_reuse": ("learned_split_reuse", "plain_fp32_reuse"),
              "learned_vs_plain_literal": ("learned_split_literal", "plain_fp32_literal"),
              "learned_vs_plain_contact": ("learned_split_contact", "plain_fp32_contact"),
              "continuity_vs_plain": ("continuity_split", "plain_fp32"),
              "learned_vs_continuity": ("learned_split", "continuity_split")}
    for key, (a, b) in ratios.items():
        if a in variants and b in variants:
            line[key] = variants[a]["value"] / variants[b]["value"]
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
