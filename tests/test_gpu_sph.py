"""GPU SPH parity against the FP64 oracle: the reference's sph-verify known-answer configuration
(ie/tools/sph-verify/src/main.rs:260-396) and the Newton smoothing-length solve."""
import numpy as np
import pytest

from botheadstudios_amd import ics
from botheadstudios_amd.gpu_sph import GpuSph
from oracle import oracle
from tests.helpers import rms_rel, sph_verify_config

pytestmark = pytest.mark.gpu

SOFT = 1.0e4  # sph-verify:293


def _oracle_state(parts, eos, gravity_mode=1):
    return oracle.Hydro(parts["pos"].astype(np.float64), parts["vel"].astype(np.float64), parts["u"].astype(np.float64),
                        parts["h"].astype(np.float64), parts["mass"].astype(np.float64), parts["mat"].astype(np.int32),
                        eos, SOFT, gravity_mode=gravity_mode, cs_fd=1, brute=True)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "separate"])
@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_sph_verify_force_eval(precision, fused):
    """sph-verify:304-343: density max rel, acc RMS < 3e-2, du/dt RMS < 5e-2 (measured 5.6e-7 / 1.9e-6 /
    4.4e-6 on the reference's GPU); here bounded at 1e-5 / 1e-4 / 1e-4. Both neighbour paths: the lists written
    by the gravity walk (fused) and by the separate neighbour build."""
    parts, eos = sph_verify_config(300)
    with GpuSph(512, leaf_size=16, precision=precision, h_mode="fixed", theta=1e-6, fused_walk=fused) as sph:
        sph.upload(parts, eos, SOFT)
        sph.force_eval()
        rho = sph.field("rho")
        asph = sph.field("asph")
        ag = sph.field("agrav")
    o = _oracle_state(parts, eos)
    rho_ref = o.density()
    acc_ref, dudt_ref = o.forces(rho_ref)
    d_rel = np.max(np.abs(rho - rho_ref) / np.maximum(rho_ref, 1.0))
    acc = asph[:, :3].astype(np.float64) + ag[:, :3]
    a_rms, a_max = rms_rel(acc, acc_ref)
    u_rms, _ = rms_rel(asph[:, 3], dudt_ref)
    assert np.all(np.isfinite(acc))
    assert d_rel < 1e-5, d_rel
    assert a_rms < 1e-4, (a_rms, a_max)
    assert u_rms < 1e-4, u_rms


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_sph_verify_kdk_50_steps(precision):
    """sph-verify:356-396: 50 KDK steps at dt = 0.5 s vs the f64 CpuState; pos/vel/u RMS < 5e-3
    (reference measured 3.1e-4 / 5.7e-4 / 5.1e-4)."""
    parts, eos = sph_verify_config(300)
    with GpuSph(512, leaf_size=16, precision=precision, h_mode="fixed", theta=1e-6) as sph:
        sph.upload(parts, eos, SOFT)
        sph.set_dt(0.5)
        sph.encode_kdk(50)
        out = sph.readback()
        pos64 = sph.field("pos64")
    o = _oracle_state(parts, eos)
    o.kdk(50, dt=0.5)
    p0 = parts["pos"].astype(np.float64)
    disp = o.xyz - p0
    pos_rms = np.sqrt(((pos64 - o.xyz) ** 2).sum() / (disp ** 2).sum())
    vel_rms, _ = rms_rel(out["vel"], o.vel)
    u_rms, _ = rms_rel(out["u"], o.u)
    assert pos_rms < 5e-3 and vel_rms < 5e-3 and u_rms < 5e-3, (pos_rms, vel_rms, u_rms)


def test_newton_h_residual_and_neighbours():
    """Newton h-solve (new; no reference counterpart): every converged h meets |f|/rho <= 1e-4 when the
    residual is re-evaluated in FP64 by the oracle; the bulk averages ~48 neighbours within h."""
    ic = ics.config_ic(1, n=16384)
    with GpuSph(ic.n, leaf_size=16, h_mode="newton") as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.force_eval()
        h = sph.field("h").astype(np.float64)
        st = sph.stats()
        pos = sph.field("pos64")
    assert st["h_unconverged"] == 0 and st["nbr_overflow"] == 0 and st["walk_overflow"] == 0
    assert st["h_capped"] == 0, st
    resid, nngb = oracle.hresidual(pos, ic.m, h, ics.ETA_48)
    assert np.max(resid) <= 1.2e-4, np.max(resid)
    r = np.sqrt((pos ** 2).sum(1))
    bulk = nngb[(r < 0.5 * ic.meta["r_target"])]
    assert abs(bulk.mean() - 47.0) < 3.0, bulk.mean()  # 48 incl. self -> ~47 others


def test_support_masks_do_not_change_the_forces():
    """The density pass marks each list's entries inside the kernel support and the force pass fetches only those.
    Lists too long for the masks (nbr_capacity > 256) take the unmasked path, which visits every entry and skips
    the same ones by distance: the SPH accelerations and du/dt agree bit for bit."""
    ic = ics.config_ic(1, n=16384)
    out = []
    for cap in (192, 320):
        with GpuSph(ic.n, leaf_size=1, h_mode="newton", nbr_capacity=cap) as sph:
            sph.upload(ic.particles48(), ic.eos, ic.softening)
            sph.set_dt(5.0)
            sph.encode_kdk(3)
            sph.force_eval()
            out.append(sph.field("asph"))
    assert np.array_equal(out[0], out[1])


def test_newton_matches_oracle_solution():
    """GPU Newton h equals the oracle's FP64 safeguarded Newton root to the tolerance-implied precision."""
    ic = ics.config_ic(1, n=4096)
    with GpuSph(ic.n, leaf_size=16, h_mode="newton") as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.force_eval()
        h_gpu = sph.field("h").astype(np.float64)
        pos = sph.field("pos64")
    h_ref, it, res, bad = oracle.hsolve(pos, ic.m, ic.particles48()["h"].astype(np.float64), ics.ETA_48, tol=1e-10,
                                        h_max=4.0 * ic.h.max())
    assert bad == 0
    rel = np.abs(h_gpu - h_ref) / h_ref
    assert np.max(rel) < 1e-4, np.max(rel)


def test_kdk_conservation_impact():
    """20 adaptive-dt KDK steps of the impact IC: momentum conserved to round-off of the FP64 reductions,
    energy drift well under the reference's 0.3-0.5% over a whole run, no overflow or non-finite values."""
    ic = ics.config_ic(1)
    with GpuSph(ic.n, leaf_size=16, h_mode="newton", precision="split") as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.force_eval()
        s0 = sph.stats(reset=True)
        sph.encode_kdk(20)
        s1 = sph.stats()
    for k in ("nbr_overflow", "walk_overflow", "h_unconverged", "nonfinite"):
        assert s1[k] == 0, (k, s1)
    e0 = s0["etot"]
    assert abs(s1["etot"] - e0) / abs(e0) < 2e-3, (s0, s1)
    pscale = ic.m.sum() * ic.meta["v0"]
    assert np.linalg.norm(np.array(s1["p"]) - np.array(s0["p"])) / pscale < 1e-5
    assert s1["steps"] == 20 and s1["dt"] > 0


@pytest.mark.parametrize("mode", ["newton", "learned"])
def test_reuse_positions_same_trajectory(mode):
    """reuse_positions: the first evaluation of a step keeps the previous evaluation's tree, lists, h, density and
    gravity (same positions) and recomputes only the EOS and the velocity-dependent forces. The trajectory is the
    one of the full recomputation up to the h-solve tolerance and FP32 summation order. Steps replay a captured
    graph, whose buffer parity differs from the full step's (one state swap per step instead of two)."""
    from botheadstudios_amd import mlp
    ic = ics.config_ic(1, n=16384)
    w = mlp.default_weights_path()
    if mode == "learned" and w is None:
        pytest.skip("no trained weights")
    out = {}
    for reuse in (False, True):
        with GpuSph(ic.n, leaf_size=1, h_mode=mode, reuse_positions=reuse) as sph:
            if mode == "learned":
                sph.load_mlp(w)
            sph.upload(ic.particles48(), ic.eos, ic.softening)
            sph.set_dt(5.0)
            sph.encode_kdk(20)
            st = sph.stats()
            out[reuse] = (sph.field("pos64"), sph.readback(), st)
    (p0, r0, s0), (p1, r1, s1) = out[False], out[True]
    for s_ in (s0, s1):
        assert s_["nonfinite"] == 0 and s_["h_unconverged"] == 0 and s_["h_capped"] == 0, s_
    disp = np.linalg.norm(p0 - ic.pos, axis=1)
    assert np.sqrt(np.mean(np.sum((p1 - p0) ** 2, 1))) / np.sqrt(np.mean(disp ** 2)) < 1e-4
    u0 = r0["u"].astype(np.float64)
    # The full recomputation also re-solves h at unchanged positions, moving it within the 1e-4 tolerance: a few
    # particles at the shock front carry that into u at the 1e-4 level, the bulk to ~1e-5 (learned seeds, whose
    # first Newton step is often the last, leave h anywhere inside the tolerance).
    du = np.abs(r1["u"] - u0)
    assert np.sqrt(np.mean(du ** 2)) < 1e-4 * u0.mean() and du.max() < 2e-3 * u0.mean()
    assert abs(s1["etot"] - s0["etot"]) < 1e-5 * abs(s0["etot"])
