"""GPU: the device Courant dt and the FP64 reductions against the oracle and FP64 sums of the device's own state.

- dt: hydrostatic.rs:361-374 courant_dt, with the WGSL sound speed of cs_signal (sph_step.wgsl:259-269) as the
  impact-run loop reads it (impact-run:345-356, 403-410): dt = cfl min_i h_i / max(c_i + |v_i|, 1), cfl = 0.1.
- KE, IE, PE, momentum and angular momentum: gpu_sph.rs:624-649 total_energy, impact-run:452-472 and
  impact-run:528-536 (PE here is 1/2 sum m phi over the walk's potential; the reference sums pairs directly).
"""
import math

import numpy as np
import pytest

from botheadstudios_amd import ics
from botheadstudios_amd.gpu_sph import GpuSph
from oracle import oracle

pytestmark = pytest.mark.gpu

G = 6.674e-11
CFL = 0.1


def _device_signal_dt(back, h, rpc):
    """The definition on the device's own sound speed, in the FP32 arithmetic the density pass uses."""
    v = back["vel"].astype(np.float32)
    speed = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    sig = h.astype(np.float32) / np.maximum(rpc[:, 2] + speed, np.float32(1.0))
    return CFL * float(sig.min())


def _oracle_dt(back, h, rho, eos, softening):
    o = oracle.Hydro(back["pos"].astype(np.float64), back["vel"].astype(np.float64), back["u"].astype(np.float64),
                     h.astype(np.float64), back["mass"].astype(np.float64), back["mat"].astype(np.int32), eos,
                     softening, cs_fd=1)
    return o.courant_dt(CFL, rho=rho.astype(np.float64))


@pytest.mark.parametrize("h_mode", ["fixed", "newton"])
@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_device_courant_dt_matches_oracle(precision, h_mode):
    """The step's dt equals cfl min h / max(c + |v|, 1) on the state its first evaluation saw: to FP32 rounding
    against the device's own c, and to the FP32 sound speed's finite-difference error (2e-4) against the oracle's
    FP64 c on the device's density and u."""
    ic = ics.config_ic(1, n=8192)
    with GpuSph(ic.n, leaf_size=16, precision=precision, h_mode=h_mode) as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)
        sph.force_eval()  # the evaluation the step's first evaluation repeats at the same positions
        back = sph.readback()
        h = sph.field("h")
        rpc = sph.field("rpc")
        sph.encode_kdk(1)
        dt = sph.segment_clock(0)["dt"]
    assert dt > 0
    assert dt == pytest.approx(_device_signal_dt(back, h, rpc), rel=1e-6 if h_mode == "fixed" else 1e-5)
    assert dt == pytest.approx(_oracle_dt(back, h, rpc[:, 0], ic.eos, ic.softening), rel=2e-4)


def test_device_courant_dt_per_scenario_of_a_batch():
    """Each scenario of a batch steps with its own dt from its own particles."""
    ics_ = [ics.config_ic(1, n=n) for n in (3072, 5120, 8192)]
    ics_[1] = ics.impact_ic(5120, m_impactor=0.2, angle_deg=10.0, v_over_vesc=1.4, seed=77)
    counts = [ic.n for ic in ics_]
    start = np.concatenate([[0], np.cumsum(counts)])
    with GpuSph(sum(counts), leaf_size=16, h_mode="fixed") as sph:
        sph.upload_batch([ic.particles48() for ic in ics_], ics_[0].eos, [ic.softening for ic in ics_])
        sph.set_dt(0.0)
        sph.force_eval()
        back = sph.readback()
        h = sph.field("h")
        rpc = sph.field("rpc")
        sph.encode_kdk(1)
        dts = [sph.segment_clock(s)["dt"] for s in range(3)]
    assert len(set(dts)) == 3
    for s, ic in enumerate(ics_):
        sl = slice(start[s], start[s + 1])
        assert dts[s] == pytest.approx(_device_signal_dt(back[sl], h[sl], rpc[sl]), rel=1e-6)
        assert dts[s] == pytest.approx(_oracle_dt(back[sl], h[sl], rpc[sl, 0], ic.eos, ic.softening), rel=2e-4)


def _fp64_sums(m, pos, vel, u):
    ke = 0.5 * float((m * (vel * vel).sum(1)).sum())
    ie = float((m * u).sum())
    p = (m[:, None] * vel).sum(0)
    L = (m[:, None] * np.cross(pos, vel)).sum(0)
    return ke, ie, p, L


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_energy_momentum_angular_momentum_equal_fp64_sums_of_the_state(precision):
    """ihs_stats KE, IE, p and L equal FP64 sums over the device's own state (positions in FP64 as the device
    holds them); PE = 1/2 sum m phi equals the oracle's FP64 direct-sum potential to the walk's tolerance."""
    ic = ics.config_ic(1, n=8192)
    with GpuSph(ic.n, leaf_size=16, precision=precision, h_mode="newton", theta=0.5) as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)
        sph.encode_kdk(3)
        sph.force_eval()  # the potential of the current positions
        st = sph.stats()
        seg = sph.segment_energy(0)
        back = sph.readback()
        pos = sph.field("pos64")
    m = back["mass"].astype(np.float64)
    vel = back["vel"].astype(np.float64)
    u = back["u"].astype(np.float64)
    ke, ie, p, L = _fp64_sums(m, pos, vel, u)
    assert st["ke"] == pytest.approx(ke, rel=1e-12)
    # Split precision keeps u in FP64; the readback rounds it to FP32.
    assert st["ie"] == pytest.approx(ie, rel=1e-12 if precision == "fp32" else 1e-7)
    scale_p = float((m * np.linalg.norm(vel, axis=1)).sum())
    scale_L = float((m * np.linalg.norm(pos, axis=1) * np.linalg.norm(vel, axis=1)).sum())
    assert np.abs(np.array(st["p"]) - p).max() <= 1e-12 * scale_p
    assert np.abs(np.array(st["L"]) - L).max() <= 1e-12 * scale_L
    # L about the origin: a centre-of-mass-frame IC with a 45 degree impact has L along +z, nonzero.
    assert abs(L[2]) > 1e3 * max(abs(L[0]), abs(L[1]))
    _, phi = oracle.direct_f64(pos, m, ic.softening, with_pot=True)
    pe_ref = 0.5 * float((m * phi).sum())
    assert st["pe"] < 0
    assert st["pe"] == pytest.approx(pe_ref, rel=1e-3)
    # One scenario's sums are the handle's totals.
    assert (seg["ke"], seg["ie"], seg["pe"]) == (st["ke"], st["ie"], st["pe"])
    assert seg["L"] == st["L"] and seg["p"] == st["p"]


def test_batch_energy_per_scenario_and_totals():
    """ihs_segment_energy sums one scenario's particles; ihs_stats_get's energies are the totals over all."""
    ics_ = [ics.config_ic(1, n=n) for n in (3072, 8192)]
    counts = [ic.n for ic in ics_]
    start = np.concatenate([[0], np.cumsum(counts)])
    with GpuSph(sum(counts), leaf_size=16, h_mode="newton") as sph:
        sph.upload_batch([ic.particles48() for ic in ics_], ics_[0].eos, [ic.softening for ic in ics_])
        sph.set_dt(0.0)
        sph.encode_kdk(2)
        sph.force_eval()
        st = sph.stats()
        segs = [sph.segment_energy(s) for s in range(2)]
        back = sph.readback()
        pos = sph.field("pos64")
    for s in range(2):
        sl = slice(start[s], start[s + 1])
        ke, ie, p, L = _fp64_sums(back["mass"][sl].astype(np.float64), pos[sl], back["vel"][sl].astype(np.float64),
                                  back["u"][sl].astype(np.float64))
        assert segs[s]["ke"] == pytest.approx(ke, rel=1e-12)
        assert segs[s]["ie"] == pytest.approx(ie, rel=1e-12)
    for k in ("ke", "ie", "pe"):
        assert st[k] == pytest.approx(segs[0][k] + segs[1][k], rel=1e-12)


def test_potential_energy_of_a_uniform_sphere():
    """gpu_sph.rs:624-649's PE on a uniform random sphere: -3 G M^2 / 5 R to the sampling noise (2%), and the
    oracle's FP64 direct sum over the same bodies to the walk's tolerance."""
    rng = np.random.default_rng(3)
    n, R, M = 32768, 1.0e6, 1.0e22
    pts = rng.uniform(-1.0, 1.0, (3 * n, 3))
    pts = pts[(pts * pts).sum(1) <= 1.0][:n] * R
    assert len(pts) == n
    p = np.zeros(n, ics.SPH_PARTICLE_DTYPE)
    p["pos"] = pts
    p["h"] = 0.05 * R
    p["u"] = 1.0e5
    p["mass"] = M / n
    p["rho"] = 2680.0
    eos = ics.eos_table(["granite"])
    soft = 1.0e-3 * R
    with GpuSph(n, leaf_size=16, h_mode="fixed", theta=0.5) as sph:
        sph.upload(p, eos, soft)
        sph.force_eval()
        pe = sph.stats()["pe"]
        pos = sph.field("pos64")
    assert pe / (-3.0 * G * M * M / (5.0 * R)) == pytest.approx(1.0, abs=0.02)
    _, phi = oracle.direct_f64(pos, np.full(n, M / n), soft, with_pot=True)
    assert pe == pytest.approx(0.5 * (M / n) * float(phi.sum()), rel=1e-3)
    assert math.isfinite(pe)


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_energy_read_after_steps_forms_the_potential(precision):
    """KDK steps sum no potential; an energy read after them forms it with one gravity walk over the last tree. Its PE
    equals the oracle's FP64 direct-sum potential of the same positions to the walk's tolerance, and the PE an explicit
    force evaluation at those positions sums (the fused walk's potential); the read leaves the trajectory untouched: a
    twin that does not read the energy between steps ends in the same state, bit for bit."""
    ic = ics.config_ic(1, n=8192)

    def run(mode):
        with GpuSph(ic.n, leaf_size=1, precision=precision, h_mode="newton", theta=0.9) as sph:
            sph.upload(ic.particles48(), ic.eos, ic.softening)
            sph.set_dt(0.0)
            sph.encode_kdk(3)
            if mode == "force_eval":
                sph.force_eval()  # the same positions: the fused walk sums the potential
                return sph.stats(), None, None
            st = sph.stats() if mode == "read" else None
            pos = sph.field("pos64")
            sph.encode_kdk(2)
            return st, pos, sph.readback()

    st, pos, back = run("read")
    _, _, back_twin = run("no read")
    st_fe, _, _ = run("force_eval")
    m = ic.m.astype(np.float32).astype(np.float64)
    _, phi = oracle.direct_f64(pos, m, float(np.float32(ic.softening)), with_pot=True)
    pe_ref = 0.5 * float((m * phi).sum())
    assert st["pe"] < 0
    assert st["pe"] == pytest.approx(pe_ref, rel=1e-3)
    assert st["pe"] == pytest.approx(st_fe["pe"], rel=1e-3)
    assert (st["ke"], st["ie"]) == (st_fe["ke"], st_fe["ie"])
    for k in back.dtype.names:
        a, b = np.ascontiguousarray(back[k]), np.ascontiguousarray(back_twin[k])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
