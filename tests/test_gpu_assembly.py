"""GPU: IC assembly around the relaxation (ihs_body_bulk, ihs_assemble, ihs_relax) against the oracle.

- body_bulk on the device equals the oracle's (gpu_sph.rs:279-316) to 1e-12 relative: clean bodies, strays,
  two bodies in one state, FP32 and split precision, a scenario of a batch.
- In-place assembly equals the oracle's assemble_from_relaxed_n records (gpu_sph.rs:318-339): every field but
  position and velocity bit-exact, positions and velocities within one FP32 ulp (the robust centres differ from
  the oracle's sequential sums in the last FP64 bits), softening and dt equal; a missing placement is an error.
- One relaxation step equals the reference's update v' = (v + (a + w^2 (x, y, 0)) dt) damp, x' = x + v' dt
  (sph_step.wgsl:244-257) applied to the device's own accelerations.
- The whole path: bodies far apart, relaxed without viscosity, measured, placed on the canonical impact in place,
  stepped, measured.
"""
import math

import numpy as np
import pytest

from botheadstudios_amd import assembly, ics
from botheadstudios_amd import measure as M
from botheadstudios_amd.gpu_sph import GpuSph, IhsError
from botheadstudios_amd.materials import eos_table
from oracle import measure_oracle as mo

pytestmark = pytest.mark.gpu

EOS = eos_table(["granite", "iron"])
PM = 1.0e20


def _pair(nt=3000, ni=600, sep=4.0e7):
    bt = assembly.from_ics_body(ics.make_body(nt, PM), PM, 1.0e5, EOS)
    bi = assembly.from_ics_body(ics.make_body(ni, PM), PM, 1.0e5, EOS)
    return assembly.far_apart_pair(bt, bi, sep)


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_body_bulk_matches_oracle(precision):
    parts, soft, _ = _pair()
    R = 1.0e7
    strays = parts[:30].copy()
    strays["pos"] = np.array([0.0, 2.0e8, 0.0], np.float32)  # 1% of the target flung far away
    strays["prov"] = 0
    p = np.concatenate([parts, strays])
    with GpuSph(len(p), leaf_size=16, h_mode="fixed", precision=precision) as s:
        s.upload(p, EOS, soft)
        for prov in (0, 1, 5):
            c, m, r = assembly.body_bulk(s, prov)
            oc, om, orr = mo.body_bulk(p, prov)
            assert np.allclose(c, oc, rtol=0, atol=1e-12 * R), (prov, c, oc)
            assert m == pytest.approx(om, rel=1e-14, abs=0.0) and r == pytest.approx(orr, rel=1e-12, abs=0.0)


def test_body_bulk_in_a_batch():
    a, soft, _ = _pair(2992, 608)
    b, _, _ = _pair(2000, 400, sep=6.0e7)
    with GpuSph(len(a) + len(b), leaf_size=16, h_mode="fixed") as s:
        s.upload_batch([a, b], EOS, [soft, soft])
        for seg, p in ((0, a), (1, b)):
            for prov in (0, 1):
                c, m, r = assembly.body_bulk(s, prov, segment=seg)
                oc, om, orr = mo.body_bulk(p, prov)
                assert np.allclose(c, oc, rtol=0, atol=1e-5) and r == pytest.approx(orr, rel=1e-12)


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_assemble_in_place_matches_oracle(precision):
    parts, soft, _ = _pair()
    pl = [assembly.BodyPlacement((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.5e-4)),
          assembly.BodyPlacement((1.8e7, 5.0e6, 0.0), (-8000.0, 0.0, 0.0), (0.0, 2.0e-4, 0.0))]
    with GpuSph(len(parts), leaf_size=16, h_mode="fixed", precision=precision) as s:
        s.upload(parts, EOS, soft)
        with pytest.raises(IhsError, match="placement"):
            assembly.assemble_from_relaxed_n(s, pl[:1])
        gsoft, gdt = assembly.assemble_from_relaxed_n(s, pl)
        out = s.readback()
        st = s.stats()
    arr = np.array([[*p.offset, *p.vel, *p.spin] for p in pl])
    ref, osoft, odt = mo.assemble_n(parts, arr)
    assert gsoft == osoft and gdt == odt
    assert st["t"] == 0.0 and st["steps"] == 0
    for k in ("h", "u", "mass", "mat", "rho", "prov"):
        assert np.array_equal(out[k], ref[k]), k
    for k in ("pos", "vel"):
        d = _ulps(out[k], ref[k])
        assert d.max() <= 1 and (d == 0).mean() > 0.99, (k, d.max(), (d == 0).mean())


def test_relax_step_is_the_reference_update():
    parts, soft, rdt = _pair(2000, 400)
    rng = np.random.default_rng(2)
    parts["vel"] = rng.normal(scale=10.0, size=(len(parts), 3)).astype(np.float32)
    w, damp = 1.0e-4, 0.98
    with GpuSph(len(parts), leaf_size=16, h_mode="fixed") as s:
        s.upload(parts, EOS, soft)
        s.force_eval()
        a = s.field("asph")[:, :3].astype(np.float32) + s.field("agrav")[:, :3].astype(np.float32)
        x0 = s.readback()
        s.set_omega(w)
        s.set_dt(rdt, damp)
        s.encode_relax(1)
        x1 = s.readback()
    f32 = np.float32
    acf = np.zeros_like(a)
    acf[:, 0] = f32(w) * f32(w) * x0["pos"][:, 0]
    acf[:, 1] = f32(w) * f32(w) * x0["pos"][:, 1]
    v = (x0["vel"] + (a + acf) * f32(rdt)) * f32(damp)
    x = x0["pos"] + v * f32(rdt)
    assert np.allclose(x1["vel"], v, rtol=1e-5, atol=1e-5 * np.abs(v).max())
    assert np.allclose(x1["pos"], x, rtol=1e-6, atol=0)
    assert np.array_equal(x1["u"], x0["u"])  # relaxation is mechanical


def test_relax_assemble_impact_flow():
    parts, soft, rdt = _pair(6000, 1200, sep=6.0e7)
    with GpuSph(len(parts), leaf_size=16, h_mode="newton", precision="split") as s:
        s.upload(parts, EOS, soft)
        assembly.relax(s, 20, rdt, damp=0.99)
        e1 = s.stats()
        assert np.isfinite(e1["ke"]) and e1["nonfinite"] == 0
        gsoft, gdt, geo = assembly.assemble_from_relaxed_with(s, v_esc_multiple=1.1, start_separation=1.5,
                                                              impact_parameter=0.7)
        c1, m1, r1 = assembly.body_bulk(s, 1)
        assert np.allclose(c1, [geo["d0"], geo["b"], 0.0], rtol=0, atol=1e-5 * geo["d0"])
        v = s.readback()["vel"][len(parts) - 1200:].astype(np.float64)
        assert np.allclose(v.mean(0), [-geo["v_esc"], 0, 0], rtol=0, atol=1e-3 * geo["v_esc"])
        s.set_dt(gdt)
        s.encode_kdk(5)
        st = s.stats()
        g = M.measure(s, 6000)
    assert st["nonfinite"] == 0 and st["steps"] == 5 and np.isfinite(st["ke"])
    # the target at rest, the impactor inbound: the momentum the assembly set is kept by the steps
    p = np.asarray(st["p"])
    assert np.allclose(p, [-m1 * geo["v_esc"], 0.0, 0.0], rtol=0, atol=1e-3 * m1 * geo["v_esc"])
    assert g.m_total == pytest.approx(m1 + geo["m_target"], rel=1e-12)
