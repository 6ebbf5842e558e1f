"""GPU: the fused gravity + neighbour walk (ihs_config.fused_walk = 1, the default) against the separate neighbour
build (fused_walk = 0) on the same state.

- Same lists: every particle's list length is equal (both paths test r^2 < max(hs_i^2, hs_j^2) on the same FP32
  positions with the same nested-FMA r^2), with no overflow. With fixed h the searched radii are bit-identical
  too; the Newton modes update them from the solved h, which carries the order-dependent round-off below.
- Same SPH: the lists differ only in order and the passes sum each list in its order, so h, densities, SPH
  accelerations and du/dt agree to FP32 summation round-off (measured on MI355X: density 6e-7 with fixed h,
  2.5e-5 after a Newton solve to 1e-4; accelerations 4e-6 / 1.2e-4 of the mean). Velocities get seeded noise so
  that du/dt and the artificial viscosity take part (the bodies' bulk motion alone gives du/dt = 0).
- Gravity: the fused walk opens every node that passes the neighbour box test, so its sum is a different
  Barnes-Hut sum at the same theta, never less accurate here; both meet 1e-3 RMS against the FP64 direct sum.
- A short adaptive-dt KDK run gives the same time, energy and momentum to round-off.
The separate build's own parity against the oracle is test_gpu_sph.py::test_sph_verify_force_eval[*-separate].
"""
import numpy as np
import pytest

from botheadstudios_amd import ics, mlp
from botheadstudios_amd.gpu_sph import GpuSph

pytestmark = pytest.mark.gpu


def _particles(ic, seed=3):
    p = ic.particles48().copy()
    rng = np.random.default_rng(seed)
    p["vel"] = (p["vel"].astype(np.float64) + rng.normal(scale=300.0, size=(len(p), 3))).astype(np.float32)
    return p


def _eval(ic, parts, fused, **kw):
    with GpuSph(ic.n, fused_walk=fused, **kw) as s:
        if kw.get("h_mode") == "learned":
            s.load_mlp(mlp.default_weights_path())
        s.upload(parts, ic.eos, ic.softening)
        s.force_eval()
        out = {k: s.field(k) for k in ("hs", "ncount", "h", "rho", "asph")}
        out["stats"] = s.stats()
        out["gerr"] = s.gravity_error(n_targets=4096)
    return out


@pytest.mark.parametrize("leaf,precision,h_mode", [(1, "fp32", "fixed"), (1, "split", "newton"),
                                                    (16, "fp32", "newton"), (4, "split", "learned")])
def test_fused_walk_matches_separate_build(leaf, precision, h_mode):
    ic = ics.config_ic(1, n=16384)
    parts = _particles(ic)
    kw = dict(leaf_size=leaf, precision=precision, h_mode=h_mode, theta=0.7, quadrupole=True)
    a = _eval(ic, parts, True, **kw)
    b = _eval(ic, parts, False, **kw)
    for r in (a, b):
        st = r["stats"]
        assert st["nbr_overflow"] == 0 and st["walk_overflow"] == 0 and st["nonfinite"] == 0, st
    assert np.array_equal(a["ncount"], b["ncount"])
    assert a["stats"]["mean_neighbours"] == b["stats"]["mean_neighbours"]
    fixed = h_mode == "fixed"
    if fixed:
        assert np.array_equal(a["hs"], b["hs"]) and np.array_equal(a["h"], b["h"])
    # FP32 sums of ~50 terms in another order: round-off grows with each particle's own sum of |terms|, so the
    # bounds are an RMS ratio and a per-particle bound on |a - b| / (|b| + mean |b|).
    tol_rho, tol_rms, tol_max = (5e-6, 2e-5, 2e-4) if fixed else (1.5e-4, 3e-4, 3e-3)
    assert np.allclose(a["h"], b["h"], rtol=tol_rho, atol=0)
    assert np.allclose(a["rho"], b["rho"], rtol=tol_rho, atol=0)
    sa, sb = a["asph"].astype(np.float64), b["asph"].astype(np.float64)
    assert np.all(np.isfinite(sa)) and np.all(np.isfinite(sb))
    for d, m in ((np.sqrt(((sa[:, :3] - sb[:, :3]) ** 2).sum(1)), np.sqrt((sb[:, :3] ** 2).sum(1))),
                 (np.abs(sa[:, 3] - sb[:, 3]), np.abs(sb[:, 3]))):
        assert m.mean() > 0.0
        rms = np.sqrt((d ** 2).mean() / (m ** 2).mean())
        worst = (d / (m + m.mean())).max()
        print(f"rms {rms:.2e} worst {worst:.2e}")
        assert rms < tol_rms and worst < tol_max, (rms, worst)
    assert a["gerr"]["rms"] < 1e-3 and b["gerr"]["rms"] < 1e-3, (a["gerr"], b["gerr"])
    assert a["gerr"]["rms"] <= 1.02 * b["gerr"]["rms"], (a["gerr"], b["gerr"])


def test_fused_walk_kdk_matches_separate_build():
    ic = ics.config_ic(1, n=16384)
    res = []
    for fused in (True, False):
        with GpuSph(ic.n, leaf_size=1, precision="split", h_mode="newton", fused_walk=fused) as s:
            s.upload(ic.particles48(), ic.eos, ic.softening)
            s.force_eval()
            s.stats(reset=True)
            s.encode_kdk(5)
            res.append(s.stats())
    a, b = res
    assert a["steps"] == b["steps"] == 5
    assert a["nonfinite"] == b["nonfinite"] == 0
    assert a["t"] == pytest.approx(b["t"], rel=1e-4)
    assert a["etot"] == pytest.approx(b["etot"], rel=5e-5)
    assert np.allclose(a["p"], b["p"], rtol=0, atol=1e-6 * ic.m.sum() * ic.meta["v0"])


def test_kept_lists_at_unchanged_positions():
    """theta_settled: an evaluation at positions and h unchanged since the last one keeps that evaluation's neighbour
    lists and walk groups and walks gravity without building lists. On the 262,144-particle impact after the upload's
    solve: the kept-list walk runs (its interactions are counted), its force error against the FP64 direct sum stays
    within the 1e-3 bound, and the density and SPH accelerations equal those of an evaluation that rebuilds the lists
    (the same pairs inside the kernel support; only the order of the list entries may differ)."""
    from botheadstudios_amd import ics

    ic = ics.config_ic(2)
    out = {}
    for ts in (0.0, 0.8):
        with GpuSph(ic.n, leaf_size=1, group_max=64, theta=0.9, quadrupole=True, theta_mono=0.6, quad_tol=0.03,
                    theta_settled=ts, h_mode="newton") as sph:
            sph.upload(ic.particles48(), ic.eos, ic.softening)
            sph.stats(reset=True)
            sph.force_eval()  # positions and h as the upload's last solve left them
            st = sph.stats(reset=True)
            err = sph.gravity_error(4096)
            out[ts] = dict(st=st, err=err, rho=sph.field("rho").copy(), asph=sph.field("asph").copy(),
                           ncount=sph.field("ncount").copy())
    a, b = out[0.0], out[0.8]
    print(f"rebuilt: {a['st']['grav_cell_interactions'] / ic.n:.0f} cells, RMS {a['err']['rms']:.2e}; kept lists: "
          f"{b['st']['grav_cell_interactions'] / ic.n:.0f} cells, RMS {b['err']['rms']:.2e}")
    assert b["st"]["grav_cell_interactions"] > 0.5 * a["st"]["grav_cell_interactions"]
    assert b["st"]["nbr_candidates"] == 0 and a["st"]["nbr_candidates"] > 0
    assert b["err"]["rms"] < 1e-3 and a["err"]["rms"] < 1e-3
    assert not np.array_equal(a["ncount"], b["ncount"])  # the kept lists were searched at the solve's wider radius
    np.testing.assert_allclose(b["rho"], a["rho"], rtol=2e-6)
    scale = np.abs(a["asph"][:, :3]).max(1, keepdims=True)
    assert np.max(np.abs(b["asph"][:, :3] - a["asph"][:, :3]) / scale) < 1e-4
