"""CPU: the measurement oracle (oracle/measure_oracle.cpp) against closed forms, independent libraries and the
reference's own accretion tests (accretion.rs:216-296). No GPU.

What each check pins:
  - sum_oi: math.fsum (the correctly rounded sum) without cancellation, the compensated-summation error bound
    with it; order independence.
  - perigee: Kepler closed forms for both variants (main.rs:232-247 and orbit.rs:96-116).
  - measure / engine frame: a scene with known classes (tests/helpers.py disk_scene); the category masses
    partition m_total; r_remnant is the ceil(0.85 n)-th radius for equal masses; the Moon candidate is the cold blob.
  - friends-of-friends: components equal scipy's connected components of the brute-force link graph, for the
    O(k^2) moon_candidate loop and for find_clumps' grid (above 512 bodies) alike.
  - clump energies: numpy pair sums (scipy pdist), kinetic energy about the mass-weighted velocity.
  - accretion.rs's tests: two cold blobs accrete and conserve mass, momentum and centre of mass; the Roche gate;
    a hot group is one spatial clump but unbound.
"""
import math

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial.distance import pdist

from oracle import measure_oracle as mo
from tests.helpers import G, cold_blob, disk_scene


def _components(pos, link, f32_diff=False):
    """Labels (lowest member index per component) of the brute-force link graph d2 <= link^2."""
    n = pos.shape[0]
    if f32_diff:
        p = pos.astype(np.float32)
        d = (p[:, None, :] - p[None, :, :]).astype(np.float64)
    else:
        d = pos[:, None, :] - pos[None, :, :]
    d2 = (d * d).sum(-1)
    i, j = np.nonzero(np.triu(d2 <= link * link, 1))
    g = coo_matrix((np.ones(i.size), (i, j)), shape=(n, n))
    _, lab = connected_components(g, directed=False)
    first = np.full(lab.max() + 1, n)
    np.minimum.at(first, lab, np.arange(n))
    return first[lab]


# ---- sum_oi ----

@pytest.mark.parametrize("kind", ["uniform", "wide", "cancel"])
def test_sum_oi_is_correctly_rounded_and_order_free(kind):
    rng = np.random.default_rng(7)
    if kind == "uniform":
        t = rng.random(20000) * 1e20
    elif kind == "wide":
        t = rng.standard_normal(20000) * 10.0 ** rng.integers(-10, 25, 20000)
    else:
        a = rng.standard_normal(10000) * 1e22
        t = np.concatenate([a, -a * (1 + 1e-12), rng.standard_normal(100)])
    exact = math.fsum(t)
    s = mo.sum_oi(t)
    if kind == "uniform":
        assert abs(s - exact) <= math.ulp(exact)  # no cancellation: correctly rounded
    # compensated summation's bound (2u + O(n u^2)) sum |t|; sorting does not make it exact under cancellation
    assert abs(s - exact) <= 2.5 * 2.0 ** -53 * math.fsum(np.abs(t))
    for seed in range(3):
        assert mo.sum_oi(np.random.default_rng(seed).permutation(t)) == s


# ---- perigee ----

MU = G * 6.0e24


def test_perigee_kepler_ellipse():
    # Apoapsis of an ellipse with a = 2e7, e = 0.3: r_a = a (1 + e), v_a = sqrt(mu (1 - e) / r_a) tangential.
    a, e = 2.0e7, 0.3
    ra = a * (1 + e)
    va = math.sqrt(MU * (1 - e) / ra)
    for which in ("impact_run", "orbit"):
        q = mo.perigee(which, [ra, 0.0, 0.0], [0.0, va, 0.0], MU)
        assert q == pytest.approx(a * (1 - e), rel=1e-12)
    # circular: q = r; inclined plane does not matter
    r = 1.5e7
    v = math.sqrt(MU / r)
    for which in ("impact_run", "orbit"):
        assert mo.perigee(which, [0.0, r, 0.0], [0.0, 0.0, v], MU) == pytest.approx(r, rel=1e-9)


def test_perigee_unbound_and_degenerate():
    r = 1.0e7
    vesc = math.sqrt(2 * MU / r)
    # hyperbola, tangential at r with v = 1.5 vesc: q = r (periapsis), e = v^2 r / mu - 1
    v = 1.5 * vesc
    assert mo.perigee("impact_run", [r, 0, 0], [0, v, 0], MU) is None
    q = mo.perigee("orbit", [r, 0, 0], [0, v, 0], MU)
    assert q == pytest.approx(r, rel=1e-12)
    # radial escape: no periapsis in either variant
    assert mo.perigee("impact_run", [r, 0, 0], [2 * vesc, 0, 0], MU) is None
    assert mo.perigee("orbit", [r, 0, 0], [2 * vesc, 0, 0], MU) is None
    # radial fall while bound: e = 1, q = 0
    assert mo.perigee("impact_run", [r, 0, 0], [-0.1 * vesc, 0, 0], MU) == pytest.approx(0.0, abs=1e-6 * r)
    # at the centre
    assert mo.perigee("impact_run", [0, 0, 0], [1, 2, 3], MU) == 0.0
    assert mo.perigee("orbit", [0, 0, 0], [1, 2, 3], MU) == 0.0


# ---- measure (impact-run) ----

def test_measure_scene_classes_partition_and_candidate():
    p, nt, cls = disk_scene()
    d, c = mo.measure_ir(p, nt, with_classes=True)
    assert np.array_equal(c, cls)
    m = p["mass"].astype(np.float64)
    assert d["m_total"] == math.fsum(m)
    parts = [d[k] for k in ("e_disk", "t_disk", "e_rem", "t_rem", "e_esc", "t_esc")]
    assert math.fsum(parts) == pytest.approx(d["m_total"], rel=1e-15)
    # equal masses: the ceil(0.85 n)-th smallest radius about the (mass-weighted) centre
    pos = p["pos"].astype(np.float64)
    com = (pos * m[:, None]).sum(0) / m.sum()
    r = np.sort(np.sqrt(((pos - com) ** 2).sum(1)))
    k = math.ceil(0.85 * len(p))
    assert d["r_remnant"] == pytest.approx(r[k - 1], rel=1e-12)
    assert d["m_remnant"] == pytest.approx(k * float(np.float32(1e20)), rel=1e-15)
    # disk split by index < n_target
    disk = cls == 1
    idx = np.arange(len(p))
    assert d["e_disk"] == pytest.approx(m[disk & (idx < nt)].sum(), rel=1e-14)
    assert d["earth_frac"] == pytest.approx(100 * d["e_disk"] / (d["e_disk"] + d["t_disk"]), rel=1e-14)
    # the Moon candidate contains the whole cold blob (indices 32000..32299) and is target-free
    assert d["clump_n"] >= 300 and d["clump_earth_frac"] < 1.0
    assert d["n_bound"] == d["n_clumps"] - 1  # the hot blob


def test_moon_candidate_components_and_energies():
    """The FoF partition of the disk equals scipy's components with the impact-run FP32 differences; the candidate's
    mass, membership and potential follow from it."""
    p, nt, cls = disk_scene(n_ball=30000, n_disk=1500, seed=3)
    d = mo.measure_ir(p, nt)
    disk = np.nonzero(cls == 1)[0]
    pos = p["pos"][disk].astype(np.float64)
    link = 2.0 * p["h"][disk].astype(np.float64).mean()
    assert d["link"] == pytest.approx(link, rel=1e-14)
    lab = _components(pos, link, f32_diff=True)
    sizes = np.bincount(np.unique(lab, return_inverse=True)[1])
    assert d["n_clumps"] == int((sizes >= 2).sum())
    # candidate = the component of the blob's first body; its potential by pdist
    blob0 = np.nonzero(disk == 30000 + 1500)[0][0]
    members = np.nonzero(lab == lab[blob0])[0]
    assert d["clump_n"] == members.size
    mem = disk[members]
    mm = p["mass"][mem].astype(np.float64)
    assert d["clump_kg"] == pytest.approx(mm.sum(), rel=1e-14)


def test_measure_small_and_empty_disks():
    p, nt, cls = disk_scene(n_ball=4000, n_disk=0, n_blob=0, n_hot=0, n_esc=50)
    d = mo.measure_ir(p, nt)
    assert d["n_disk"] == 0 and d["disk_kg"] == 0.0 and d["earth_frac"] == 0.0 and d["clump_kg"] == 0.0
    assert d["e_esc"] + d["t_esc"] == pytest.approx(50 * float(np.float32(1e20)), rel=1e-14)
    one = disk_scene(n_ball=4000, n_disk=2, n_blob=0, n_hot=0, n_esc=0)[0][:4001]
    d1 = mo.measure_ir(one, 4000)
    assert d1["n_disk"] == 1 and d1["n_clumps"] == 0 and d1["clump_kg"] == 0.0


# ---- engine frame (gpu_sph.rs) ----

def test_engine_frame_bound_disk_is_subset():
    p, nt, cls = disk_scene(seed=5)
    fb, cb = mo.engine_frame(p, True, with_classes=True)
    fa, ca = mo.engine_frame(p, False, with_classes=True)
    assert np.array_equal(cb, cls)
    assert np.all(ca[cb == 1] == 1)  # every bound disk body is also in the any-orbit disk
    m = p["mass"].astype(np.float64)
    assert fb["e_disk"] + fb["t_disk"] + fb["esc"] <= fb["m_total"] * (1 + 1e-15)
    assert fb["esc"] == pytest.approx(m[cls == 2].sum(), rel=1e-14)
    assert fb["link"] == pytest.approx(2 * p["h"].astype(np.float64).mean(), rel=1e-14)
    # hyperbolic passes: radially escaping bodies have no periapsis, so they stay out of the any-orbit disk too
    assert np.all(ca[cls == 2] != 1)


# ---- find_clumps (accretion.rs) ----

def _accrete(pos, vel, mass, rho, link, central, cm, cr):
    clumps, lab = mo.find_clumps(pos, vel, mass, rho, link, G, 1.0e4, central, cm, cr)
    return clumps, lab


def test_accretion_conserves_mass_momentum_and_com():
    """accretion.rs:225-263."""
    m_i, rho = 1.0e19, 3000.0
    p1, v1, m1, r1 = cold_blob((2.0e7, 0, 0), (0, 1500.0, 0), 3.0e5, 40, m_i, rho)
    p2, v2, m2, r2 = cold_blob((-1.5e7, 1.0e7, 0), (-800.0, 0, 300.0), 2.5e5, 30, m_i, rho)
    ps = np.array([[4.0e7 + k * 5.0e6, -3.0e7, 0.0] for k in range(5)])
    vs = np.array([[0.0, 0.0, 500.0 * k] for k in range(5)])
    pos = np.concatenate([p1, p2, ps])
    vel = np.concatenate([v1, v2, vs])
    mass = np.concatenate([m1, m2, np.full(5, m_i)])
    rr = np.concatenate([r1, r2, np.full(5, rho)])
    clumps, lab = _accrete(pos, vel, mass, rr, 5.0e5, (0, 0, 0), 5.0e24, 6.0e6)
    acc = [c for c in clumps if c["n_members"] >= 2 and c["bound"] and c["outside_roche"]]
    assert len(acc) == 2
    consumed = np.concatenate([np.nonzero(lab == clumps.index(c))[0] for c in acc])
    keep = np.setdiff1d(np.arange(len(pos)), consumed)
    fm = np.concatenate([[c["mass"] for c in acc], mass[keep]])
    fp = np.concatenate([[c["com_pos"] for c in acc], pos[keep]])
    fv = np.concatenate([[c["com_vel"] for c in acc], vel[keep]])
    m0, m1_ = mass.sum(), fm.sum()
    assert abs(m1_ - m0) / m0 < 1e-12
    mom0, mom1 = (vel * mass[:, None]).sum(0), (fv * fm[:, None]).sum(0)
    assert np.linalg.norm(mom1 - mom0) / np.linalg.norm(mom0) < 1e-12
    com0, com1 = (pos * mass[:, None]).sum(0) / m0, (fp * fm[:, None]).sum(0) / m1_
    assert np.linalg.norm(com1 - com0) / np.linalg.norm(com0) < 1e-12
    assert len(fm) == 2 + 5


def test_roche_gate():
    """accretion.rs:266-281."""
    m_planet, r_planet, rho_clump = 5.0e24, 6.0e6, 3000.0
    d_roche = 2.44 * r_planet * (m_planet / (4.0 / 3.0 * math.pi * r_planet ** 3) / rho_clump) ** (1 / 3)
    for d, want in ((0.6 * d_roche, 0), (2.0 * d_roche, 1)):
        p, v, m, r = cold_blob((d, 0, 0), (0, 0, 0), 2.0e5, 30, 1.0e19, rho_clump)
        clumps, _ = _accrete(p, v, m, r, 5.0e5, (0, 0, 0), m_planet, r_planet)
        assert sum(1 for c in clumps if c["n_members"] >= 2 and c["bound"] and c["outside_roche"]) == want


def test_hot_group_is_unbound():
    """accretion.rs:284-296."""
    p, v, m, r = cold_blob((3.0e7, 0, 0), (0, 0, 0), 2.0e5, 30, 1.0e15, 3000.0)
    s = np.where(np.arange(30) % 2 == 0, 1.0, -1.0)
    v = np.stack([s * 5000.0, s * -4000.0, s * 3000.0], axis=1)
    clumps, _ = _accrete(p, v, m, r, 5.0e5, (0, 0, 0), 5.0e24, 6.0e6)
    big = max(clumps, key=lambda c: c["n_members"])
    assert big["n_members"] >= 25 and not big["bound"]


@pytest.mark.parametrize("n,link", [(300, 4.0e5), (3000, 2.5e5), (3000, 0.0)])
def test_find_clumps_components_and_energies(n, link):
    """Grid (n >= 512) and brute-force pair sets give scipy's components; per-clump sums and the softened pair
    potential match numpy."""
    rng = np.random.default_rng(n)
    pos = rng.random((n, 3)) * 5.0e6
    pos[: n // 10] = pos[0]  # coincident bodies link even at zero linking length
    vel = rng.normal(scale=100.0, size=(n, 3))
    mass = rng.random(n) * 1e19 + 1e18
    rho = rng.random(n) * 2000 + 1000
    clumps, lab = mo.find_clumps(pos, vel, mass, rho, link, G, 1.0e4, (0, 0, 0), 5e24, 6e6)
    ref = _components(pos, link)
    # same partition, clumps ordered by lowest member
    firsts = np.array([c["first"] for c in clumps])
    assert np.array_equal(firsts, np.unique(ref))
    assert np.array_equal(firsts[lab], ref)
    big = max(clumps, key=lambda c: c["n_members"])
    mem = np.nonzero(firsts[lab] == big["first"])[0]
    mm = mass[mem]
    assert big["mass"] == pytest.approx(mm.sum(), rel=1e-13)
    cv = (vel[mem] * mm[:, None]).sum(0) / mm.sum()
    assert big["internal_ke"] == pytest.approx((0.5 * mm * ((vel[mem] - cv) ** 2).sum(1)).sum(), rel=1e-10)
    r = np.sqrt(pdist(pos[mem]) ** 2 + 1.0e8)
    mi, mj = np.triu_indices(mem.size, 1)
    assert big["self_pe"] == pytest.approx(-(G * mm[mi] * mm[mj] / r).sum(), rel=1e-10)
    assert big["rho"] == pytest.approx(mm.sum() / (mm / rho[mem]).sum(), rel=1e-13)
