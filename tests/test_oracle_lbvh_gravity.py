"""CPU: the oracle's Morton/sort/Karras/upsweep/walk restatement and FP64 gravity, pinned against the numbers the
reference records (docs/37:14-23, bhtree.rs:205-249, gpu-bh-verify:461-569) and against brute force."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import bh_cloud, rms_rel, splitmix_stream
from tests.test_gpu_lbvh import verify_tree

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_published.json")))
FIX = os.path.join(os.path.dirname(__file__), "golden", "lbvh_cloud1500.npz")


@pytest.fixture(scope="module")
def cloud():
    pos, mass = bh_cloud(1500)
    return pos, mass, oracle.direct_f64(pos, mass, 5.0e3)


def test_expand_bits_known_values():
    assert [oracle.expand_bits(v) for v in (0, 1, 2, 3, 1023)] == [0, 1, 8, 9, 0x09249249]
    for v in range(0, 1024, 37):
        bits = [(v >> b) & 1 for b in range(10)]
        assert oracle.expand_bits(v) == sum(bit << (3 * b) for b, bit in enumerate(bits))


def test_morton_is_the_bit_interleave_and_bbox_exact():
    rng = np.random.default_rng(3)
    xyz = (rng.random((5000, 3)) * 1e6 - 3e5).astype(np.float32)
    lo, hi = oracle.bbox_f32(xyz)
    assert np.array_equal(lo, xyz.min(0)) and np.array_equal(hi, xyz.max(0))
    codes = oracle.morton_f32(xyz, lo, hi)
    ext = np.maximum(hi - lo, np.float32(1e-30))
    u = np.clip((xyz - lo) / ext, 0, 1).astype(np.float32)
    q = np.clip(np.floor(u * np.float32(1024)), 0, 1023).astype(np.uint32)
    ref = np.zeros(len(xyz), np.uint32)
    for b in range(10):
        for ax, sh in ((0, 2), (1, 1), (2, 0)):
            ref |= ((q[:, ax] >> b) & 1) << (3 * b + sh)
    assert np.array_equal(codes, ref)
    xyz[7] = xyz[3]
    c2 = oracle.morton_f32(xyz, lo, hi)
    assert c2[7] == c2[3]


def test_sort_is_lexicographic_code_index():
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 50, 4000).astype(np.uint32)  # many ties
    sc, order = oracle.sort_pairs(codes)
    ref = np.lexsort((np.arange(4000), codes))
    assert np.array_equal(order, ref) and np.array_equal(sc, codes[ref])


def _lcp(a, b, i, j):
    return 32 + (32 - int(i ^ j).bit_length()) if a == b else 32 - int(a ^ b).bit_length()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_karras_tree_valid_and_split_is_the_prefix_change(seed):
    rng = np.random.default_rng(seed)
    cc = np.sort(rng.integers(0, 1 << 30, 700).astype(np.uint32))
    left, right, parent = oracle.karras(cc)
    L = len(cc)
    assert verify_tree(left, right, parent, L)
    # Each internal node's children split its leaf range where the common prefix with the first leaf drops.
    rng_of = {}

    def span(x):
        if x >= L - 1:
            return (x - (L - 1),) * 2
        return rng_of[x]
    order = []
    st = [0]
    while st:
        x = st.pop()
        order.append(x)
        st += [c for c in (left[x], right[x]) if c < L - 1]
    for x in reversed(order):
        a, b = span(left[x]), span(right[x])
        assert a[1] + 1 == b[0]
        rng_of[x] = (a[0], b[1])
        f, l = rng_of[x]
        full = _lcp(cc[f], cc[l], f, l)
        assert _lcp(cc[f], cc[a[1]], f, a[1]) > full
        assert _lcp(cc[f], cc[b[0]], f, b[0]) <= full


def test_karras_duplicate_runs():
    """Runs of equal cluster codes: the index-augmented split always yields a valid tree; the reference's midpoint
    split (bh_gravity.wgsl:159-160) agrees for runs of two and breaks for longer runs."""
    cc = np.array([5, 9, 9, 12, 12, 12, 12, 12, 12, 12, 12, 12, 40, 41, 41, 99], np.uint32)
    lf, rf, pf = oracle.karras(cc, fix_dups=True)
    assert verify_tree(lf, rf, pf, len(cc))
    l0, r0, p0 = oracle.karras(cc, fix_dups=False)
    with pytest.raises(AssertionError):
        assert verify_tree(l0, r0, p0, len(cc))
    pairs = np.array([1, 4, 4, 7, 9, 9, 30, 31, 31, 50], np.uint32)
    a, b = oracle.karras(pairs, True), oracle.karras(pairs, False)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_upsweep_root_and_bbox(cloud):
    pos, mass, _ = cloud
    lb = oracle.Lbvh(pos.astype(np.float32), mass.astype(np.float32), 1)
    m = mass.astype(np.float32).astype(np.float64)
    com = (pos.astype(np.float32).astype(np.float64) * m[:, None]).sum(0) / m.sum()
    b = GOLD["upsweep_root"]["bounds"]
    assert abs(lb.com4[0, 3] - m.sum()) / m.sum() < b["mass_rel"]
    assert np.linalg.norm(lb.com4[0, :3] - com) / max(np.linalg.norm(com), 1.0) < b["com_rel"]
    assert np.array_equal(lb.bmin[0], lb.lo) and np.array_equal(lb.bmax[0], lb.hi)


def test_walk_reproduces_the_reference_accuracy_table(cloud):
    """The restated cs_gravity_bh walk on the reference's own N=1500 cloud reproduces its published accuracy:
    theta=0.5 RMS 0.70% (max 6.2%), 0.25 -> 0.12%, 0.1 -> 0.009%, theta->0 -> 1.8e-6."""
    pos, mass, ref = cloud
    lb = oracle.Lbvh(pos.astype(np.float32), mass.astype(np.float32), 1)
    g = GOLD["bh_walk_n1500"]
    for th, want in g["rms"].items():
        rms, worst = rms_rel(lb.walk(float(th), 5.0e3), ref)
        assert rms == pytest.approx(want, rel=0.08), (th, rms)
    _, worst = rms_rel(lb.walk(0.5, 5.0e3), ref)
    assert worst == pytest.approx(g["max"]["0.5"], rel=0.05)


def test_octree_bounds(cloud):
    """bhtree.rs:205-249: theta=0.5 RMS < 1%, max < 10%; theta->0 max rel < 1e-9."""
    pos, mass, ref = cloud
    b = GOLD["octree_n1500"]["bounds"]
    rms, worst = rms_rel(oracle.octree_f64(pos, mass, 0.5, 5.0e3), ref)
    assert rms < b["rms"] and worst < b["max"]
    _, worst0 = rms_rel(oracle.octree_f64(pos, mass, 1e-6, 5.0e3), ref)
    assert worst0 < b["theta_0_max"]


def test_direct_closed_forms():
    G = oracle.G
    pos = np.array([[0.0, 0, 0], [3.0e6, 4.0e6, 0]])
    m = np.array([2e20, 5e20])
    soft = 1e5
    acc, pot = oracle.direct_f64(pos, m, soft, with_pot=True)
    r2 = 25e12 + soft ** 2
    assert acc[0] == pytest.approx(G * m[1] * pos[1] / r2 ** 1.5, rel=1e-14)
    assert pot[0] == pytest.approx(-G * m[1] / math.sqrt(r2), rel=1e-14)
    # Momentum conservation of the pairwise sum.
    p2, m2 = bh_cloud(400, seed=9)
    a = oracle.direct_f64(p2, m2, 5e3)
    assert np.linalg.norm((m2[:, None] * a).sum(0)) / np.abs(m2[:, None] * a).sum() < 1e-13


def test_uniform_sphere_potential_energy():
    """PE of a uniform sphere -> -3GM^2/(5R) (softening small against the spacing)."""
    n_side = 24
    ax = (np.arange(n_side) + 0.5) / n_side * 2 - 1
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    p = p[(p ** 2).sum(1) <= 1.0] * 1e6
    m = np.full(len(p), 1e18)
    _, pot = oracle.direct_f64(p, m, 1e3, with_pot=True)
    pe = 0.5 * (m * pot).sum()
    R = (3 * len(p) / (4 * math.pi)) ** (1 / 3) * (2e6 / n_side)
    assert pe == pytest.approx(-0.6 * oracle.G * m.sum() ** 2 / R, rel=0.01)


def test_grid_pairs_equal_brute_force():
    """neighbors.rs:103-145: the grid yields exactly the brute-force pair set within each cell size."""
    r = splitmix_stream(0x12345678, 3 * 800).reshape(800, 3) * 100.0
    for cell in (3.0, 7.5, 20.0, 50.0):
        got = {tuple(p) for p in oracle.grid_pairs(r, cell, radius=cell)}
        d = np.sqrt(((r[:, None] - r[None]) ** 2).sum(-1))
        i, j = np.where(np.triu(d < cell, 1))
        assert got == set(zip(i.tolist(), j.tolist()))


def test_frozen_lbvh_fixture(cloud):
    """Regression pin: the oracle's tree for the N=1500 cloud equals the committed snapshot."""
    pos, mass, _ = cloud
    z = np.load(FIX)
    for k in (1, 16):
        lb = oracle.Lbvh(pos.astype(np.float32), mass.astype(np.float32), k)
        assert np.array_equal(lb.sorted_codes, z["sorted_codes"])
        assert np.array_equal(lb.order, z["order"])
        assert np.array_equal(lb.left, z[f"left_k{k}"]) and np.array_equal(lb.right, z[f"right_k{k}"])
        assert np.array_equal(lb.parent, z[f"parent_k{k}"])


def test_walk_counts_are_consistent():
    """The oracle walk's counts (accepted nodes, leaf-particle interactions, opened nodes) on the 1500-cloud with
    one-particle leaves: theta -> 0 opens every internal node and meets every other particle; a larger theta
    opens fewer nodes and interacts less; each target's work is accounted (every opened node pushes two children)."""
    pos, mass = bh_cloud(1500)
    t = oracle.Lbvh(pos.astype(np.float32), mass.astype(np.float32), 1)
    tg = np.arange(0, 1500, 7)
    c0 = t.walk_counts(1e-6, 1.0e3, tg)
    assert np.all(c0[:, 0] == 0) and np.all(c0[:, 1] == 1499) and np.all(c0[:, 2] == 1499)
    c5 = t.walk_counts(0.5, 1.0e3, tg)
    c9 = t.walk_counts(0.9, 1.0e3, tg)
    assert c9.sum(0)[2] < c5.sum(0)[2] < c0.sum(0)[2]
    assert (c9[:, 0] + c9[:, 1]).sum() < (c5[:, 0] + c5[:, 1]).sum()
    # Every node visited is either accepted, a leaf (one particle; the target's own leaf is skipped) or opened:
    # visits = 1 (root) + 2 * opened.
    for c in (c5, c9):
        own = 1  # the target's own one-particle leaf contributes no interaction
        assert np.all(c[:, 0] + c[:, 1] + own + c[:, 2] == 1 + 2 * c[:, 2])


def test_monopole_cell_thresholds():
    """The walk lets an accepted quadrupole cell act as a monopole on two tests, both resting on facts checked here
    in FP64 on random cells:
    (1) the quadrupole part of a cell's acceleration, differentiated numerically from the expansion's potential
        phi = -M [1/r + (d.q.d) / (2 r^5)] (q traceless, mass-normalised; d = com - x), is at most 3.5 |q|_F M / r^4,
        a fraction 3.5 |q|_F / r^2 of the monopole's M / r^2 (so d^2 >= 3.5 |q|_F / tol keeps it below tol); the
        largest ratio met is about 0.35 (an eigenvector of q: 3 |lambda| against 3.5 sqrt(6) |lambda|);
    (2) a critical radius rc = b / theta + delta scaled by theta / theta_mono is at least b / theta_mono + delta, so
        the test d >= rc theta / theta_mono is the opening test at theta_mono or a stricter one."""
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(400):
        a = rng.normal(size=(3, 3)) * rng.uniform(0.1, 10.0)
        q = a + a.T
        q -= np.trace(q) / 3.0 * np.eye(3)
        qf = np.sqrt((q * q).sum())
        d = rng.normal(size=3) * rng.uniform(5.0, 50.0)
        M = rng.uniform(0.5, 2.0)

        def phi(x, quad):
            dd = d - x
            r = np.linalg.norm(dd)
            return -M * (1.0 / r + (quad * (dd @ q @ dd) / (2.0 * r ** 5)))

        def grad(quad):
            h = 1e-4 * np.linalg.norm(d)
            return np.array([(phi(h * e, quad) - phi(-h * e, quad)) / (2 * h) for e in np.eye(3)])

        a_quad = -(grad(1.0) - grad(0.0))  # a = -grad_x phi, quadrupole part only
        r = np.linalg.norm(d)
        bound = 3.5 * qf * M / r ** 4
        ratio = np.linalg.norm(a_quad) / bound
        assert ratio <= 1.0 + 1e-6, ratio
        worst = max(worst, ratio)
    assert worst > 0.3, worst
    for b, delta, th, tm in rng.uniform([1.0, 0.0, 0.5, 0.1], [10.0, 5.0, 1.2, 0.5], size=(1000, 4)):
        rc = b / th + delta
        assert rc * th / tm >= b / tm + delta - 1e-12
