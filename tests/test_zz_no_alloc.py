"""GPU: the allocation-free force-error probe (ihs_direct_f64_noalloc) against ihs_direct_f64 through the C ABI.

The no-allocation entry point has no counterpart in the reference; it is for callers that cannot allocate device
memory. Its results must equal the allocating probe's bit for bit, call after call, with a smaller second call
reusing the same per-GPU scratch."""
import numpy as np
import pytest

from botheadstudios_amd import ics
from botheadstudios_amd.gpu_sph import GpuSph, IhsError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_noalloc_probe_equals_direct_f64_on_two_calls(precision):
    ic = ics.config_ic(1, n=16384)
    rng = np.random.default_rng(12)
    with GpuSph(ic.n, leaf_size=16, precision=precision, h_mode="newton") as sph:
        sph.upload(ic.particles48(), ic.eos, ic.softening)
        sph.set_dt(0.0)
        sph.encode_kdk(2)
        big = rng.choice(ic.n, size=8192, replace=False).astype(np.uint32)
        small = rng.choice(ic.n, size=300, replace=False).astype(np.uint32)
        a1, b1 = sph.direct_f64(big), sph.direct_f64_noalloc(big)
        sph.encode_kdk(1)  # the state moves between the calls
        a2, b2 = sph.direct_f64(small), sph.direct_f64_noalloc(small)
    assert np.array_equal(a1, b1)
    assert np.array_equal(a2, b2)
    assert np.all(np.isfinite(b1)) and np.abs(b1).min() > 0
    assert not np.array_equal(b1[:300], b2)


def test_noalloc_probe_shares_its_scratch_across_handles_and_rejects_oversize_calls():
    """Two handles on one GPU use the same scratch one after the other; more targets than the scratch holds are
    refused on the host, before any device work."""
    ic1 = ics.config_ic(1, n=4096)
    ic2 = ics.impact_ic(6144, m_impactor=0.2, angle_deg=20.0, seed=5)
    with GpuSph(ic1.n, leaf_size=16) as s1, GpuSph(ic2.n, leaf_size=16, precision="split") as s2:
        s1.upload(ic1.particles48(), ic1.eos, ic1.softening)
        s2.upload(ic2.particles48(), ic2.eos, ic2.softening)
        ids1 = np.arange(0, ic1.n, 3, dtype=np.uint32)
        ids2 = np.arange(0, ic2.n, 5, dtype=np.uint32)
        r1 = s1.direct_f64_noalloc(ids1)
        r2 = s2.direct_f64_noalloc(ids2)
        assert np.array_equal(r1, s1.direct_f64(ids1))
        assert np.array_equal(r2, s2.direct_f64(ids2))
        with pytest.raises(IhsError, match="at most 65536 targets"):
            s1.direct_f64_noalloc(np.zeros(65537, np.uint32))
        assert np.array_equal(s1.direct_f64_noalloc(ids1), r1)  # still usable after the refusal
