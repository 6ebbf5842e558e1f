"""GPU: checkpoint / resume. A run resumed from a checkpoint continues bit for bit as the run that wrote it: every
48-byte record, the FP64 positions of split precision, the FP64 energies and the scenario clocks, for the plain
Newton FP32 step, the learned seed with split precision, and a batch of two scenarios. A checkpoint written under
another configuration, or truncated, is refused."""
import numpy as np
import pytest

from botheadstudios_amd import checkpoint, ics
from botheadstudios_amd import mlp as mlp_mod
from botheadstudios_amd.gpu_sph import GpuSph, IhsError
from tests.helpers import checkpoint_sections as _sections

pytestmark = pytest.mark.gpu


def _state(s, split, nseg):
    out = {"rec": s.readback().view(np.uint8).copy(), "clk": [s.segment_clock(k) for k in range(nseg)]}
    if split:
        out["pos64"] = s.field("pos64").copy()
    st = s.stats()
    out["energy"] = (st["ke"], st["ie"], st["pe"], tuple(st["p"]), tuple(st["L"]))
    return out


@pytest.mark.parametrize("variant", ["fp32_newton", "split_learned", "batch"])
def test_resume_is_bit_exact(variant, tmp_path):
    ic = ics.config_ic(1, n=8192)
    split = variant == "split_learned"
    kw = dict(leaf_size=1, h_mode="learned" if split else "newton", precision="split" if split else "fp32",
              theta=0.85, quadrupole=True)
    path = str(tmp_path / "run.ihsckpt")
    nseg = 2 if variant == "batch" else 1
    print(f"checkpoint variant={variant}", flush=True)
    with GpuSph(ic.n * nseg, **kw) as a:
        if split:
            a.load_mlp(mlp_mod.default_weights_path())
        if nseg == 2:
            b2 = ics.config_ic(1, n=8192)
            b2.pos[:, 0] *= 1.01
            a.upload_batch([ic.particles48(), b2.particles48()], ic.eos, [ic.softening, ic.softening])
        else:
            a.upload(ic.particles48(), ic.eos, ic.softening)
        a.set_dt(0.0)
        a.encode_kdk(3)
        nbytes = checkpoint.save(a, path)
        a.encode_kdk(4)
        ref = _state(a, split, nseg)
    hd = checkpoint.header(open(path, "rb").read())
    assert hd["n"] == ic.n * nseg and hd["scenarios"] == nseg and hd["total_bytes"] == nbytes and hd["split"] == split
    with checkpoint.load(path) as b:
        b.encode_kdk(4)
        got = _state(b, split, nseg)
    assert np.array_equal(got["rec"], ref["rec"])
    if split:
        assert np.array_equal(got["pos64"], ref["pos64"])
    assert got["clk"] == ref["clk"]
    assert got["energy"] == ref["energy"]


def test_checkpoint_refusals(tmp_path):
    ic = ics.config_ic(1, n=4096)
    with GpuSph(ic.n, leaf_size=16) as a:
        with pytest.raises(IhsError, match="no particles"):
            checkpoint.to_bytes(a)
        a.upload(ic.particles48(), ic.eos, ic.softening)
        data = checkpoint.to_bytes(a)
    with GpuSph(ic.n, leaf_size=8) as b:
        buf = np.frombuffer(data, np.uint8)
        with pytest.raises(IhsError, match="configuration"):
            b._check(b._L.ihs_checkpoint_load(b._h, buf.ctypes.data, len(data)))
    with pytest.raises(IhsError, match="size|truncated"):
        checkpoint.from_bytes(data[:-100])
    with pytest.raises(ValueError, match="not a checkpoint"):
        checkpoint.from_bytes(b"x" * 200)


@pytest.mark.parametrize("corruption", ["id_duplicate", "id_range", "mat", "seg_start", "h"])
def test_checkpoint_corrupt_sections_are_refused_on_the_host(corruption):
    """Sections the kernels index with are checked before anything reaches the device: the ids must be a permutation,
    materials inside the uploaded EOS table, the scenario table must span the particles, h positive. The refused
    handle takes a fresh upload afterwards."""
    ic = ics.config_ic(1, n=4096)
    with GpuSph(ic.n, leaf_size=16) as a:
        a.upload(ic.particles48(), ic.eos, ic.softening)
        a.encode_kdk(1)
        data = bytearray(checkpoint.to_bytes(a))
    at, n = _sections(bytes(data))
    if corruption == "id_duplicate":
        data[at["id"] + 4 * 5:at["id"] + 4 * 6] = data[at["id"] + 4 * 6:at["id"] + 4 * 7]
    elif corruption == "id_range":
        data[at["id"]:at["id"] + 4] = np.uint32(n + 7).tobytes()
    elif corruption == "mat":
        data[at["mat"] + 3] = len(ic.eos)  # one past the uploaded table
    elif corruption == "seg_start":
        data[at["seg_start"] + 4:at["seg_start"] + 8] = np.uint32(n - 1).tobytes()
    else:
        data[at["h"] + 4 * 9:at["h"] + 4 * 10] = np.float32(-1.0).tobytes()
    with GpuSph(ic.n, leaf_size=16) as b:
        buf = np.frombuffer(bytes(data), np.uint8)
        with pytest.raises(IhsError, match="corrupt checkpoint"):
            b._check(b._L.ihs_checkpoint_load(b._h, buf.ctypes.data, len(data)))
        b.upload(ic.particles48(), ic.eos, ic.softening)
        b.encode_kdk(1)
        assert b.stats()["nonfinite"] == 0
