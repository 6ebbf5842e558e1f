"""GPU: the two-phase readback (gpu_store.rs:126-152) as the browser's dynamics loop uses it (lib.rs:3151-3166:
step, take, begin, every frame): take hands over the state as it was at begin, however many steps were queued
after it; it reports "not ready" while idle or pending and never blocks."""
import time

import numpy as np
import pytest

from botheadstudios_amd import ics
from botheadstudios_amd.gpu_sph import GpuSph

pytestmark = pytest.mark.gpu


def _handle(ic):
    sph = GpuSph(ic.n, leaf_size=16, h_mode="newton")
    sph.upload(ic.particles48(), ic.eos, ic.softening)
    sph.set_dt(0.0)
    return sph


def _take(sph, timeout_s=60.0):
    t0 = time.monotonic()
    while time.monotonic() - t0 < timeout_s:
        snap = sph.take_readback()
        if snap is not None:
            return snap
        time.sleep(0.001)
    raise AssertionError("the snapshot never completed")


def test_take_returns_the_state_at_begin_despite_later_steps():
    ic = ics.config_ic(1, n=8192)
    with _handle(ic) as a, _handle(ic) as b:
        assert a.take_readback() is None  # idle: nothing begun
        a.encode_kdk(5)
        a.begin_readback()
        a.begin_readback()  # in flight: a no-op, as in the reference
        a.encode_kdk(7)  # queued behind the snapshot
        snap = _take(a)
        assert a.take_readback() is None  # handed over once
        b.encode_kdk(5)
        twin = b.readback()
        now = a.readback()
    assert snap.tobytes() == twin.tobytes()
    assert not np.array_equal(now["pos"], snap["pos"])


def test_frame_loop_sees_every_completed_snapshot():
    """step, take, begin per frame: each snapshot taken equals the blocking readback of a twin run at the step count
    where it was begun."""
    ic = ics.config_ic(1, n=4096)
    with _handle(ic) as a, _handle(ic) as b:
        taken = []
        begun_at = None
        steps = 0
        for frame in range(6):
            a.encode_kdk(2)
            steps += 2
            a.sync()  # let the previous frame's copy complete, so that every frame hands one over
            snap = a.take_readback()
            if begun_at is not None:
                assert snap is not None
                taken.append((begun_at, snap))
            a.begin_readback()
            begun_at = steps
        assert len(taken) == 5
        done = 0
        for at, snap in taken:
            b.encode_kdk(at - done)
            done = at
            assert snap.tobytes() == b.readback().tobytes()
