"""Checkpoint / resume on disk: the device state as a versioned binary image (ihs_checkpoint_* in the C ABI).

The reference keeps its snapshots in memory (impact-run:394-434 reads one back at each physical-time checkpoint;
`VIZ_OUT` writes a JSON dump, hydrostatic.rs:1164-1218) and has no restart. Here `save` writes the whole device
state (scenario tables and clocks, step control, EOS table, network weights, FP64 positions in split precision)
and `load` creates a handle from the configuration stored in the file and restores it, so that a resumed run
continues bit for bit (with reuse_positions off).

    checkpoint.save(sph, "impact.ihsckpt")
    sph2 = checkpoint.load("impact.ihsckpt")
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .gpu_sph import GpuSph

MAGIC = b"IHSCKPT1"
CONFIG_OFFSET = 40  # magic 8, version 4, header size 4, n 8, total size 8, scenarios 4, precision 4


def to_bytes(sph: GpuSph) -> bytes:
    n = sph._check(sph._L.ihs_checkpoint_bytes(sph._h))
    buf = np.empty(n, np.uint8)
    sph._check(sph._L.ihs_checkpoint_save(sph._h, buf.ctypes.data, n))
    return buf.tobytes()


def save(sph: GpuSph, path: str) -> int:
    data = to_bytes(sph)
    with open(path, "wb") as f:
        f.write(data)
    return len(data)


def header(data: bytes) -> dict:
    if data[:8] != MAGIC:
        raise ValueError("not a checkpoint")
    version, hbytes = np.frombuffer(data, np.uint32, 2, 8)
    n, total = np.frombuffer(data, np.int64, 2, 16)
    nseg, split = np.frombuffer(data, np.int32, 2, 32)
    cfg = _lib.IhsConfig.from_buffer_copy(data, CONFIG_OFFSET)
    return {"version": int(version), "header_bytes": int(hbytes), "n": int(n), "total_bytes": int(total),
            "scenarios": int(nseg), "split": bool(split), "config": cfg}


def from_bytes(data: bytes, device: int = 0, capacity: int | None = None) -> GpuSph:
    hd = header(data)
    cfg = hd["config"]
    cfg.capacity = int(max(hd["n"], capacity or cfg.capacity))
    sph = GpuSph.from_config(cfg, device)
    buf = np.frombuffer(data, np.uint8)
    sph._check(sph._L.ihs_checkpoint_load(sph._h, buf.ctypes.data, len(data)))
    return sph


def load(path: str, device: int = 0, capacity: int | None = None) -> GpuSph:
    with open(path, "rb") as f:
        data = f.read()
    return from_bytes(data, device, capacity)
