"""Tillotson parameter sets — the table of ie/data/materials.json `tillotson` blocks.

The numbers are data and must equal the reference catalogue exactly (pinned in eos.rs:352-377;
re-pinned by tests/test_oracle_eos.py). Order of the 10-field block (SphEos, gpu_sph.rs:39-55):
rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta.
"""
from __future__ import annotations

import numpy as np

# ie/data/materials.json:69-82 (granite), 155-168 (basalt), 237-251 (peridotite), 520-534 (iron),
# 1280-1293 (water), 1633-1646 (ice).
TILLOTSON = {
    "granite": (2680.0, 0.5, 1.3, 1.8e10, 1.8e10, 1.6e7, 3.5e6, 1.8e7, 5.0, 5.0),
    "basalt": (2700.0, 0.5, 1.5, 2.67e10, 2.67e10, 4.87e8, 4.72e6, 1.82e7, 5.0, 5.0),
    "peridotite": (3500.0, 0.5, 1.4, 1.31e11, 4.9e10, 5.5e8, 4.5e6, 1.45e7, 5.0, 5.0),
    "iron": (7850.0, 0.5, 1.28, 1.28e11, 1.815e11, 1.425e7, 2.4e6, 8.67e6, 5.0, 5.0),
    "water": (998.0, 0.7, 0.15, 2.18e9, 1.325e10, 7.0e6, 4.19e5, 2.69e6, 10.0, 5.0),
    "ice": (917.0, 0.3, 0.1, 9.47e9, 9.47e9, 1.0e7, 7.73e5, 3.04e6, 10.0, 5.0),
}

FIELDS = ("rho0", "a", "b", "A", "B", "E0", "E_iv", "E_cv", "alpha", "beta")

# Up to 16 distinct materials per upload (gpu_sph.rs:114 MAX_MATERIALS).
MAX_MATERIALS = 16


def eos_table(names) -> np.ndarray:
    """(n_mat, 10) float64 table in the given material order; particle `mat` indexes it."""
    return np.array([TILLOTSON[n] for n in names], dtype=np.float64)


def sph_eos48(names) -> np.ndarray:
    """The 48-byte SphEos records (10 f32 + 2 pad) for the C ABI."""
    t = eos_table(names).astype(np.float32)
    out = np.zeros((t.shape[0], 12), np.float32)
    out[:, :10] = t
    return out
