"""Tillotson parameter sets — the table of ie/data/materials.json `tillotson` blocks.

The numbers are data and must equal the reference catalogue exactly (pinned in eos.rs:352-377;
re-pinned by tests/test_oracle_eos_kernel.py). Order of the 10-field block (SphEos, gpu_sph.rs:39-55):
rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta.
"""
from __future__ import annotations

import math

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


def tillotson_pressure(eos10, rho: float, u: float) -> float:
    """Host FP64 Tillotson pressure (ie/crates/engine/src/eos.rs:52-85), for set-up arithmetic such as the
    relaxation time step. eos10 = rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta."""
    rho0, a, b, cap_a, cap_b, e0, e_iv, e_cv, alpha, beta = (float(x) for x in eos10[:10])
    rho = max(rho, 1.0e-9)
    eta = rho / rho0
    mu = eta - 1.0
    omega = u / (e0 * eta * eta) + 1.0
    p_c = (a + b / omega) * rho * u + cap_a * mu + cap_b * mu * mu
    if eta >= 1.0 or u <= e_iv:
        return p_c
    z = rho0 / rho - 1.0
    p_e = a * rho * u + (b * rho * u / omega + cap_a * mu * math.exp(-beta * z)) * math.exp(-alpha * z * z)
    if u >= e_cv:
        return p_e
    return ((u - e_iv) * p_e + (e_cv - u) * p_c) / (e_cv - e_iv)


def tillotson_c2(eos10, rho: float, u: float) -> float:
    """Adiabatic sound speed squared by central differences (eos.rs:88-96)."""
    rho = max(rho, 1.0e-9)
    dr = rho * 1.0e-4
    du = abs(u) * 1.0e-4 + 1.0
    dp_drho = (tillotson_pressure(eos10, rho + dr, u) - tillotson_pressure(eos10, rho - dr, u)) / (2.0 * dr)
    dp_du = (tillotson_pressure(eos10, rho, u + du) - tillotson_pressure(eos10, rho, u - du)) / (2.0 * du)
    p = tillotson_pressure(eos10, rho, u)
    return max(dp_drho + p / (rho * rho) * dp_du, 0.0)
