"""Learned smoothing-length seed: host side (weights file, training driver)."""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "hseed_mlp.npz")


def default_weights_path():
    """The shipped trained weights, or None when none have been trained yet."""
    return WEIGHTS if os.path.exists(WEIGHTS) else None
