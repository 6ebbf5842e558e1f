"""Writes the small golden fixtures under tests/golden/ from the oracle (run once; the outputs are committed).

    python tests/golden/make_fixtures.py

lbvh_cloud1500.npz: the reference verifier's N=1500 splitmix cloud (gpu-bh-verify:116-137, seed 0xABCD1234)
through the oracle's Morton leg and Karras build at K=1 and K=16 (u32 arrays). The published accuracy
numbers the oracle is checked against live in reference_published.json (transcribed data).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle  # noqa: E402
from tests.helpers import bh_cloud  # noqa: E402


def main():
    pos, mass = bh_cloud(1500)
    out = {}
    for k in (1, 16):
        lb = oracle.Lbvh(pos.astype(np.float32), mass.astype(np.float32), k)
        out["sorted_codes"] = lb.sorted_codes
        out["order"] = lb.order
        out[f"left_k{k}"] = lb.left
        out[f"right_k{k}"] = lb.right
        out[f"parent_k{k}"] = lb.parent
    np.savez_compressed(os.path.join(HERE, "lbvh_cloud1500.npz"), **out)


if __name__ == "__main__":
    main()
