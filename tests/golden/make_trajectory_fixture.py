"""Writes trajectory_checkered.npz: the 'Checkered' (deterministic) DISORDER
order of an elliptical-shutter phase-encode grid, from the oracle's literal
restatement of Acquisition/samplingDISORDER.m.

  python -m tests.golden.make_trajectory_fixture
"""
import os

import numpy as np

from oracle import trajectory_oracle as TO


def shutter_points(K1, K2):
    """Grid points -K/2..K/2-1 inside the ellipse inscribed in the grid, in
    column-major order."""
    g1 = np.arange(K1) - K1 // 2
    g2 = np.arange(K2) - K2 // 2
    pts = [(a, b) for b in g2 for a in g1 if (a / (K1 / 2)) ** 2 + (b / (K2 / 2)) ** 2 <= 1.0]
    return np.array(pts, np.int64)


CASES = [((32, 32), (4, 4), "AlternatingZigZag"), ((24, 20), (4, 4), "ZigZag"), ((30, 26), (3, 2), "AlternatingZigZag")]


def main():
    out = {}
    for i, (K, U, so) in enumerate(CASES):
        k = shutter_points(*K)
        if so == "ZigZag":
            k = k[: len(k) - len(k) % (U[0] * U[1])]
        ko, NP, Ko, _, shot, echo = TO.sampling_disorder(k, U, "Checkered", so, K=K)
        out[f"k_in_{i}"] = k
        out[f"k_out_{i}"] = ko
        out[f"shot_{i}"] = shot
        out[f"echo_{i}"] = echo
    path = os.path.join(os.path.dirname(__file__), "trajectory_checkered.npz")
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
