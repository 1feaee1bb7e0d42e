"""Shared helpers for the parity tests (seeded inputs, error metrics)."""
import numpy as np


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def smooth_volume(rng, n, nblob=6):
    """Band-limited-ish test volume: sum of Gaussian blobs with random phase."""
    g = np.meshgrid(*[np.arange(v) - v / 2 for v in n], indexing="ij")
    x = np.zeros(n, np.complex128)
    for _ in range(nblob):
        c = [rng.uniform(-v / 4, v / 4) for v in n]
        s = rng.uniform(0.08, 0.2) * min(n)
        r2 = sum((gi - ci) ** 2 for gi, ci in zip(g, c))
        x += rng.uniform(0.3, 1.0) * np.exp(-r2 / (2 * s * s)) * np.exp(1j * rng.uniform(-np.pi, np.pi))
    return x


def random_motion(rng, S, rot_deg=5.0, tra=2.0):
    T = np.zeros((S, 6))
    T[:, :3] = rng.uniform(-tra, tra, (S, 3))
    T[:, 3:] = rng.uniform(-rot_deg, rot_deg, (S, 3)) * np.pi / 180
    return T


def checkered_sampling(rng, n1, n2, S, t1, t2):
    """Random-checkered DISORDER sampling on an (n1, n2) grid with t1 x t2 tiles
    (t1*t2 == S): per tile a random permutation assigns tile positions to
    states (Simulation/synthesizeEncoding.m:86-107), then ifftshift puts DC
    first.  Returns H (n1, n2) with states 1..S."""
    assert t1 * t2 == S and n1 % t1 == 0 and n2 % t2 == 0
    H = np.zeros((n1, n2), np.int64)
    for o in range(n1 // t1):
        for p in range(n2 // t2):
            perm = rng.permutation(S)
            for s in range(S):
                i, j = np.unravel_index(perm[s], (t1, t2), order="F")
                H[o * t1 + i, p * t2 + j] = s + 1
    return np.fft.ifftshift(H)
