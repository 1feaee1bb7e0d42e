"""Inputs of the reference's CPU-runnable case (64^3, 8 coils, 16 shots),
generated with numpy in FP64 so that the fixture script and the GPU test
build bit-identical inputs on any host:

* x_gt: 10 random nested ellipsoids, magnitudes U(0.2, 1), times a smooth
  quadratic phase scaled into [-pi/4, pi/4];
* S: 8 Gaussian-blob sensitivities (sigma 24 voxels) centred on a sphere of
  radius 48 voxels about the FOV centre, each with a linear phase,
  normalised to unit sum of squares;
* sampling: 64 x 64 phase-encode plane in 16 shots, DISORDER random-checkered
  4 x 4 tiles (disorder_amd.sampling, Simulation/synthesizeEncoding.m:86-107);
* T_gt: rotations U(-5, 5) degrees and translations U(-3, 3) voxels with the
  mean over shots removed (disorder_amd.synth.uniform_motion,
  Simulation/synthesizeT.m:36-40);
* y = E(T_gt) x_gt from the FP64 oracle (unitary phase-encode DFT, readout in
  image space), plus complex Gaussian noise at 30 dB by SNRToLevels.m:23-33 in
  that convention (disorder_amd.synth.noise_level).

Only the arithmetic operations +, -, *, / and comparisons decide the
ellipsoid membership, so the phantom's support is identical on every host."""
from __future__ import annotations

import math

import numpy as np

N = (64, 64, 64)
C = 8
S = 16
TILES = (4, 4)
SNR_DB = 30.0
SEED = 2024
# driver schedule of the run (the 64^3 case of the reference: 20 CG
# iterations with the known motion, then 3 pyramid levels of joint LM+CG)
CG_KNOWN = 20
PYR = dict(nlevels=3, n_outer=10, n_cg=8, final_cg=300)


def _rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def phantom(n, n_ell, seed, mag=(0.2, 1.0), phase_max=math.pi / 4):
    rng = np.random.default_rng(seed)
    X = np.meshgrid(*[np.linspace(-1.0, 1.0, v) for v in n], indexing="ij")
    x = np.zeros(n)
    for k in range(n_ell):
        scale = 0.85 * (1 - k / (n_ell + 1))
        ax = rng.uniform(0.3, 1.0, 3) * scale
        c = rng.uniform(-0.15, 0.15, 3) * scale
        R = _rotation(rng)
        d = [X[i] - c[i] for i in range(3)]
        u = [R[0, i] * d[0] + R[1, i] * d[1] + R[2, i] * d[2] for i in range(3)]
        inside = (u[0] / ax[0]) ** 2 + (u[1] / ax[1]) ** 2 + (u[2] / ax[2]) ** 2 <= 1
        x = x + rng.uniform(*mag) * inside
    a = rng.uniform(-1, 1, 3)
    B = rng.uniform(-1, 1, (3, 3))
    ph = sum(a[i] * X[i] for i in range(3)) + sum(B[i, j] * X[i] * X[j] for i in range(3) for j in range(3))
    ph = ph * (phase_max / max(np.abs(ph).max(), 1e-12))
    return x * np.exp(1j * ph)


def coil_maps(n, C, seed, sigma=24.0, radius=48.0):
    rng = np.random.default_rng(seed)
    X = np.meshgrid(*[np.arange(v) - (v - 1) / 2 for v in n], indexing="ij")
    Nmax = max(n)
    gold = math.pi * (3 - math.sqrt(5))
    maps = []
    for c in range(C):
        zc = 1 - 2 * (c + 0.5) / C
        rc = math.sqrt(max(0.0, 1 - zc * zc))
        th = gold * c
        ctr = radius * np.array([rc * math.cos(th), rc * math.sin(th), zc])
        r2 = sum((X[i] - ctr[i]) ** 2 for i in range(3))
        kv = rng.uniform(-1, 1, 3) * math.pi / Nmax
        maps.append(np.exp(-r2 / (2 * sigma * sigma)) * np.exp(1j * sum(kv[i] * X[i] for i in range(3))))
    Sm = np.stack(maps, axis=-1)
    return Sm / np.sqrt(np.sum(np.abs(Sm) ** 2, axis=-1, keepdims=True))


def make_inputs():
    """dict with x_gt (complex128), Sm (complex64-representable complex128),
    k1, k2, off, T_gt, y (complex64-representable complex128), sigma."""
    import torch
    from disorder_amd import sampling, synth
    from oracle import disorder_oracle as O
    rng = np.random.default_rng(SEED)
    x_gt = phantom(N, 10, SEED + 1)
    Sm = coil_maps(N, C, SEED + 2).astype(np.complex64).astype(np.complex128)
    perms = sampling.permutation_table(rng, (N[0] // TILES[0]) * (N[1] // TILES[1]), S)
    H = sampling.random_checkered(N[0], N[1], TILES[0], TILES[1], perms)
    k1, k2, off = sampling.harmonic_sampling(H, S)
    T_gt = synth.uniform_motion(S, SEED + 3)
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    y = O.encode(x_gt, Sm, samp, T_gt, N)
    sigma = synth.noise_level(torch.from_numpy(x_gt), torch.from_numpy(Sm), SNR_DB, unitary=True)
    nrng = np.random.default_rng(SEED + 4)
    y = y + sigma * (nrng.standard_normal(y.shape) + 1j * nrng.standard_normal(y.shape))  # plugNoise.m:17
    y = y.astype(np.complex64).astype(np.complex128)
    return {"x_gt": x_gt, "Sm": Sm, "k1": k1, "k2": k2, "off": off, "T_gt": T_gt, "y": y, "sigma": sigma}


def aligned_nrmse(x, T, x_gt, T_gt, transform):
    """NRMSE against the ground truth after removing the common motion: the
    estimate is determined only up to a rigid motion shared by all shots,
    so it is moved by the mean of T - T_gt (the mean-motion removal of
    Simulation/pyramidDISORDER.m:124-131) before the comparison;
    transform(x, Tmean) applies sincRigidTransform."""
    d = np.mean(np.asarray(T) - np.asarray(T_gt), axis=0, keepdims=True)
    xa = transform(x, d)
    return float(np.linalg.norm(xa - x_gt) / np.linalg.norm(x_gt))
