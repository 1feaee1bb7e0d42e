"""Seeded synthetic inputs of the benchmark workloads (the simulator side of
the hot path: Simulation/synthesizeEncoding.m, synthesizeT.m, synthesizeY.m,
SNRToLevels.m, Methods/plugNoise.m).  Volumes are generated on the device
with torch; integer sampling comes from ``disorder_amd.sampling``.
All random parameters come from numpy ``Generator(PCG64(seed))``."""
from __future__ import annotations

import math

import numpy as np
import torch

from .ops import colmajor
from . import sampling


def _grid(n, device):
    ax = [torch.linspace(-1, 1, v, device=device, dtype=torch.float32) for v in n]
    return torch.meshgrid(*ax, indexing="ij")


def _rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def phantom(n, n_ell, seed, device, mag=(0.1, 1.0), phase_max=math.pi / 4):
    """Sum of n_ell random nested ellipsoids (magnitudes U(mag)) times a smooth
    quadratic phase bounded by phase_max; complex64, column-major."""
    rng = np.random.default_rng(seed)
    X = _grid(n, device)
    x = torch.zeros(n, dtype=torch.float32, device=device)
    for k in range(n_ell):
        scale = 0.85 * (1 - k / (n_ell + 1))  # nested: shrinking envelopes
        ax = rng.uniform(0.3, 1.0, 3) * scale
        c = rng.uniform(-0.15, 0.15, 3) * scale
        R = torch.tensor(_rotation(rng), dtype=torch.float32, device=device)
        d = [X[i] - float(c[i]) for i in range(3)]
        u = [R[0, i] * d[0] + R[1, i] * d[1] + R[2, i] * d[2] for i in range(3)]
        inside = (u[0] / ax[0]) ** 2 + (u[1] / ax[1]) ** 2 + (u[2] / ax[2]) ** 2 <= 1
        x = x + float(rng.uniform(*mag)) * inside.float()
    a = rng.uniform(-1, 1, 3)
    B = rng.uniform(-1, 1, (3, 3))
    ph = sum(float(a[i]) * X[i] for i in range(3)) + sum(float(B[i, j]) * X[i] * X[j] for i in range(3)
                                                          for j in range(3))
    ph = ph * (phase_max / max(ph.abs().max().item(), 1e-12))
    return colmajor(torch.polar(x, ph))


def coil_maps(n, C, seed, device, sigma=None, radius=None):
    """C Gaussian-blob sensitivities (sigma, radius in voxels; defaults scale
    24 and 48 voxels of a 64^3 grid with the grid) centred on a sphere around
    the FOV centre, each with a small linear phase, normalised to unit
    sum-of-squares per voxel; (n1,n2,n3,C) complex64 column-major."""
    rng = np.random.default_rng(seed)
    N = max(n)
    sigma = sigma if sigma is not None else 24.0 * N / 64
    radius = radius if radius is not None else 48.0 * N / 64
    idx = [torch.arange(v, device=device, dtype=torch.float32) - (v - 1) / 2 for v in n]
    X = torch.meshgrid(*idx, indexing="ij")
    maps = []
    gold = math.pi * (3 - math.sqrt(5))
    for c in range(C):
        zc = 1 - 2 * (c + 0.5) / C
        rc = math.sqrt(max(0.0, 1 - zc * zc))
        th = gold * c
        ctr = radius * np.array([rc * math.cos(th), rc * math.sin(th), zc])
        r2 = sum((X[i] - float(ctr[i])) ** 2 for i in range(3))
        mag = torch.exp(-r2 / (2 * sigma * sigma))
        kv = rng.uniform(-1, 1, 3) * math.pi / N
        ph = sum(float(kv[i]) * X[i] for i in range(3))
        maps.append(torch.polar(mag, ph))
    S = torch.stack(maps, dim=-1)
    S = S / torch.sqrt((S.abs() ** 2).sum(dim=-1, keepdim=True))
    return colmajor(S.to(torch.complex64))


def random_walk_motion(S, seed, step_deg=0.5, step_vox=0.5, clip_deg=8.0, clip_vox=8.0):
    """Per-shot random walk: steps N(0, step) in every parameter, clipped;
    radians for rotations.  [S x 6] float64."""
    rng = np.random.default_rng(seed)
    T = np.zeros((S, 6))
    for s in range(1, S):
        T[s, :3] = np.clip(T[s - 1, :3] + rng.normal(0, step_vox, 3), -clip_vox, clip_vox)
        T[s, 3:] = np.clip(T[s - 1, 3:] + rng.normal(0, step_deg, 3) * math.pi / 180,
                           -clip_deg * math.pi / 180, clip_deg * math.pi / 180)
    return T


def uniform_motion(S, seed, rot_deg=5.0, tra=3.0, remove_mean=True):
    """Simulation/synthesizeT.m:36-40 for one motion state per shot: rotations
    pi theta_r (U - 0.5) / 180 and translations theta_t (U - 0.5) with
    theta_r = 2 rot_deg, theta_t = 2 tra (so U(-rot, rot) degrees and
    U(-tra, tra) voxels), then the mean over shots subtracted from every
    parameter (:38, as the reference does for the identifiability of the
    common motion).  [S x 6] float64, radians for rotations."""
    rng = np.random.default_rng(seed)
    T = np.zeros((S, 6))
    T[:, :3] = 2 * tra * (rng.uniform(0, 1, (S, 3)) - 0.5)
    T[:, 3:] = math.pi * 2 * rot_deg * (rng.uniform(0, 1, (S, 3)) - 0.5) / 180
    if remove_mean:
        T -= T.mean(axis=0, keepdims=True)
    return T


def disorder_sampling(n1, n2, S, t1, t2, seed):
    """DISORDER random-checkered sampling and its per-state sample lists."""
    rng = np.random.default_rng(seed)
    perms = sampling.permutation_table(rng, (n1 // t1) * (n2 // t2), S)
    H = sampling.random_checkered(n1, n2, t1, t2, perms)
    k1, k2, off = sampling.harmonic_sampling(H, S)
    return H, k1, k2, off


def noise_level(x, Smaps, snr_db, unitary=True):
    """Simulation/SNRToLevels.m:23-33: sigma such that the SENSE-combined image
    noise power 2 sigma^2 mean(1/(sum|S|^2 + 1e-3)) / K equals the mean signal
    power over 10^(dB/10).  unitary=False is the reference's formula,
    K = numel(x), for data y = F S x with the simulator's unnormalised DFT.
    unitary=True gives the same image SNR for data in the reconstruction
    path's unitary convention (F / sqrt(numel), solveXT.m:342): those data
    and their noise are the simulator's divided by sqrt(numel(x)), so the
    level is the reference's divided by sqrt(numel(x)), i.e. K = 1."""
    meanS = (x.abs() ** 2).mean().item()
    meanN = meanS / (10 ** (snr_db / 10))
    P = (1.0 / ((Smaps.abs() ** 2).sum(dim=-1) + 1e-3)).mean().item() / (1 if unitary else x.numel())
    return math.sqrt(meanN / (2 * P))


def add_noise(y, sigma, seed):
    """Methods/plugNoise.m:17: complex N(0,1) scaled by sigma."""
    g = torch.Generator(device=y.device).manual_seed(seed)
    nz = torch.randn(tuple(y.shape), dtype=torch.complex64, device=y.device, generator=g)
    # torch's complex randn has variance 1 split over re/im; plugNoise draws
    # N(0,1) for each part
    return colmajor(y + sigma * math.sqrt(2.0) * nz)


def truncate_shots(H, frac, seed, keep=(0.3, 0.8)):
    """Uneven shots: a fraction `frac` of the shots, chosen at random, lose the
    tail of their samples (they keep a U(keep) share, in column-major order);
    the dropped positions become unsampled (H = 0)."""
    rng = np.random.default_rng(seed)
    H = np.array(H, copy=True)
    S = int(H.max())
    nt = max(1, int(round(frac * S)))
    for s in rng.choice(np.arange(1, S + 1), nt, replace=False):
        idx = np.nonzero(H.reshape(-1, order="F") == s)[0]
        k = int(round(rng.uniform(*keep) * len(idx)))
        drop = idx[k:]
        Hf = H.reshape(-1, order="F")
        Hf[drop] = 0
        H = Hf.reshape(H.shape, order="F")
    return H


def batched_volumes():
    """The batched form's six volumes [PE1, PE2, RO] and shot tilings (t1 x t2
    rectangular DISORDER tiles, 40-96 shots)."""
    return [((160, 120, 160), (5, 8)), ((192, 144, 192), (8, 9)), ((192, 144, 192), (8, 12)),
            ((224, 160, 224), (7, 8)), ((256, 192, 256), (8, 8)), ((256, 192, 256), (8, 6))]


def batched_problem(dev, C=8, seed=0, volumes=None, trunc=0.1, snr_db=25.0):
    """Six independent seeded reconstructions of uneven sizes: phantom, coils,
    rectangular-tile DISORDER sampling with 10 % of shots truncated,
    random-walk motion, noisy data.  Returns a list of dicts."""
    from . import ops
    out = []
    for i, (n, (t1, t2)) in enumerate(volumes or batched_volumes()):
        sd = seed + 100 * i
        S = t1 * t2
        rng = np.random.default_rng(sd)
        perms = sampling.permutation_table(rng, (n[0] // t1) * (n[1] // t2), S)
        H = sampling.random_checkered(n[0], n[1], t1, t2, perms)
        H = truncate_shots(H, trunc, sd + 1)
        k1, k2, off = sampling.harmonic_sampling(H, S)
        x = phantom(n, 40, seed=sd + 2, device=dev)
        Sm = coil_maps(n, C, seed=sd + 3, device=dev)
        T = random_walk_motion(S, seed=sd + 4)
        E = ops.Encoding(Sm, k1, k2, off)
        y = E.encode(x, T)
        y = add_noise(y, noise_level(x, Sm, snr_db), seed=sd + 5)
        out.append({"n": n, "S": S, "tiles": (t1, t2), "E": E, "x_gt": x, "y": y, "T": T, "Smaps": Sm,
                    "Se": ops.coil_precond(Sm), "lines": np.diff(off)})
    return out
