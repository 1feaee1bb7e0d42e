"""Synthetic-input generators (disorder_amd/synth.py) against what the
reference's simulator defines them to be (CPU, torch tensors on the host):

* noise_level: Simulation/SNRToLevels.m:23-33 defines the SNR as the mean
  image power over the mean power of the SENSE-combined noise image; the
  level is checked by drawing plugNoise-style noise at that level, taking it
  to the image domain the way each convention's decoder does, combining the
  coils and measuring the ratio (not by restating the formula);
* add_noise: Methods/plugNoise.m:17 draws N(0, 1) real and imaginary parts,
  scaled by the level;
* uniform_motion: Simulation/synthesizeT.m:36-40 (uniform draws, the mean
  over shots removed);
* random_walk_motion: the benchmark's random walk (steps N(0, step),
  clipped)."""
import math

import numpy as np
import pytest
import torch

from disorder_amd import synth


def _maps(rng, n, C):
    S = rng.standard_normal(n + (C,)) + 1j * rng.standard_normal(n + (C,))
    S = S / np.sqrt(np.sum(np.abs(S) ** 2, axis=-1, keepdims=True))  # unit sum of squares, as the generators
    return torch.from_numpy(S.astype(np.complex64))


@pytest.mark.parametrize("unitary", [False, True])
def test_noise_level_gives_the_prescribed_image_snr(unitary):
    rng = np.random.default_rng(0)
    n, C, db = (24, 20, 16), 4, 20.0
    x = torch.from_numpy((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64))
    S = _maps(rng, n, C)
    sigma = synth.noise_level(x, S, db, unitary=unitary)
    Sn = S.numpy().astype(np.complex128)
    powers = []
    for rep in range(6):
        nz = sigma * (rng.standard_normal(n + (C,)) + 1j * rng.standard_normal(n + (C,)))  # plugNoise.m:17
        if unitary:  # reconstruction path: unitary DFT along the phase encodes, readout in image space
            img = np.fft.ifft2(nz, axes=(0, 1), norm="ortho")
        else:        # simulator: unnormalised DFT, ifft (1/N) back (decodeDISORDER.m)
            img = np.fft.ifftn(nz, axes=(0, 1, 2))
        comb = np.sum(np.conj(Sn) * img, axis=-1) / (np.sum(np.abs(Sn) ** 2, axis=-1) + 1e-3)
        powers.append(np.mean(np.abs(comb) ** 2))
    snr = np.mean(np.abs(x.numpy()) ** 2) / np.mean(powers)
    assert abs(10 * math.log10(snr) - db) < 0.05, 10 * math.log10(snr)


def test_noise_level_conventions_differ_by_sqrt_numel():
    rng = np.random.default_rng(1)
    n = (10, 8, 6)
    x = torch.from_numpy((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64))
    S = _maps(rng, n, 3)
    a = synth.noise_level(x, S, 25.0, unitary=False)
    b = synth.noise_level(x, S, 25.0, unitary=True)
    assert b == pytest.approx(a / math.sqrt(np.prod(n)), rel=1e-12)


def test_add_noise_draws_unit_normal_parts_scaled_by_sigma():
    y = torch.zeros((200, 64, 8), dtype=torch.complex64)
    out = synth.add_noise(y, 0.3, seed=4).numpy()
    re, im = out.real.ravel(), out.imag.ravel()
    assert abs(re.std() - 0.3) < 0.003 and abs(im.std() - 0.3) < 0.003
    assert abs(re.mean()) < 0.003 and abs(im.mean()) < 0.003
    assert abs(np.corrcoef(re, im)[0, 1]) < 0.01


def test_uniform_motion_follows_synthesizeT():
    S = 20000
    T = synth.uniform_motion(S, seed=3, rot_deg=5.0, tra=3.0)
    np.testing.assert_allclose(T.mean(axis=0), 0.0, atol=1e-12)  # synthesizeT.m:38
    raw = synth.uniform_motion(S, seed=3, rot_deg=5.0, tra=3.0, remove_mean=False)
    np.testing.assert_allclose(T, raw - raw.mean(axis=0), atol=1e-15)
    assert np.abs(raw[:, :3]).max() <= 3.0 and np.abs(raw[:, 3:]).max() <= 5.0 * math.pi / 180
    # U(-a, a): variance a^2 / 3, both halves of the range populated
    np.testing.assert_allclose(raw[:, :3].var(axis=0), 3.0, rtol=0.04)
    np.testing.assert_allclose(raw[:, 3:].var(axis=0), (5 * math.pi / 180) ** 2 / 3, rtol=0.04)
    assert (raw[:, 0] > 2.5).any() and (raw[:, 0] < -2.5).any()


def test_random_walk_steps_and_clip():
    T = synth.random_walk_motion(5000, seed=2, step_deg=0.5, step_vox=0.5, clip_deg=8.0, clip_vox=8.0)
    assert np.all(T[0] == 0)
    assert np.abs(T[:, :3]).max() <= 8.0 and np.abs(T[:, 3:]).max() <= 8.0 * math.pi / 180 + 1e-15
    d = np.diff(T, axis=0)
    free = (np.abs(T[1:, :3]) < 8.0) & (np.abs(T[:-1, :3]) < 8.0)
    assert abs(d[:, :3][free].std() - 0.5) < 0.02
    freer = (np.abs(T[1:, 3:]) < 8 * math.pi / 180) & (np.abs(T[:-1, 3:]) < 8 * math.pi / 180)
    assert abs(d[:, 3:][freer].std() * 180 / math.pi - 0.5) < 0.02
