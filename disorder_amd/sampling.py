"""Integer sampling bookkeeping (host side, bit-exact).

* ``random_checkered`` -- the 'Random-checkered' DISORDER mask of
  Simulation/synthesizeEncoding.m:86-107 for one repeat and no multiband:
  the (n1, n2) phase-encode plane is cut into tiles of t1 x t2 positions and
  inside every tile a permutation assigns the t1*t2 positions to the S = t1*t2
  shots; ``ifftshift`` then puts DC first (:127).  The permutation table is an
  explicit input (the reference draws it with randperm and no seed), so the
  mask is a pure function of it.
* ``harmonic_sampling`` -- Methods/buildHarmonicSampling.m:36-49: per state
  the sampled positions in MATLAB column-major linear order, states
  concatenated, unsampled positions last as an extra (outer) segment.
"""
from __future__ import annotations

import numpy as np


def permutation_table(rng: np.random.Generator, ntiles: int, S: int) -> np.ndarray:
    """[ntiles x S] table; row t is the randperm(S) of tile t (0-based)."""
    return np.stack([rng.permutation(S) for _ in range(ntiles)]).astype(np.int64)


def random_checkered(n1: int, n2: int, t1: int, t2: int, perms: np.ndarray) -> np.ndarray:
    """H[k1, k2] = shot (1..S) that samples phase-encode position (k1, k2),
    DC first.  Tiles are visited with the dim-2 tile index fastest, as the
    (o, p) loops of synthesizeEncoding.m:90-91; inside a tile, shot s takes
    position ind2sub([t1 t2], perm(s)) (:92-99)."""
    S = t1 * t2
    if n1 % t1 or n2 % t2:
        raise ValueError("tile sizes must divide the phase-encode grid")
    nb1, nb2 = n1 // t1, n2 // t2
    perms = np.asarray(perms, dtype=np.int64)
    if perms.shape != (nb1 * nb2, S):
        raise ValueError("permutation table must be [ntiles x S]")
    H = np.zeros((n1, n2), np.int64)
    t = 0
    for o in range(nb1):
        for p in range(nb2):
            pr = perms[t]
            i = pr % t1   # ind2sub, column-major
            j = pr // t1
            H[o * t1 + i, p * t2 + j] = np.arange(1, S + 1)
            t += 1
    return np.fft.ifftshift(H)


def harmonic_sampling(H: np.ndarray, nstates: int, with_outer: bool = False):
    """k1, k2 (0-based FFT-order indices), state offsets [nstates+1] (or
    [nstates+2] with the outer segment of unsampled positions)."""
    n1, _ = H.shape
    Hl = np.asarray(H).reshape(-1, order="F")
    order = np.argsort(np.where(Hl == 0, nstates + 1, Hl), kind="stable")
    counts = np.bincount(Hl, minlength=nstates + 1)
    ngroups = nstates + (1 if with_outer else 0)
    if not with_outer:
        order = order[: len(order) - counts[0]]
    off = np.zeros(ngroups + 1, np.int64)
    off[1:nstates + 1] = np.cumsum(counts[1:nstates + 1])
    if with_outer:
        off[nstates + 1] = off[nstates] + counts[0]
    k1 = (order % n1).astype(np.int32)
    k2 = (order // n1).astype(np.int32)
    return k1, k2, off.astype(np.int32)


def harmonic_sampling_repeats(H: np.ndarray, nstates: int):
    """buildHarmonicSampling over a state map with repeats, H [n1, n2] or
    [n1, n2, R] (positions taken in MATLAB linear order over all three
    dimensions): (k1, k2, repeat, off), repeat the 0-based repeat of each
    sample."""
    H = np.asarray(H)
    n1, n2 = H.shape[:2]
    Hl = H.reshape(-1, order="F")
    order = np.argsort(np.where(Hl == 0, nstates + 1, Hl), kind="stable")
    counts = np.bincount(Hl, minlength=nstates + 1)
    order = order[: len(order) - counts[0]]
    off = np.zeros(nstates + 1, np.int64)
    off[1:] = np.cumsum(counts[1:nstates + 1])
    k1 = (order % n1).astype(np.int32)
    k2 = ((order // n1) % n2).astype(np.int32)
    rep = (order // (n1 * n2)).astype(np.int32)
    return k1, k2, rep, off.astype(np.int32)


def signed_freq(k, n):
    """Signed frequency of FFT-order index k of a length-n transform (the
    Nyquist bin of an even length is negative)."""
    k = np.asarray(k)
    return np.where(k < (n + 1) // 2, k, k - n)


def level_sampling(k1, k2, off, N12, n12):
    """Samples of a pyramid level (Methods/downsampleOperators.m:36, the
    sampling resampled in k-space with resampling(A, NYRes(1:2), 1), then
    Methods/buildHarmonicSampling.m:36-49 on the cropped plane).  A sample
    survives when both signed frequencies lie in [-ceil((n-1)/2),
    floor((n-1)/2)] of the level size; its level index is the frequency mod n.
    Column-major order within a state is preserved by the crop, so the kept
    samples stay in their full-resolution order.  Returns (rows, k1, k2, off)
    with rows indexing the full-resolution samples."""
    k1 = np.asarray(k1)
    k2 = np.asarray(k2)
    off = np.asarray(off)
    f1 = signed_freq(k1, N12[0])
    f2 = signed_freq(k2, N12[1])
    keep = np.ones(len(k1), bool)
    for f, n in ((f1, n12[0]), (f2, n12[1])):
        keep &= (f >= -(n // 2)) & (f <= (n - 1) // 2)
    rows = np.nonzero(keep)[0]
    state = np.searchsorted(off, np.arange(len(k1)), side="right") - 1
    offl = np.zeros(len(off), np.int64)
    offl[1:] = np.cumsum(np.bincount(state[rows], minlength=len(off) - 1))
    return (rows.astype(np.int32), np.mod(f1[rows], n12[0]).astype(np.int32),
            np.mod(f2[rows], n12[1]).astype(np.int32), offl.astype(np.int32))
