"""Motion-state binning for the image solve (Methods/compressMotion.m:27-80).

Each of the six motion parameters, as a series over the S shots, is taken to
NL = floor(log2 S) levels of the orthonormal Haar transform with MATLAB's
default 'sym' boundary (an odd-length series repeats its last value), every
coefficient (approximation included) below the binning limit is zeroed —
traLimX(1) * res voxels for translations, traLimX(2) * res degrees for
rotations, res the isotropic resolution of the finest level — and the series
is rebuilt.  Runs of consecutive shots whose rebuilt parameters agree with
the run's first shot to 1e-6 share one bin, whose motion is their mean
rebuilt motion and whose samples are theirs, concatenated in shot order.

The CG image solve encodes with the binned motion (solveXT.m:392: E.Tr, E.Fs,
E.nEc); LMsolver keeps the per-shot motion (:496).  Shots sit contiguously
in sample order, so a bin's samples are the union of its shots' ranges and
the data need no reordering: only the state offsets change.
"""
from __future__ import annotations

import math

import numpy as np

TRA_LIM_X = (0.05, 0.02)


def _haar_analysis(x):
    """One level over axis 0 of [n, P]: pairs (x_{2k}, x_{2k+1}) with the
    last row repeated for odd n; returns (approximation, detail)."""
    if x.shape[0] % 2:
        x = np.concatenate([x, x[-1:]], axis=0)
    e, o = x[0::2], x[1::2]
    return (e + o) * math.sqrt(0.5), (e - o) * math.sqrt(0.5)


def _haar_synthesis(a, d, n):
    out = np.empty((2 * a.shape[0],) + a.shape[1:])
    out[0::2] = (a + d) * math.sqrt(0.5)
    out[1::2] = (a - d) * math.sqrt(0.5)
    return out[:n]


def haar_denoise(T, thresholds, levels):
    """Hard-threshold every Haar coefficient of each column of T [S, P]
    (column p against thresholds[p]) over `levels` levels; rebuilt [S, P]."""
    a = np.asarray(T, np.float64)
    th = np.asarray(thresholds, np.float64)[None, :]
    lengths, details = [], []
    for _ in range(levels):
        lengths.append(a.shape[0])
        a, d = _haar_analysis(a)
        details.append(np.where(np.abs(d) < th, 0.0, d))
    a = np.where(np.abs(a) < th, 0.0, a)
    for n, d in zip(reversed(lengths), reversed(details)):
        a = _haar_synthesis(a, d, n)
    return a


def bin_motion(T, off, res, traLimX=TRA_LIM_X):
    """(T_bin [B x 6], off_bin [B + 1], bin_of_shot [S]) for shot motion
    T [S x 6] with sample offsets off [S + 1]."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    off = np.asarray(off, np.int64)
    S = T.shape[0]
    if off.shape != (S + 1,):
        raise ValueError("off must have one entry more than T has rows")
    if S <= 1:
        return T.copy(), off.copy(), np.zeros(S, np.int64)
    th = [traLimX[0] * res] * 3 + [traLimX[1] * res * math.pi / 180.0] * 3
    R = haar_denoise(T, th, int(math.floor(math.log2(S))))
    bin_of_shot = np.empty(S, np.int64)
    first = 0
    b = 0
    for s in range(S):
        if s > first and (np.abs(R[s] - R[first]) > 1e-6).any():
            first = s
            b += 1
        bin_of_shot[s] = b
    B = b + 1
    counts = np.bincount(bin_of_shot, minlength=B)
    T_bin = np.add.reduceat(R, np.concatenate([[0], np.cumsum(counts)[:-1]]), axis=0) / counts[:, None]
    starts = np.flatnonzero(np.diff(np.concatenate([[-1], bin_of_shot])))
    off_bin = np.concatenate([off[starts], off[-1:]])
    return T_bin, off_bin, bin_of_shot
